// experiment_compare: one campaign workload simulated twice and compared
// with experiment::compareStates (reference examples/experiment_compare.cpp).
// The reference run is the Hip backend in the exact mode (the dense
// backend's arithmetic bit for bit); the candidate is the configuration the
// options give: --precision complex64 ("lossy"), the default fused
// complex128 ("fused": within 1e-12) or --exact ("reference": bit-equal).
// Prints the reference program's CSV header and row.
#include <iomanip>

#include "cli_options.h"

namespace {

using namespace tmfqs;
constexpr unsigned kRandomPhaseSeed = 123456u;

const char *kWorkloads[] = {"qft_pattern",          "qft_high_entropy",      "grover_single_n8",       "grover_single_n2",
                            "grover_single_7n8",    "grover_multi",          "grover_lazy_single_n8", "grover_lazy_single_n2",
                            "grover_lazy_single_7n8", "grover_lazy_multi"};

bool known(const std::string &w) {
	for(const char *k : kWorkloads)
		if(w == k) return true;
	return false;
}

// The campaign's marked states: N/8, N/2, 7N/8 (scripts/run_thesis_experiments.py:216-229).
BasisStateList markedFor(const std::string &w, StateIndex N) {
	if(w.find("multi") != std::string::npos) return BasisStateList({N / 8, N / 2, (7 * N) / 8});
	if(w.find("7n8") != std::string::npos) return BasisStateList({(7 * N) / 8});
	if(w.find("n2") != std::string::npos) return BasisStateList({N / 2});
	return BasisStateList({N / 8});
}

state::QuantumState runWorkload(unsigned n, const std::string &w, const RegisterConfig &cfg) {
	sim::ExecutionConfig ec;
	ec.backend = cfg;
	if(w.rfind("grover", 0) == 0) {
		circuit::GroverCircuitOptions go;
		go.markedStates = markedFor(w, checkedStateCount(n));
		go.materializedDiffusion = w.rfind("grover_lazy", 0) != 0;
		return sim::Simulator(ec).run(circuit::makeGrover(n, std::move(go)), state::QuantumState::basis(n)).state;
	}
	const state::QuantumState input =
	    w == "qft_high_entropy" ? state::QuantumState::randomPhase(n, kRandomPhaseSeed) : state::QuantumState::sparsePattern(n);
	return sim::Simulator(ec).run(circuit::makeQft(n), input).state;
}

}  // namespace

int main(int argc, char *argv[]) {
	using namespace tmfqs_cli;
	try {
		unsigned n = 0;
		std::string workload = "qft_pattern";
		Options o;
		o.config.strategy = StorageStrategyKind::Hip;
		bool ok = true;
		for(int i = 1; i < argc && ok; ++i) {
			const std::string arg = argv[i];
			if(arg == "--qubits" && i + 1 < argc) {
				n = static_cast<unsigned>(toNumber(argv[++i], "--qubits"));
			} else if(arg == "--workload" && i + 1 < argc) {
				workload = argv[++i];
				if(!known(workload)) throw std::invalid_argument("Unknown workload: " + workload);
			} else {
				ok = takeCommon(argc, argv, i, o);
			}
		}
		if(!ok || n == 0) {
			std::cout << "Usage: ./experiment_compare --qubits N --workload NAME " << commonUsage() << "\nWorkloads:";
			for(const char *k : kWorkloads) std::cout << " " << k;
			std::cout << "\n";
			return 1;
		}
		if(n > maxSupportedQubitsForU32States()) throw std::invalid_argument("Number of qubits is too large");
		o.config.workloadHint = workload.rfind("grover", 0) == 0 ? StorageWorkloadHint::Grover : StorageWorkloadHint::Qft;
		RegisterConfig ref = o.config;
		ref.hip.precision = HipPrecision::Complex128;
		ref.hip.exactPhases = true;
		const bool c64 = o.config.hip.precision == HipPrecision::Complex64;
		const bool exact = !c64 && o.config.hip.exactPhases;
		const experiment::StateComparison m =
		    experiment::compareStates(runWorkload(n, workload, ref), runWorkload(n, workload, o.config));
		std::cout << std::setprecision(17)
		          << "qubits,workload,strategy,accuracy_kind,bitwise_equal,max_abs_amplitude_error,"
		          << "max_abs_component_error,rel_l2,rmse_amplitude,max_abs_probability_error,"
		          << "total_probability_diff,worst_state\n"
		          << n << ',' << workload << ',' << (c64 ? "hip-complex64" : "hip") << ','
		          << (c64 ? "lossy" : exact ? "reference" : "fused") << ',' << (m.bitwiseEqual ? "true" : "false") << ','
		          << m.maxAbsAmplitudeError << ',' << m.maxAbsComponentError << ',' << m.relL2 << ',' << m.rmseAmplitude << ','
		          << m.maxAbsProbabilityError << ',' << m.totalProbabilityDiff << ',' << m.worstState << '\n';
		return 0;
	} catch(const std::exception &ex) {
		std::cerr << "experiment_compare error: " << ex.what() << '\n';
		return 2;
	}
}
