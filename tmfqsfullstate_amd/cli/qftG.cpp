// qftG: makeQft(n) on the pattern input (uniform over states 8x + x mod 2)
// or on the random-phase input of seed 123456, through sim::Simulator, then
// one measurement and the period estimate N/k (reference examples/qftG.cpp).
#include "cli_options.h"

namespace {
constexpr unsigned kRandomPhaseSeed = 123456u;  // the reference CLI's seed
}

int main(int argc, char *argv[]) {
	using namespace tmfqs;
	using namespace tmfqs_cli;
	auto usage = [] { std::cout << "Usage: ./qftG <num_qubits> [--input-family pattern|random-phase] " << commonUsage() << "\n"; };
	try {
		if(argc < 2) {
			usage();
			return 1;
		}
		unsigned long long n64 = 0;
		if(!parseUnsigned(argv[1], n64) || n64 == 0) {
			std::cout << "Number of qubits must be an integer >= 1\n";
			return 1;
		}
		if(n64 > maxSupportedQubitsForU32States()) {
			std::cout << "Number of qubits is too large for this example.\n";
			return 1;
		}
		const unsigned n = static_cast<unsigned>(n64);
		bool randomPhase = false;
		Options o;
		for(int i = 2; i < argc; ++i) {
			const std::string arg = argv[i];
			if(arg == "--input-family" && i + 1 < argc) {
				const std::string f = argv[++i];
				if(f != "pattern" && f != "random-phase") throw std::invalid_argument("Unknown input family: " + f);
				randomPhase = f == "random-phase";
				continue;
			}
			if(!takeCommon(argc, argv, i, o)) {
				std::cout << "Unknown option: " << arg << "\n";
				usage();
				return 1;
			}
		}
		if(o.config.workloadHint == StorageWorkloadHint::Generic) o.config.workloadHint = StorageWorkloadHint::Qft;
		const state::QuantumState input =
		    randomPhase ? state::QuantumState::randomPhase(n, kRandomPhaseSeed) : state::QuantumState::sparsePattern(n);
		const sim::RunResult run = simulate(o, circuit::makeQft(n), input);
		Mt19937RandomSource rng = measurementSource(o);
		const StateIndex k = run.state.measure(rng);
		std::cout << "Input family: " << (randomPhase ? "random-phase" : "pattern") << "\n";
		if(randomPhase) std::cout << "Random seed: " << kRandomPhaseSeed << "\n";
		std::cout << "Measured state (k): " << k << "\n";
		std::cout << "Calculated r (N/k): ";
		if(k == 0)
			std::cout << "undefined (division by zero)\n";
		else
			std::cout << static_cast<double>(checkedStateCount(n)) / static_cast<double>(k) << "\n";
		if(o.verbose) std::cout << "Execution seconds: " << run.report.executionSeconds << "\n";
		return 0;
	} catch(const std::exception &ex) {
		std::cout << "qftG error: " << ex.what() << "\n";
		printAvailableStrategies(std::cout);
		return 2;
	}
}
