// grover / groverLazy: makeGrover(n, marked, materialized) from basis(n)
// through sim::Simulator, then one measurement.
#pragma once

#include <sstream>

#include "cli_options.h"

namespace tmfqs_cli {

inline int groverMain(int argc, char *argv[], const char *program, const char *mode, bool materialized) {
	using namespace tmfqs;
	auto usage = [&] {
		std::cout << "Usage: ./" << program << " <num_qubits> <marked_state[,marked_state...]> " << commonUsage() << "\n";
	};
	try {
		if(argc < 3) {
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
		BasisStateList marked;
		bool parsed = true;
		std::stringstream list(argv[2]);
		for(std::string item; parsed && std::getline(list, item, ',');) {
			unsigned long long m = 0;
			parsed = parseUnsigned(item, m);
			if(parsed) marked.push_back(m);
		}
		if(!parsed || marked.empty()) {
			std::cout << "Marked states must be a comma-separated list of integers.\n";
			return 1;
		}
		Options o;
		for(int i = 3; i < argc; ++i)
			if(!takeCommon(argc, argv, i, o)) {
				std::cout << "Unknown option: " << argv[i] << "\n";
				usage();
				return 1;
			}
		if(o.config.workloadHint == StorageWorkloadHint::Generic) o.config.workloadHint = StorageWorkloadHint::Grover;
		circuit::GroverCircuitOptions go;
		go.markedStates = marked;
		go.materializedDiffusion = materialized;
		const circuit::Circuit c = circuit::makeGrover(n, std::move(go));
		sim::ExecutionConfig ec;
		ec.backend = o.config;
		ec.observability.traceOperations = o.verbose || !o.reportJson.empty();
		const sim::RunResult run = sim::Simulator(ec).run(c, state::QuantumState::basis(n));
		if(!o.reportJson.empty()) experiment::writeJsonReport(run.report, o.reportJson);
		Mt19937RandomSource rng = measurementSource(o);
		const StateIndex result = run.state.measure(rng);
		std::cout << "Grover mode: " << mode << "\n";
		std::cout << "Resolved strategy: " << StorageStrategyRegistry::toString(StorageStrategyRegistry::resolve(n, o.config)) << "\n";
		std::cout << "Grover search result: " << result << "\n";
		if(o.verbose) {
			double p = 0.0;
			for(StateIndex m : marked) {
				const Amplitude a = run.state.amplitude(m);
				p += a.real * a.real + a.imag * a.imag;
			}
			std::cout << "Circuit operations: " << run.report.operationCount << "\n";
			std::cout << "Execution seconds: " << run.report.executionSeconds << "\n";
			std::cout.precision(17);
			std::cout << "Marked-state probability: " << p << "\n";
		}
		return 0;
	} catch(const std::exception &ex) {
		std::cout << program << " error: " << ex.what() << "\n";
		printAvailableStrategies(std::cout);
		return 2;
	}
}

}  // namespace tmfqs_cli
