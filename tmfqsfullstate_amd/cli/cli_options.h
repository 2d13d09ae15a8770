// Command-line drivers of the device backend: the reference's grover,
// groverLazy, qftG and experiment_compare programs
// (examples/grover_cli.h:102-259, examples/qftG.cpp:114-291,
// examples/experiment_compare.cpp:228-268) with the Hip strategy.  Output
// lines keep the reference's wording, so campaign scripts that parse them
// (scripts/run_thesis_experiments.py in the reference) read these too.
//
// Options every driver takes:
//   --strategy hip|hip-chunked|auto|dense|blosc|zfp   (this library ships hip and
//                                         hip-chunked; auto resolves to hip)
//   --chunk-states N  --cache-slots N     (hip-chunked: amplitudes per host chunk and
//                                         chunks per device window, powers of two; the
//                                         reference's Blosc options of the same names)
//   --precision complex128|complex64      (device storage precision)
//   --device N  --tile-qubits N  --fuse-qubits K  --exact
//   --seed S                              (measurement seed; default: std::random_device)
//   --verbose  --report-json PATH         (RunReport with per-operation device times)
#pragma once

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmfqsfs.h"

namespace tmfqs_cli {

struct Options {
	tmfqs::RegisterConfig config;
	bool verbose = false;
	std::optional<uint32_t> seed;
	std::string reportJson;
};

// Decimal digits only (no sign, no spaces, no trailing text).
inline bool parseUnsigned(const std::string &text, unsigned long long &out) {
	if(text.empty() || text.find_first_not_of("0123456789") != std::string::npos) return false;
	errno = 0;
	out = std::strtoull(text.c_str(), nullptr, 10);
	return errno == 0;
}

inline unsigned long long toNumber(const std::string &text, const char *what) {
	unsigned long long v = 0;
	if(!parseUnsigned(text, v)) throw std::invalid_argument(std::string(what) + " must be a non-negative integer: " + text);
	return v;
}

inline tmfqs::StorageStrategyKind strategyNamed(const std::string &name) {
	for(tmfqs::StorageStrategyKind k : {tmfqs::StorageStrategyKind::Dense, tmfqs::StorageStrategyKind::Blosc, tmfqs::StorageStrategyKind::Zfp,
	                                    tmfqs::StorageStrategyKind::Auto, tmfqs::StorageStrategyKind::Hip,
	                                    tmfqs::StorageStrategyKind::HipChunked})
		if(tmfqs::StorageStrategyRegistry::toString(k) == name) return k;
	throw std::invalid_argument("Unknown strategy: " + name);
}

// Consumes argv[i] (and its value) when it is a common option; false when
// it is not one.
inline bool takeCommon(int argc, char *argv[], int &i, Options &o) {
	const std::string arg = argv[i];
	auto value = [&]() -> std::string {
		if(i + 1 >= argc) throw std::invalid_argument(arg + " needs a value");
		return argv[++i];
	};
	if(arg == "--strategy") {
		o.config.strategy = strategyNamed(value());
	} else if(arg == "--precision") {
		const std::string p = value();
		if(p == "complex128")
			o.config.hip.precision = tmfqs::HipPrecision::Complex128;
		else if(p == "complex64")
			o.config.hip.precision = tmfqs::HipPrecision::Complex64;
		else
			throw std::invalid_argument("Unknown precision: " + p);
	} else if(arg == "--chunk-states") {
		o.config.hip.chunkStates = static_cast<size_t>(toNumber(value(), "--chunk-states"));
	} else if(arg == "--cache-slots") {
		o.config.hip.windowChunks = static_cast<size_t>(toNumber(value(), "--cache-slots"));
	} else if(arg == "--device") {
		o.config.hip.device = static_cast<int>(toNumber(value(), "--device"));
	} else if(arg == "--tile-qubits") {
		o.config.hip.tileQubits = static_cast<unsigned>(toNumber(value(), "--tile-qubits"));
	} else if(arg == "--fuse-qubits") {
		o.config.hip.fuseQubits = static_cast<unsigned>(toNumber(value(), "--fuse-qubits"));
	} else if(arg == "--exact") {
		o.config.hip.exactPhases = true;
	} else if(arg == "--seed") {
		o.seed = static_cast<uint32_t>(toNumber(value(), "--seed"));
	} else if(arg == "--verbose") {
		o.verbose = true;
	} else if(arg == "--report-json") {
		o.reportJson = value();
	} else {
		return false;
	}
	return true;
}

inline tmfqs::Mt19937RandomSource measurementSource(const Options &o) {
	return o.seed ? tmfqs::Mt19937RandomSource(*o.seed) : tmfqs::Mt19937RandomSource();
}

inline void printAvailableStrategies(std::ostream &os) {
	os << "Available strategies: ";
	const std::vector<std::string> names = tmfqs::StorageStrategyRegistry::listAvailable();
	for(size_t i = 0; i < names.size(); ++i) os << names[i] << (i + 1 < names.size() ? ", " : "");
	os << "\n";
}

inline const char *commonUsage() {
	return "[--strategy hip|hip-chunked|auto|dense|blosc|zfp] [--chunk-states N] [--cache-slots N] "
	       "[--precision complex128|complex64] [--device N] [--tile-qubits N] "
	       "[--fuse-qubits K] [--exact] [--seed S] [--verbose] [--report-json PATH]";
}

// Simulator::run with the per-operation trace when a report is requested.
inline tmfqs::sim::RunResult simulate(const Options &o, const tmfqs::circuit::Circuit &c, const tmfqs::state::QuantumState &input) {
	tmfqs::sim::ExecutionConfig ec;
	ec.backend = o.config;
	ec.observability.traceOperations = !o.reportJson.empty();
	tmfqs::sim::RunResult r = tmfqs::sim::Simulator(ec).run(c, input);
	if(!o.reportJson.empty()) tmfqs::experiment::writeJsonReport(r.report, o.reportJson);
	return r;
}

}  // namespace tmfqs_cli
