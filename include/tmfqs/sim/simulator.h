// Circuit runner (reference: include/tmfqs/sim/simulator.h:16-41).
#ifndef TMFQS_SIM_SIMULATOR_H
#define TMFQS_SIM_SIMULATOR_H

#include <cstddef>
#include <string>
#include <vector>

#include "tmfqs/circuit/circuit.h"
#include "tmfqs/config/register_config.h"
#include "tmfqs/experiment/report.h"
#include "tmfqs/state/quantum_state.h"

namespace tmfqs {

class QuantumRegister;

namespace sim {

struct ObservabilityConfig {
	bool collectMetrics = false;
	/** Per-operation seconds; with a device backend each operation is waited for. */
	bool traceOperations = false;
};

struct ExecutionConfig {
	RegisterConfig backend;
	ObservabilityConfig observability;
};

struct RunResult {
	state::QuantumState state;
	experiment::RunReport report;
};

class Simulator {
	public:
		Simulator() = default;
		explicit Simulator(ExecutionConfig config);

		const ExecutionConfig &config() const noexcept;
		RunResult run(const circuit::Circuit &circuit, const state::QuantumState &initialState) const;
		/** Runs the circuit's operations on an existing register inside one batch
		 *  (the timed region of run(), simulator.cpp:66-93 in the reference). */
		experiment::RunReport runOn(const circuit::Circuit &circuit, QuantumRegister &reg) const;

	private:
		ExecutionConfig config_;
};

}  // namespace sim
}  // namespace tmfqs

#endif
