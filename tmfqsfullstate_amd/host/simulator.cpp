// Circuit runner (reference: src/tmfqs/sim/simulator.cpp:17-96).
#include "tmfqs/sim/simulator.h"

#include <chrono>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "tmfqs/register/quantum_register.h"
#include "tmfqs/storage/factory/storage_strategy_registry.h"

namespace tmfqs {
namespace sim {
namespace {

using Clock = std::chrono::steady_clock;

// reference :17-44
void executeOperation(QuantumRegister &reg, const circuit::Operation &operation) {
	std::visit(
	    [&](const auto &op) {
		    using Op = std::decay_t<decltype(op)>;
		    if constexpr(std::is_same<Op, circuit::Hadamard>::value) {
			    reg.applyHadamard(op.target);
		    } else if constexpr(std::is_same<Op, circuit::PauliX>::value) {
			    reg.applyPauliX(op.target);
		    } else if constexpr(std::is_same<Op, circuit::ControlledNot>::value) {
			    reg.applyControlledNot(op.control, op.target);
		    } else if constexpr(std::is_same<Op, circuit::ControlledPhaseShift>::value) {
			    reg.applyControlledPhaseShift(op.control, op.target, op.theta);
		    } else if constexpr(std::is_same<Op, circuit::Swap>::value) {
			    reg.applySwap(op.first, op.second);
		    } else if constexpr(std::is_same<Op, circuit::MatrixGate>::value) {
			    reg.applyGate(op.gate, op.targets);
		    } else if constexpr(std::is_same<Op, circuit::PhaseFlipBasisState>::value) {
			    reg.applyPhaseFlipBasisState(op.state);
		    } else if constexpr(std::is_same<Op, circuit::InversionAboutMean>::value) {
			    if(op.materialized)
				    reg.applyInversionAboutMeanMaterialized();
			    else
				    reg.applyInversionAboutMean();
		    }
	    },
	    operation);
}

}  // namespace

Simulator::Simulator(ExecutionConfig config) : config_(std::move(config)) {}

const ExecutionConfig &Simulator::config() const noexcept { return config_; }

// reference :66-93: one batch around the op loop, closed on exceptions; with
// tracing on, each operation is waited for so that its seconds cover the
// device work too.
experiment::RunReport Simulator::runOn(const circuit::Circuit &circuit, QuantumRegister &reg) const {
	experiment::RunReport report;
	report.qubits = circuit.qubitCount();
	report.operationCount = circuit.operationCount();
	report.strategy = reg.storageStrategy();
	const auto start = Clock::now();
	reg.beginOperationBatch();
	bool open = true;
	try {
		const auto &ops = circuit.operations();
		for(size_t i = 0; i < ops.size(); ++i) {
			const auto t0 = Clock::now();
			executeOperation(reg, ops[i]);
			if(config_.observability.traceOperations) {
				reg.backend()->synchronize();
				report.operations.push_back({i, circuit::operationName(ops[i]), std::chrono::duration<double>(Clock::now() - t0).count()});
			}
		}
		reg.endOperationBatch();
		open = false;
	} catch(...) {
		if(open) {
			try {
				reg.endOperationBatch();
			} catch(...) {
			}
		}
		throw;
	}
	report.executionSeconds = std::chrono::duration<double>(Clock::now() - start).count();
	return report;
}

// reference :54-96: materialisation (host -> backend) is outside the timed region.
RunResult Simulator::run(const circuit::Circuit &circuit, const state::QuantumState &initialState) const {
	if(circuit.qubitCount() != initialState.qubitCount()) throw std::invalid_argument("Circuit and initial state qubit counts differ");
	RegisterConfig cfg = config_.backend;
	QuantumRegister reg = initialState.materialize(cfg);
	experiment::RunReport report = runOn(circuit, reg);
	report.strategy = StorageStrategyRegistry::resolve(circuit.qubitCount(), cfg);
	return {state::QuantumState::fromRegister(std::move(reg)), std::move(report)};
}

}  // namespace sim
}  // namespace tmfqs
