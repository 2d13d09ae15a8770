// Circuit runner (reference: src/tmfqs/sim/simulator.cpp:17-96).
#include "tmfqs/sim/simulator.h"

#include <chrono>
#include <vector>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "tmfqs/register/quantum_register.h"
#include "tmfqs/storage/factory/storage_strategy_registry.h"
#include "tmfqs/storage/hip/hip_state_backend.h"
#include "tmfqs_hip.h"

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

// reference :66-93: one batch around the op loop, closed on exceptions.  With
// tracing on, each operation's queued work is executed right after it,
// bracketed by stream events on a device backend (its seconds are the
// device time of its kernels, read once at the end) or waited for on others.
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
		tmfqs_hip_state *device = config_.observability.traceOperations ? hipBackendHandle(reg.backend()) : nullptr;
		for(size_t i = 0; i < ops.size(); ++i) {
			const auto t0 = Clock::now();
			executeOperation(reg, ops[i]);
			if(!config_.observability.traceOperations) continue;
			if(device) {
				raiseHipStatus(tmfqs_hip_flush_timed(device));
				report.operations.push_back({i, circuit::operationName(ops[i]), 0.0});
			} else {
				reg.backend()->synchronize();
				report.operations.push_back({i, circuit::operationName(ops[i]), std::chrono::duration<double>(Clock::now() - t0).count()});
			}
		}
		reg.endOperationBatch();
		open = false;
		if(device) {
			std::vector<double> ms(report.operations.size());
			size_t got = 0;
			raiseHipStatus(tmfqs_hip_timed_read(device, ms.data(), ms.size(), &got));
			for(size_t i = 0; i < report.operations.size() && i < got; ++i) report.operations[i].seconds = ms[i] * 1e-3;
		}
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
