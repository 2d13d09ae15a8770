// sim::adjointGradient / sim::expectation over tmfqs_hip_adjoint and
// tmfqs_hip_expectation (include/tmfqs_hip.h).
#include "tmfqs/sim/adjoint.h"

#include <chrono>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>

#include "tmfqs/register/quantum_register.h"
#include "tmfqs/storage/hip/hip_state_backend.h"
#include "tmfqs_hip.h"

namespace tmfqs {
namespace sim {
namespace {

void raise(int rc) {
	if(rc == TMFQS_HIP_OK) return;
	const std::string msg = tmfqs_hip_last_error();
	switch(rc) {
		case TMFQS_HIP_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
		case TMFQS_HIP_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
		case TMFQS_HIP_ERR_LOGIC: throw std::logic_error(msg);
		case TMFQS_HIP_ERR_OUT_OF_MEMORY: throw std::bad_alloc();
		default: throw std::runtime_error(msg);
	}
}

tmfqs_hip_state *deviceHandle(QuantumRegister &reg, const char *what) {
	tmfqs_hip_state *h = hipBackendHandle(reg.beginExternalUpdate());
	if(!h) throw std::runtime_error(std::string(what) + " requires a register on StorageStrategyKind::Hip");
	return h;
}

std::vector<tmfqs_hip_diag_term> toTerms(const DiagonalObservable &o) {
	std::vector<tmfqs_hip_diag_term> t;
	for(const auto &z : o.zStrings) {
		tmfqs_hip_diag_term d;
		std::memset(&d, 0, sizeof d);
		d.kind = TMFQS_HIP_DIAG_Z;
		d.coef = z.coef;
		for(QubitIndex q : z.qubits) {
			if(q >= 64) throw std::out_of_range("diagonal observable qubit index out of range");
			d.bits ^= 1ull << q;  // Z_q Z_q = I
		}
		t.push_back(d);
	}
	for(const auto &p : o.projectors) {
		tmfqs_hip_diag_term d;
		std::memset(&d, 0, sizeof d);
		d.kind = TMFQS_HIP_DIAG_PROJECTOR;
		d.coef = p.coef;
		d.bits = p.state;
		t.push_back(d);
	}
	return t;
}

struct OpList {
	std::vector<tmfqs_hip_op> ops;
	std::vector<double> mats;
	std::vector<unsigned> qubits;
	size_t params = 0;
};

OpList toOpList(const circuit::Circuit &c) {
	OpList L;
	for(const circuit::Operation &operation : c.operations()) {
		tmfqs_hip_op r;
		std::memset(&r, 0, sizeof r);
		std::visit(
		    [&](const auto &op) {
			    using Op = std::decay_t<decltype(op)>;
			    if constexpr(std::is_same<Op, circuit::Hadamard>::value) {
				    r.kind = TMFQS_HIP_OP_H;
				    r.q0 = op.target;
			    } else if constexpr(std::is_same<Op, circuit::PauliX>::value) {
				    r.kind = TMFQS_HIP_OP_X;
				    r.q0 = op.target;
			    } else if constexpr(std::is_same<Op, circuit::ControlledNot>::value) {
				    r.kind = TMFQS_HIP_OP_CX;
				    r.q0 = op.control;
				    r.q1 = op.target;
			    } else if constexpr(std::is_same<Op, circuit::ControlledPhaseShift>::value) {
				    r.kind = TMFQS_HIP_OP_CPHASE;
				    r.q0 = op.control;
				    r.q1 = op.target;
				    r.theta = op.theta;
				    ++L.params;
			    } else if constexpr(std::is_same<Op, circuit::Swap>::value) {
				    r.kind = TMFQS_HIP_OP_SWAP;
				    r.q0 = op.first;
				    r.q1 = op.second;
			    } else if constexpr(std::is_same<Op, circuit::MatrixGate>::value) {
				    const size_t k = op.targets.size();
				    const size_t dim = op.gate.dimension();
				    if(k == 0 || k > 5 || dim != (size_t(1) << k))
					    throw std::invalid_argument("adjointGradient: matrix gates must act on 1 to 5 qubits");
				    r.kind = TMFQS_HIP_OP_MATRIX;
				    r.k = static_cast<uint32_t>(k);
				    r.q0 = static_cast<uint32_t>(L.qubits.size());
				    r.state = L.mats.size();
				    for(QubitIndex q : op.targets) L.qubits.push_back(q);
				    for(size_t row = 0; row < dim; ++row)
					    for(size_t col = 0; col < dim; ++col) {
						    L.mats.push_back(op.gate[static_cast<unsigned>(row)][col].real);
						    L.mats.push_back(op.gate[static_cast<unsigned>(row)][col].imag);
					    }
			    } else {
				    throw std::invalid_argument("adjointGradient: only gate operations can be differentiated");
			    }
		    },
		    operation);
		L.ops.push_back(r);
	}
	return L;
}

}  // namespace

AdjointResult adjointGradient(const circuit::Circuit &circuit, QuantumRegister &reg, const DiagonalObservable &observable,
                              unsigned int checkpointGates) {
	if(circuit.qubitCount() != reg.qubitCount()) throw std::invalid_argument("adjointGradient: circuit and register sizes differ");
	const OpList L = toOpList(circuit);
	const std::vector<tmfqs_hip_diag_term> terms = toTerms(observable);
	tmfqs_hip_state *h = deviceHandle(reg, "adjointGradient");
	AdjointResult out;
	out.gradients.assign(L.params, 0.0);
	const auto t0 = std::chrono::steady_clock::now();
	raise(tmfqs_hip_adjoint(h, L.ops.data(), L.ops.size(), L.mats.empty() ? nullptr : L.mats.data(),
	                        L.qubits.empty() ? nullptr : L.qubits.data(), terms.data(), terms.size(), checkpointGates,
	                        &out.expectation, out.gradients.data(), out.gradients.size()));
	raise(tmfqs_hip_synchronize(h));
	out.executionSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	return out;
}

double expectation(QuantumRegister &reg, const DiagonalObservable &observable) {
	const std::vector<tmfqs_hip_diag_term> terms = toTerms(observable);
	tmfqs_hip_state *h = deviceHandle(reg, "expectation");
	double e = 0.0;
	raise(tmfqs_hip_expectation(h, terms.data(), terms.size(), &e));
	return e;
}

}  // namespace sim
}  // namespace tmfqs
