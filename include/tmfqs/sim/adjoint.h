// Adjoint differentiation of diagonal-observable expectations on the device
// backend.  No counterpart in the reference, which only runs and compares
// circuits; the circuit, register and configuration types are the
// reference's.
#ifndef TMFQS_SIM_ADJOINT_H
#define TMFQS_SIM_ADJOINT_H

#include <vector>

#include "tmfqs/circuit/circuit.h"
#include "tmfqs/core/types.h"

namespace tmfqs {

class QuantumRegister;

namespace sim {

/** O = sum of coef * prod_{q in qubits} Z_q plus sum of coef * |state><state|
 *  (Z_q is +1 when qubit q is 0; qubit numbering as everywhere in the API). */
struct DiagonalObservable {
	struct ZString {
		QubitList qubits;
		double coef = 1.0;
	};
	struct Projector {
		StateIndex state = 0;
		double coef = 1.0;
	};
	std::vector<ZString> zStrings;
	std::vector<Projector> projectors;
};

struct AdjointResult {
	/** <psi|O|psi> of the final state. */
	double expectation = 0.0;
	/** dE/dtheta of every ControlledPhaseShift operation, in circuit order. */
	std::vector<double> gradients;
	/** Wall seconds of the forward and backward sweeps (device synchronised). */
	double executionSeconds = 0.0;
};

/** Runs the circuit on the register (which must use StorageStrategyKind::Hip
 *  and must hold the initial state) and returns E and its gradient.  The
 *  register holds the final state afterwards, as after Simulator::run.
 *  checkpointGates = 0 stores a forward checkpoint about every sqrt(G) gates.
 *  Gate operations only (H, X, CX, controlled phase, swap, matrix gates). */
AdjointResult adjointGradient(const circuit::Circuit &circuit, QuantumRegister &reg, const DiagonalObservable &observable,
                              unsigned int checkpointGates = 0);

/** <psi|O|psi> of the register's current state (Hip strategy). */
double expectation(QuantumRegister &reg, const DiagonalObservable &observable);

}  // namespace sim
}  // namespace tmfqs

#endif
