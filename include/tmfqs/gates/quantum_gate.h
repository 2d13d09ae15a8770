// Square complex gate matrix, row-major (reference: include/tmfqs/gates/quantum_gate.h:17-137).
// Fused k-qubit blocks are composed on the host in FP64 with operator*.
#ifndef TMFQS_GATES_QUANTUM_GATE_H
#define TMFQS_GATES_QUANTUM_GATE_H

#include <iosfwd>
#include <vector>

#include "tmfqs/core/types.h"

namespace tmfqs {

class QuantumGate {
	public:
		QuantumGate() = default;
		explicit QuantumGate(unsigned int dimension);
		QuantumGate(const QuantumGate &) = default;
		QuantumGate(QuantumGate &&) noexcept = default;
		QuantumGate &operator=(const QuantumGate &) = default;
		QuantumGate &operator=(QuantumGate &&) noexcept = default;
		~QuantumGate() = default;

		/** Row i; throws std::out_of_range if i >= dimension(). */
		Amplitude *operator[](unsigned int i);
		const Amplitude *operator[](unsigned int i) const;
		unsigned int dimension() const noexcept;
		/** Row-major entries (dimension()^2). */
		const std::vector<Amplitude> &entries() const noexcept { return matrix_; }

		QuantumGate operator*(Amplitude x) const;
		/** Matrix product this * qg; throws std::invalid_argument on a size mismatch. */
		QuantumGate operator*(const QuantumGate &qg) const;

		void printQuantumGate() const;
		friend std::ostream &operator<<(std::ostream &os, const QuantumGate &qg);

		static QuantumGate Identity(unsigned int dimension);
		static QuantumGate Hadamard();
		static QuantumGate PauliX();
		static QuantumGate PauliY();
		static QuantumGate PauliZ();
		static QuantumGate PhaseShift(double theta);
		static QuantumGate PiOverEight();
		static QuantumGate ControlledNot();
		static QuantumGate Toffoli();
		static QuantumGate ControlledPhaseShift(double theta);
		static QuantumGate Swap();
		static QuantumGate Ising(double theta);

	private:
		std::vector<Amplitude> matrix_;
		unsigned int dimension_ = 0;
};

QuantumGate operator*(Amplitude x, const QuantumGate &U);

}  // namespace tmfqs

#endif
