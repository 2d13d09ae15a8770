// Storage plugin interface (reference: include/tmfqs/storage/i_state_backend.h:22-164).
// The 24 virtuals keep the reference's signatures and semantics.  The optional
// virtuals after them let a device backend answer register-level reads without
// streaming the state to the host; their defaults report "not supported" and
// QuantumRegister then falls back to the reference's host loops.
#ifndef TMFQS_STORAGE_I_STATE_BACKEND_H
#define TMFQS_STORAGE_I_STATE_BACKEND_H

#include <functional>
#include <iosfwd>
#include <memory>

#include "tmfqs/core/types.h"
#include "tmfqs/gates/quantum_gate.h"

namespace tmfqs {

using AmplitudeChunkVisitor = std::function<void(StateIndex baseState, const double *data, size_t elemCount)>;

class IStateBackend {
	public:
		virtual ~IStateBackend() = default;

		virtual std::unique_ptr<IStateBackend> clone() const = 0;

		virtual void initZero(unsigned int numQubits) = 0;
		virtual void initBasis(unsigned int numQubits, StateIndex initState, Amplitude amp) = 0;
		virtual void initUniformSuperposition(unsigned int numQubits, const BasisStateList &basisStates) = 0;
		virtual void loadAmplitudes(unsigned int numQubits, AmplitudesVector amplitudes) = 0;
		virtual void exportAmplitudes(AmplitudesVector &out) const = 0;
		virtual void forEachAmplitudeChunk(const AmplitudeChunkVisitor &visitor) const = 0;

		virtual Amplitude amplitude(StateIndex state) const = 0;
		virtual void setAmplitude(StateIndex state, Amplitude amp) = 0;
		virtual double probability(StateIndex state) const = 0;
		virtual double totalProbability() const = 0;
		virtual StateIndex sampleMeasurement(double rnd) const = 0;
		virtual void printNonZeroStates(std::ostream &os, double epsilon) const = 0;

		virtual void phaseFlipBasisState(StateIndex state) = 0;
		virtual void inversionAboutMean() = 0;
		virtual void inversionAboutMean(Amplitude mean) = 0;
		virtual void applyAffineTransform(Amplitude scale, Amplitude bias) = 0;
		virtual void applyHadamard(QubitIndex qubit) = 0;
		virtual void applyPauliX(QubitIndex qubit) = 0;
		virtual void applyControlledPhaseShift(QubitIndex controlQubit, QubitIndex targetQubit, double theta) = 0;
		virtual void applyControlledNot(QubitIndex controlQubit, QubitIndex targetQubit) = 0;
		virtual void applyGate(const QuantumGate &gate, const QubitList &qubits) = 0;

		virtual void beginOperationBatch() {}
		virtual void endOperationBatch() {}

		// ---- optional accelerated operations (no counterpart in the reference) ----
		/** Sum of all stored amplitudes; false if unsupported. */
		virtual bool amplitudeSum(Amplitude &) const { return false; }
		/** Sequential-order sum of |scale*a + bias|^2; false if unsupported. */
		virtual bool totalProbabilityAffine(Amplitude, Amplitude, double &) const { return false; }
		/** First state whose sequential cumulative |scale*a + bias|^2 reaches rnd; false if unsupported. */
		virtual bool sampleMeasurementAffine(double, Amplitude, Amplitude, StateIndex &) const { return false; }
		/** Swap of two qubits, identical to three controlled-NOTs; false if unsupported. */
		virtual bool applySwap(QubitIndex, QubitIndex) { return false; }
		/** Waits until every queued operation has completed. */
		virtual void synchronize() {}
};

}  // namespace tmfqs

#endif
