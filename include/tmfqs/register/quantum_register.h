// Register facade (reference: include/tmfqs/register/quantum_register.h:24-251).
#ifndef TMFQS_REGISTER_QUANTUM_REGISTER_H
#define TMFQS_REGISTER_QUANTUM_REGISTER_H

#include <cstddef>
#include <functional>
#include <iosfwd>
#include <memory>
#include <optional>

#include "tmfqs/config/register_config.h"
#include "tmfqs/core/types.h"
#include "tmfqs/gates/quantum_gate.h"
#include "tmfqs/storage/i_state_backend.h"

namespace tmfqs {

class IRandomSource;

class QuantumRegister {
	public:
		QuantumRegister();
		explicit QuantumRegister(unsigned int numQubits, const RegisterConfig &cfg = {});
		QuantumRegister(unsigned int numQubits, StateIndex initState, const RegisterConfig &cfg = {});
		QuantumRegister(unsigned int numQubits, StateIndex initState, Amplitude amp, const RegisterConfig &cfg = {});
		QuantumRegister(unsigned int numQubits, AmplitudesVector amplitudes, const RegisterConfig &cfg = {});
		QuantumRegister(const QuantumRegister &);
		QuantumRegister(QuantumRegister &&) noexcept = default;
		QuantumRegister &operator=(const QuantumRegister &);
		QuantumRegister &operator=(QuantumRegister &&) noexcept = default;
		~QuantumRegister();

		unsigned int qubitCount() const;
		StateIndex stateCount() const;
		size_t amplitudeElementCount() const;

		double probability(StateIndex state) const;
		double totalProbability() const;
		Amplitude amplitude(StateIndex state) const;
		AmplitudesVector amplitudes() const;
		StateIndex measure(IRandomSource &randomSource) const;

		void setAmplitude(StateIndex state, Amplitude amp);
		void loadAmplitudes(AmplitudesVector amplitudes);
		void initUniformSuperposition(const BasisStateList &basisStates);

		StorageStrategyKind storageStrategy() const;

		void printStatesVector(double epsilon = 1e-12) const;
		friend std::ostream &operator<<(std::ostream &os, const QuantumRegister &reg);

		void applyGate(const QuantumGate &gate, const QubitList &qubits);
		void applyPhaseFlipBasisState(StateIndex state);
		void applyInversionAboutMean();
		void applyInversionAboutMean(Amplitude mean);
		void applyInversionAboutMeanMaterialized();
		void applyInversionAboutMeanMaterialized(Amplitude mean);
		void applyHadamard(QubitIndex qubit);
		void applyPauliX(QubitIndex qubit);
		void applyControlledPhaseShift(QubitIndex controlQubit, QubitIndex targetQubit, double theta);
		void applyControlledNot(QubitIndex controlQubit, QubitIndex targetQubit);
		void applySwap(QubitIndex qubit1, QubitIndex qubit2);
		void beginOperationBatch();
		void endOperationBatch();

		/** The storage backend (device handles are reachable through hipBackendHandle). */
		IStateBackend *backend() const noexcept { return backend_.get(); }
		/** True when no lazy affine overlay is pending (stored == logical amplitudes). */
		bool overlayIsIdentity() const noexcept { return !overlay_.has_value(); }
		/** Hands the backend to an operation outside the register's gate set (the
		 *  adjoint gradient): flushes a pending overlay and drops the cached
		 *  amplitude sum, since the caller is about to change the state. */
		IStateBackend *beginExternalUpdate();

	private:
		// What the backend stores and what the register reports can differ by
		// an affine map, logical = scale * stored + bias: the lazy diffusion
		// keeps reflections about the mean in this map instead of sweeping
		// the state.  No map means stored == logical.
		struct AffineMap {
			Amplitude scale{1.0, 0.0};
			Amplitude bias{0.0, 0.0};
			Amplitude toLogical(Amplitude stored) const;
			Amplitude toStored(Amplitude logical) const;
			AffineMap reflectedAbout(Amplitude mean) const;
		};

		unsigned int numQubits_ = 0;
		StateIndex numStates_ = 0;
		RegisterConfig config_{};
		StorageStrategyKind resolvedStrategy_ = StorageStrategyKind::Dense;
		std::unique_ptr<IStateBackend> backend_;
		std::optional<AffineMap> overlay_;
		// The exact sequential sum of the logical amplitudes when it is known
		// (Grover's mean); any gate forgets it.
		mutable std::optional<Amplitude> logicalSum_;

		QuantumRegister(unsigned int numQubits, const RegisterConfig &cfg,
		                const std::function<void(IStateBackend &, unsigned int)> &init);
		IStateBackend &backendFor(const char *operation) const;
		void collapseOverlay();
		void gateUpdate(const char *operation, const std::function<void()> &validate,
		                const std::function<void(IStateBackend &)> &apply);
		Amplitude logicalSum() const;
		void reflectSum(Amplitude mean) const;
		bool scanLogical(const std::function<bool(StateIndex, Amplitude)> &visit) const;
		void printLogical(std::ostream &os, double epsilon) const;
};

}  // namespace tmfqs

#endif
