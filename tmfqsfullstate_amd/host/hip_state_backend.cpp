// IStateBackend over the C-ABI of libtmfqs_hip.so.  Each method forwards to
// the matching tmfqs_hip_* entry point (include/tmfqs_hip.h) and turns its
// status back into the exception the reference's backends throw.
#include "tmfqs/storage/hip/hip_state_backend.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmfqs/core/state_space.h"
#include "tmfqs/core/validation.h"
#include "tmfqs_hip.h"

namespace tmfqs {

void raiseHipStatus(int rc) {
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

tmfqs_hip_config toHipConfig(const HipConfig &c) {
	tmfqs_hip_config h;
	tmfqs_hip_config_default(&h);
	h.precision = c.precision == HipPrecision::Complex64 ? TMFQS_HIP_COMPLEX64 : TMFQS_HIP_COMPLEX128;
	h.device = c.device;
	h.fusion = c.fusion ? 1 : 0;
	h.tile_qubits = static_cast<int>(c.tileQubits);
	h.coalesce_qubits = static_cast<int>(c.coalesceQubits);
	h.grover_block = static_cast<int>(c.groverBlock);
	h.exact_phases = c.exactPhases ? 1 : 0;
	h.fuse_qubits = static_cast<int>(c.fuseQubits);
	h.fuse_min_gates = static_cast<int>(c.fuseMinGates);
	h.qft_plan = static_cast<int>(c.qftPlan);
	return h;
}

namespace {

void raise(int rc) { raiseHipStatus(rc); }

class HipStateBackend final : public IStateBackend {
	public:
		explicit HipStateBackend(tmfqs_hip_state *h) : h_(h) {}
		~HipStateBackend() override { tmfqs_hip_destroy(h_); }
		HipStateBackend(const HipStateBackend &) = delete;
		HipStateBackend &operator=(const HipStateBackend &) = delete;

		tmfqs_hip_state *handle() const { return h_; }

		std::unique_ptr<IStateBackend> clone() const override {
			tmfqs_hip_state *c = nullptr;
			raise(tmfqs_hip_clone(h_, &c));
			return std::make_unique<HipStateBackend>(c);
		}

		void initZero(unsigned int n) override { raise(tmfqs_hip_init_zero(h_, n)); }
		void initBasis(unsigned int n, StateIndex s, Amplitude a) override { raise(tmfqs_hip_init_basis(h_, n, s, a.real, a.imag)); }
		void initUniformSuperposition(unsigned int n, const BasisStateList &states) override {
			raise(tmfqs_hip_init_uniform(h_, n, states.values().data(), states.size()));
		}
		void loadAmplitudes(unsigned int n, AmplitudesVector a) override { raise(tmfqs_hip_load(h_, n, a.data(), a.size())); }
		void exportAmplitudes(AmplitudesVector &out) const override {
			unsigned n = 0;
			raise(tmfqs_hip_qubit_count(h_, &n));
			const StateIndex N = checkedStateCount(n);
			out.resize(static_cast<size_t>(2 * N));
			raise(tmfqs_hip_export(h_, 0, N, out.data()));
		}
		// Streams the device state through a host window of 2^20 states.
		void forEachAmplitudeChunk(const AmplitudeChunkVisitor &visitor) const override {
			unsigned n = 0;
			raise(tmfqs_hip_qubit_count(h_, &n));
			const StateIndex N = checkedStateCount(n);
			const StateIndex per = std::min<StateIndex>(N, StateIndex{1} << 20);
			std::vector<double> buf(static_cast<size_t>(2 * per));
			for(StateIndex s0 = 0; s0 < N; s0 += per) {
				const StateIndex cnt = std::min<StateIndex>(per, N - s0);
				raise(tmfqs_hip_export(h_, s0, cnt, buf.data()));
				visitor(s0, buf.data(), static_cast<size_t>(2 * cnt));
			}
		}
		Amplitude amplitude(StateIndex s) const override {
			Amplitude a{};
			raise(tmfqs_hip_amplitude(h_, s, &a.real, &a.imag));
			return a;
		}
		void setAmplitude(StateIndex s, Amplitude a) override { raise(tmfqs_hip_set_amplitude(h_, s, a.real, a.imag)); }
		double probability(StateIndex s) const override {
			double p = 0;
			raise(tmfqs_hip_probability(h_, s, &p));
			return p;
		}
		double totalProbability() const override {
			double p = 0;
			raise(tmfqs_hip_total_probability(h_, &p));
			return p;
		}
		StateIndex sampleMeasurement(double rnd) const override {
			uint64_t s = 0;
			raise(tmfqs_hip_sample(h_, rnd, &s));
			return s;
		}
		// dense_state_backend.cpp:190-202 semantics over a streamed copy.
		void printNonZeroStates(std::ostream &os, double epsilon) const override {
			if(epsilon < 0.0) throw std::invalid_argument("HipStateBackend::printNonZeroStates epsilon must be non-negative");
			forEachAmplitudeChunk([&](StateIndex base, const double *d, size_t elems) {
				for(size_t i = 0; i + 1 < elems; i += 2)
					if(std::abs(d[i]) > epsilon || std::abs(d[i + 1]) > epsilon)
						os << (base + i / 2) << ": " << d[i] << " + " << d[i + 1] << "i\n";
			});
		}
		void phaseFlipBasisState(StateIndex s) override { raise(tmfqs_hip_phase_flip(h_, s)); }
		void inversionAboutMean() override { raise(tmfqs_hip_inversion_about_mean(h_)); }
		void inversionAboutMean(Amplitude m) override { raise(tmfqs_hip_inversion_about_mean_value(h_, m.real, m.imag)); }
		void applyAffineTransform(Amplitude s, Amplitude b) override { raise(tmfqs_hip_affine(h_, s.real, s.imag, b.real, b.imag)); }
		void applyHadamard(QubitIndex q) override { raise(tmfqs_hip_hadamard(h_, q)); }
		void applyPauliX(QubitIndex q) override { raise(tmfqs_hip_pauli_x(h_, q)); }
		void applyControlledPhaseShift(QubitIndex c, QubitIndex t, double theta) override {
			raise(tmfqs_hip_controlled_phase(h_, c, t, theta));
		}
		void applyControlledNot(QubitIndex c, QubitIndex t) override { raise(tmfqs_hip_controlled_not(h_, c, t)); }
		void applyGate(const QuantumGate &gate, const QubitList &qubits) override {
			unsigned n = 0;
			raise(tmfqs_hip_qubit_count(h_, &n));
			validateGateTargets("HipStateBackend::applyGate", qubits, n, gate.dimension());
			std::vector<double> m;
			m.reserve(2 * gate.entries().size());
			for(const Amplitude &a : gate.entries()) {
				m.push_back(a.real);
				m.push_back(a.imag);
			}
			std::vector<unsigned> q(qubits.begin(), qubits.end());
			raise(tmfqs_hip_apply_gate(h_, m.data(), static_cast<unsigned>(q.size()), q.data()));
		}
		void beginOperationBatch() override { raise(tmfqs_hip_begin_batch(h_)); }
		// The batch ends when the device has finished it, so a caller's clock
		// around the batch measures the work (Simulator::run, simulator.cpp:66-93).
		void endOperationBatch() override {
			raise(tmfqs_hip_end_batch(h_));
			raise(tmfqs_hip_synchronize(h_));
		}

		bool amplitudeSum(Amplitude &out) const override {
			raise(tmfqs_hip_amplitude_sum(h_, &out.real, &out.imag));
			return true;
		}
		bool totalProbabilityAffine(Amplitude s, Amplitude b, double &out) const override {
			raise(tmfqs_hip_total_probability_affine(h_, s.real, s.imag, b.real, b.imag, &out));
			return true;
		}
		bool sampleMeasurementAffine(double rnd, Amplitude s, Amplitude b, StateIndex &out) const override {
			uint64_t r = 0;
			raise(tmfqs_hip_sample_affine(h_, rnd, s.real, s.imag, b.real, b.imag, &r));
			out = r;
			return true;
		}
		bool applySwap(QubitIndex a, QubitIndex b) override {
			raise(tmfqs_hip_swap(h_, a, b));
			return true;
		}
		void synchronize() override { raise(tmfqs_hip_synchronize(h_)); }

	private:
		tmfqs_hip_state *h_;
};

}  // namespace

bool isHipStateBackendAvailable() {
	int c = 0;
	return tmfqs_hip_device_count(&c) == TMFQS_HIP_OK && c > 0;
}

std::unique_ptr<IStateBackend> createHipStateBackend(const RegisterConfig &cfg) {
	if(!isHipStateBackendAvailable()) return nullptr;
	const tmfqs_hip_config hc = toHipConfig(cfg.hip);
	tmfqs_hip_state *h = nullptr;
	raise(tmfqs_hip_create(&hc, &h));
	return std::make_unique<HipStateBackend>(h);
}

tmfqs_hip_state *hipBackendHandle(IStateBackend *backend) {
	auto *b = dynamic_cast<HipStateBackend *>(backend);
	return b ? b->handle() : nullptr;
}

}  // namespace tmfqs
