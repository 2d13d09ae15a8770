// StorageStrategyKind::HipChunked: an IStateBackend over the chunked-register
// entry points of libtmfqs_hip.so (tmfqs_hip_chunked_*, include/tmfqs_hip.h).
//
// The reference's chunked strategies (CompressedStateBackend,
// include/tmfqs/storage/common/compressed_state_backend.h:103-747) apply each
// gate through their chunk cache and flush the cache after every operation
// unless an operation batch is open (batchDepth_, flushCacheIfNeeded).  Here
// gates are records queued on the host and run as one operation list when the
// outermost batch ends, or at once outside a batch, so a Simulator::run costs
// one window sweep per segment of the circuit; every read first runs what is
// queued.  Reads of single amplitudes and streamed exports come from the
// host-resident chunks.
#include <algorithm>
#include <cmath>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmfqs/core/state_space.h"
#include "tmfqs/core/validation.h"
#include "tmfqs/storage/hip/hip_state_backend.h"
#include "tmfqs_hip.h"

namespace tmfqs {
namespace {

void raise(int rc) { raiseHipStatus(rc); }

unsigned log2Exact(size_t v, const char *what) {
	if(v < 2 || (v & (v - 1)) != 0) throw std::invalid_argument(std::string("HipChunkedStateBackend: ") + what + " must be a power of two >= 2");
	unsigned k = 0;
	while((size_t{1} << k) < v) ++k;
	return k;
}

class HipChunkedStateBackend final : public IStateBackend {
	public:
		explicit HipChunkedStateBackend(const RegisterConfig &cfg) : cfg_(cfg), hc_(toHipConfig(cfg.hip)) {}
		~HipChunkedStateBackend() override {
			if(c_) tmfqs_hip_chunked_destroy(c_);
		}
		HipChunkedStateBackend(const HipChunkedStateBackend &) = delete;
		HipChunkedStateBackend &operator=(const HipChunkedStateBackend &) = delete;

		std::unique_ptr<IStateBackend> clone() const override {
			auto b = std::make_unique<HipChunkedStateBackend>(cfg_);
			if(c_) {
				flush();
				raise(tmfqs_hip_chunked_clone(c_, &b->c_));
				b->n_ = n_;
			}
			return b;
		}

		void initZero(unsigned int n) override {
			reset(n);
			raise(tmfqs_hip_chunked_init_basis(c_, 0, 0.0, 0.0));
		}
		void initBasis(unsigned int n, StateIndex s, Amplitude a) override {
			reset(n);
			raise(tmfqs_hip_chunked_init_basis(c_, s, a.real, a.imag));
		}
		// Equal weight on every distinct listed state, as the device backend does.
		void initUniformSuperposition(unsigned int n, const BasisStateList &states) override {
			std::vector<StateIndex> sel(states.begin(), states.end());
			const StateIndex N = checkedStateCount(n);
			for(StateIndex s : sel)
				if(s >= N) throw std::out_of_range("HipChunkedStateBackend::initUniformSuperposition state index out of range");
			std::sort(sel.begin(), sel.end());
			sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
			if(sel.empty()) throw std::invalid_argument("HipChunkedStateBackend::initUniformSuperposition requires at least one state");
			reset(n);
			const double amp = 1.0 / std::sqrt(static_cast<double>(sel.size()));
			raise(tmfqs_hip_chunked_init_basis(c_, sel[0], amp, 0.0));
			for(size_t i = 1; i < sel.size(); ++i) raise(tmfqs_hip_chunked_set_amplitude(c_, sel[i], amp, 0.0));
		}
		void loadAmplitudes(unsigned int n, AmplitudesVector a) override {
			if(a.size() != checkedAmplitudeElementCount(n)) throw std::invalid_argument("HipChunkedStateBackend: amplitudes size mismatch");
			reset(n);
			raise(tmfqs_hip_chunked_load(c_, a.data(), a.size()));
		}
		void exportAmplitudes(AmplitudesVector &out) const override {
			ready("export");
			const StateIndex N = checkedStateCount(n_);
			out.resize(static_cast<size_t>(2 * N));
			raise(tmfqs_hip_chunked_export(c_, 0, N, out.data()));
		}
		void forEachAmplitudeChunk(const AmplitudeChunkVisitor &visitor) const override {
			ready("amplitude chunk iteration");
			const StateIndex N = checkedStateCount(n_);
			const StateIndex per = N < (StateIndex{1} << 20) ? N : (StateIndex{1} << 20);
			std::vector<double> buf(static_cast<size_t>(2 * per));
			for(StateIndex s0 = 0; s0 < N; s0 += per) {
				const StateIndex cnt = per < N - s0 ? per : N - s0;
				raise(tmfqs_hip_chunked_export(c_, s0, cnt, buf.data()));
				visitor(s0, buf.data(), static_cast<size_t>(2 * cnt));
			}
		}
		Amplitude amplitude(StateIndex s) const override {
			ready("amplitude");
			double v[2];
			raise(tmfqs_hip_chunked_export(c_, s, 1, v));
			return {v[0], v[1]};
		}
		void setAmplitude(StateIndex s, Amplitude a) override {
			ready("setAmplitude");
			raise(tmfqs_hip_chunked_set_amplitude(c_, s, a.real, a.imag));
		}
		double probability(StateIndex s) const override {
			const Amplitude a = amplitude(s);
			return a.real * a.real + a.imag * a.imag;
		}
		double totalProbability() const override {
			ready("probability summation");
			double p = 0.0;
			raise(tmfqs_hip_chunked_total_probability(c_, &p));
			return p;
		}
		StateIndex sampleMeasurement(double rnd) const override {
			ready("measurement");
			uint64_t s = 0;
			raise(tmfqs_hip_chunked_sample(c_, rnd, &s));
			return s;
		}
		void printNonZeroStates(std::ostream &os, double epsilon) const override {
			if(epsilon < 0.0) throw std::invalid_argument("HipChunkedStateBackend::printNonZeroStates epsilon must be non-negative");
			forEachAmplitudeChunk([&](StateIndex base, const double *d, size_t elems) {
				for(size_t i = 0; i + 1 < elems; i += 2)
					if(std::abs(d[i]) > epsilon || std::abs(d[i + 1]) > epsilon)
						os << (base + i / 2) << ": " << d[i] << " + " << d[i + 1] << "i\n";
			});
		}
		void phaseFlipBasisState(StateIndex s) override {
			if(s >= checkedStateCount(initialized())) throw std::out_of_range("HipChunkedStateBackend::phaseFlipBasisState state index out of range");
			record(TMFQS_HIP_OP_PHASE_FLIP, 0, 0, 0.0, s);
		}
		void inversionAboutMean() override { record(TMFQS_HIP_OP_DIFFUSE_MATERIALIZED, 0, 0, 0.0, 0); }
		void inversionAboutMean(Amplitude m) override {
			ready("inversion about mean");
			raise(tmfqs_hip_chunked_inversion_about_mean_value(c_, m.real, m.imag));
		}
		void applyAffineTransform(Amplitude s, Amplitude b) override {
			ready("affine transform");
			raise(tmfqs_hip_chunked_affine(c_, s.real, s.imag, b.real, b.imag));
		}
		void applyHadamard(QubitIndex q) override { record(TMFQS_HIP_OP_H, qubit(q, "applyHadamard"), 0, 0.0, 0); }
		void applyPauliX(QubitIndex q) override { record(TMFQS_HIP_OP_X, qubit(q, "applyPauliX"), 0, 0.0, 0); }
		void applyControlledPhaseShift(QubitIndex c, QubitIndex t, double theta) override {
			pair(c, t, "applyControlledPhaseShift");
			record(TMFQS_HIP_OP_CPHASE, c, t, theta, 0);
		}
		void applyControlledNot(QubitIndex c, QubitIndex t) override {
			pair(c, t, "applyControlledNot");
			record(TMFQS_HIP_OP_CX, c, t, 0.0, 0);
		}
		void applyGate(const QuantumGate &gate, const QubitList &qubits) override {
			validateGateTargets("HipChunkedStateBackend::applyGate", qubits, initialized(), gate.dimension());
			if(qubits.size() > 5) throw std::invalid_argument("HipChunkedStateBackend::applyGate takes at most five qubits");
			tmfqs_hip_op o = blank(TMFQS_HIP_OP_MATRIX);
			o.q0 = static_cast<uint32_t>(qpool_.size());
			o.k = static_cast<uint32_t>(qubits.size());
			o.state = mats_.size();
			for(QubitIndex q : qubits) qpool_.push_back(q);
			for(const Amplitude &a : gate.entries()) {
				mats_.push_back(a.real);
				mats_.push_back(a.imag);
			}
			queue_.push_back(o);
			if(!batch_) flush();
		}
		void beginOperationBatch() override { ++batch_; }
		void endOperationBatch() override {
			if(batch_ && --batch_ == 0) flush();
		}
		bool amplitudeSum(Amplitude &out) const override {
			ready("amplitude chunk iteration");
			raise(tmfqs_hip_chunked_amplitude_sum(c_, &out.real, &out.imag));
			return true;
		}
		bool applySwap(QubitIndex a, QubitIndex b) override {
			pair(a, b, "applySwap");
			record(TMFQS_HIP_OP_SWAP, a, b, 0.0, 0);
			return true;
		}
		void synchronize() override {
			if(c_) flush();
		}

	private:
		RegisterConfig cfg_;
		tmfqs_hip_config hc_;
		tmfqs_hip_chunked *c_ = nullptr;
		unsigned n_ = 0;
		unsigned batch_ = 0;
		mutable std::vector<tmfqs_hip_op> queue_;
		mutable std::vector<double> mats_;
		mutable std::vector<unsigned> qpool_;

		// A fresh register of n qubits with the configured geometry.
		void reset(unsigned n) {
			checkedStateCount(n);
			unsigned c = n < 24 ? n : 24u;
			if(cfg_.hip.chunkStates) c = log2Exact(cfg_.hip.chunkStates, "chunkStates");
			if(c > n) c = n;
			unsigned g = 4;
			if(cfg_.hip.windowChunks) g = log2Exact(cfg_.hip.windowChunks, "windowChunks");
			if(c + g > n) g = n - c;
			queue_.clear();
			mats_.clear();
			qpool_.clear();
			if(c_) {
				raise(tmfqs_hip_chunked_destroy(c_));
				c_ = nullptr;
			}
			raise(tmfqs_hip_chunked_create(&hc_, n, c, g ? g : 1u, &c_));
			n_ = n;
		}
		unsigned initialized() const {
			if(!c_) throw std::logic_error("HipChunkedStateBackend is not initialized");
			return n_;
		}
		// Runs what is queued, then the caller reads or updates the chunks.
		void ready(const char *what) const {
			if(!c_) throw std::logic_error(std::string("HipChunkedStateBackend is not initialized for ") + what);
			flush();
		}
		void flush() const {
			if(queue_.empty()) return;
			std::vector<tmfqs_hip_op> ops;
			ops.swap(queue_);
			std::vector<double> mats;
			mats.swap(mats_);
			std::vector<unsigned> qpool;
			qpool.swap(qpool_);
			raise(tmfqs_hip_chunked_run(c_, ops.data(), ops.size(), mats.empty() ? nullptr : mats.data(),
			                            qpool.empty() ? nullptr : qpool.data()));
		}
		unsigned qubit(QubitIndex q, const char *what) const {
			if(q >= initialized()) throw std::out_of_range(std::string("HipChunkedStateBackend::") + what + " qubit index out of range");
			return q;
		}
		void pair(QubitIndex a, QubitIndex b, const char *what) const {
			qubit(a, what);
			qubit(b, what);
			if(a == b) throw std::invalid_argument(std::string("HipChunkedStateBackend::") + what + " requires distinct qubits");
		}
		static tmfqs_hip_op blank(int kind) {
			tmfqs_hip_op o{};
			o.kind = kind;
			return o;
		}
		void record(int kind, unsigned q0, unsigned q1, double theta, uint64_t state) {
			initialized();
			tmfqs_hip_op o = blank(kind);
			o.q0 = q0;
			o.q1 = q1;
			o.theta = theta;
			o.state = state;
			queue_.push_back(o);
			if(!batch_) flush();
		}
};

}  // namespace

std::unique_ptr<IStateBackend> createHipChunkedStateBackend(const RegisterConfig &cfg) {
	if(!isHipStateBackendAvailable()) return nullptr;
	if(cfg.hip.chunkStates) log2Exact(cfg.hip.chunkStates, "chunkStates");
	if(cfg.hip.windowChunks) log2Exact(cfg.hip.windowChunks, "windowChunks");
	return std::make_unique<HipChunkedStateBackend>(cfg);
}

}  // namespace tmfqs
