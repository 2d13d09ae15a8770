// Lazy state values (reference: src/tmfqs/state/quantum_state.cpp).  Basis,
// uniform-subset and amplitude-vector states answer queries from their host
// description; a simulation result (Register kind) forwards to its register.
#include "tmfqs/state/quantum_state.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <thread>

#include "tmfqs/config/register_config.h"
#include "tmfqs/core/constants.h"
#include "tmfqs/core/state_space.h"
#include "tmfqs/register/quantum_register.h"

namespace tmfqs {
namespace state {
namespace {

enum class StateKind { Basis, BasisAmplitude, UniformSubset, Amplitudes, Register };

// reference :27-33
double probabilityMass(const AmplitudesVector &a) {
	double t = 0.0;
	for(size_t e = 0; e + 1u < a.size(); e += 2u) t += a[e] * a[e] + a[e + 1u] * a[e + 1u];
	return t;
}

// reference :35-59
void validateOrNormalize(AmplitudesVector &a, const StateValidation &v) {
	if(a.size() % 2u != 0u) throw std::invalid_argument("Amplitude vector length must be even");
	if(v.normalization == NormalizationPolicy::AllowUnnormalized) return;
	if(v.normalization == NormalizationPolicy::Normalize) {
		const double t = probabilityMass(a);
		if(t <= 0.0) throw std::invalid_argument("Cannot normalize an all-zero amplitude vector");
		const double s = 1.0 / std::sqrt(t);
		for(double &x : a) x *= s;
		return;
	}
	if(std::abs(probabilityMass(a) - 1.0) > v.probabilityTolerance) throw std::invalid_argument("Amplitude vector is not normalized");
}

// A uniform random bit generator that replays 32-bit words drawn earlier from
// std::mt19937, so that std::uniform_real_distribution maps them exactly as it
// maps the engine's own output.
struct Replay {
	using result_type = std::mt19937::result_type;
	const result_type *p;
	static constexpr result_type min() { return std::mt19937::min(); }
	static constexpr result_type max() { return std::mt19937::max(); }
	result_type operator()() { return *p++; }
};

}  // namespace

struct QuantumState::Impl {
	unsigned int qubits = 0;
	StateKind kind = StateKind::Basis;
	StateIndex basisState = 0u;
	Amplitude basisAmplitude{1.0, 0.0};
	BasisStateList basisStates;
	AmplitudesVector amplitudes;
	std::unique_ptr<QuantumRegister> reg;

	Impl() = default;
	Impl(const Impl &o)
	    : qubits(o.qubits), kind(o.kind), basisState(o.basisState), basisAmplitude(o.basisAmplitude), basisStates(o.basisStates),
	      amplitudes(o.amplitudes) {
		if(o.reg) reg = std::make_unique<QuantumRegister>(*o.reg);
	}
	Impl(Impl &&) noexcept = default;
	Impl &operator=(Impl &&) noexcept = default;

	// Dense description of a uniform subset (dense_state_backend.cpp:92-115 semantics).
	AmplitudesVector uniformVector() const {
		const StateIndex N = checkedStateCount(qubits);
		AmplitudesVector a(static_cast<size_t>(2 * N), 0.0);
		std::vector<StateIndex> sel;
		for(StateIndex s : basisStates) {
			if(s >= N) throw std::out_of_range("QuantumRegister::initUniformSuperposition state index out of range");
			sel.push_back(s);
		}
		std::sort(sel.begin(), sel.end());
		sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
		if(sel.empty()) throw std::invalid_argument("QuantumRegister::initUniformSuperposition requires at least one basis state");
		const double v = 1.0 / std::sqrt(static_cast<double>(sel.size()));
		for(StateIndex s : sel) a[static_cast<size_t>(2 * s)] = v;
		return a;
	}
};

QuantumState::QuantumState() : impl_(std::make_unique<Impl>()) {}
QuantumState::QuantumState(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
QuantumState::QuantumState(const QuantumState &o) : impl_(std::make_unique<Impl>(*o.impl_)) {}
QuantumState &QuantumState::operator=(const QuantumState &o) {
	if(this != &o) impl_ = std::make_unique<Impl>(*o.impl_);
	return *this;
}
QuantumState::QuantumState(QuantumState &&) noexcept = default;
QuantumState &QuantumState::operator=(QuantumState &&) noexcept = default;
QuantumState::~QuantumState() = default;

unsigned int QuantumState::qubitCount() const { return impl_->qubits; }
StateIndex QuantumState::stateCount() const { return checkedStateCount(impl_->qubits); }
const QuantumRegister *QuantumState::registerPtr() const { return impl_->reg.get(); }

Amplitude QuantumState::amplitude(StateIndex s) const {
	if(s >= stateCount()) throw std::invalid_argument("Amplitude state index is out of range");
	switch(impl_->kind) {
		case StateKind::Amplitudes: return {impl_->amplitudes[2 * s], impl_->amplitudes[2 * s + 1]};
		case StateKind::Basis:
		case StateKind::BasisAmplitude: return s == impl_->basisState ? impl_->basisAmplitude : Amplitude{0.0, 0.0};
		case StateKind::Register: return impl_->reg->amplitude(s);
		case StateKind::UniformSubset: {
			const AmplitudesVector a = impl_->uniformVector();
			return {a[2 * s], a[2 * s + 1]};
		}
	}
	return {0.0, 0.0};
}

double QuantumState::totalProbability() const {
	switch(impl_->kind) {
		case StateKind::Amplitudes: return probabilityMass(impl_->amplitudes);
		case StateKind::Basis:
		case StateKind::BasisAmplitude:
			return impl_->basisAmplitude.real * impl_->basisAmplitude.real + impl_->basisAmplitude.imag * impl_->basisAmplitude.imag;
		case StateKind::Register: return impl_->reg->totalProbability();
		case StateKind::UniformSubset: return probabilityMass(impl_->uniformVector());
	}
	return 0.0;
}

AmplitudesVector QuantumState::amplitudes() const {
	switch(impl_->kind) {
		case StateKind::Amplitudes: return impl_->amplitudes;
		case StateKind::Register: return impl_->reg->amplitudes();
		case StateKind::UniformSubset: return impl_->uniformVector();
		case StateKind::Basis:
		case StateKind::BasisAmplitude: {
			AmplitudesVector a(2 * static_cast<size_t>(stateCount()), 0.0);
			a[2 * impl_->basisState] = impl_->basisAmplitude.real;
			a[2 * impl_->basisState + 1] = impl_->basisAmplitude.imag;
			return a;
		}
	}
	return {};
}

StateIndex QuantumState::measure(IRandomSource &src) const {
	switch(impl_->kind) {
		case StateKind::Basis:
		case StateKind::BasisAmplitude: return impl_->basisState;
		case StateKind::Register: return impl_->reg->measure(src);
		case StateKind::Amplitudes:
		case StateKind::UniformSubset: {
			const double rnd = src.nextUnitDouble();
			if(impl_->kind == StateKind::UniformSubset && (rnd < 0.0 || rnd >= 1.0))
				throw std::invalid_argument("DenseStateBackend::sampleMeasurement requires rnd in [0,1)");
			const AmplitudesVector own = impl_->kind == StateKind::UniformSubset ? impl_->uniformVector() : AmplitudesVector{};
			const AmplitudesVector &a = impl_->kind == StateKind::UniformSubset ? own : impl_->amplitudes;
			double cum = 0.0;
			for(StateIndex s = 0; s < stateCount(); ++s) {
				cum += a[2 * s] * a[2 * s] + a[2 * s + 1] * a[2 * s + 1];
				if(rnd <= cum) return s;
			}
			throw std::runtime_error("QuantumState::measure cumulative probability did not reach sample");
		}
	}
	return 0;
}

void QuantumState::setAmplitude(StateIndex s, Amplitude a, const StateValidation &v) {
	AmplitudesVector values = amplitudes();
	if(s >= stateCount()) throw std::invalid_argument("Amplitude state index is out of range");
	values[2 * s] = a.real;
	values[2 * s + 1] = a.imag;
	loadAmplitudes(std::move(values), v);
}

void QuantumState::loadAmplitudes(AmplitudesVector a, const StateValidation &v) {
	if(a.size() != checkedAmplitudeElementCount(impl_->qubits)) throw std::invalid_argument("Amplitude vector size does not match state qubit count");
	validateOrNormalize(a, v);
	impl_->kind = StateKind::Amplitudes;
	impl_->amplitudes = std::move(a);
	impl_->reg.reset();
}

QuantumState QuantumState::basis(unsigned int qubits, StateIndex s) {
	if(s >= checkedStateCount(qubits)) throw std::invalid_argument("Basis state is out of range");
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->kind = StateKind::Basis;
	impl->basisState = s;
	return QuantumState(std::move(impl));
}

QuantumState QuantumState::basisAmplitude(unsigned int qubits, StateIndex s, Amplitude a) {
	if(s >= checkedStateCount(qubits)) throw std::invalid_argument("Basis state is out of range");
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->kind = StateKind::BasisAmplitude;
	impl->basisState = s;
	impl->basisAmplitude = a;
	return QuantumState(std::move(impl));
}

QuantumState QuantumState::uniformSubset(unsigned int qubits, BasisStateList states) {
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->kind = StateKind::UniformSubset;
	impl->basisStates = std::move(states);
	return QuantumState(std::move(impl));
}

QuantumState QuantumState::fromAmplitudes(unsigned int qubits, AmplitudesVector a, const StateValidation &v) {
	if(a.size() != checkedAmplitudeElementCount(qubits)) throw std::invalid_argument("Amplitude vector size does not match qubit count");
	validateOrNormalize(a, v);
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->kind = StateKind::Amplitudes;
	impl->amplitudes = std::move(a);
	return QuantumState(std::move(impl));
}

// reference :264-277: phase_k from mt19937(seed) through
// uniform_real_distribution(0, 2 acos(-1)) in index order, amplitude
// 2^(-n/2) (cos, sin).  The engine's words are drawn sequentially in chunks;
// the distribution and cos/sin run on worker threads over a replay of those
// words, which yields the same values as the sequential loop.
QuantumState QuantumState::randomPhase(unsigned int qubits, unsigned int seed) {
	const StateIndex N = checkedStateCount(qubits);
	std::mt19937 gen(seed);
	const double hi = 2.0 * std::acos(-1.0);
	const double norm = 1.0 / std::sqrt(static_cast<double>(N));
	AmplitudesVector a(static_cast<size_t>(2 * N), 0.0);
	// generate_canonical<double, 53> draws two 32-bit words per value.
	const StateIndex chunk = std::min<StateIndex>(N, StateIndex{1} << 22);
	std::vector<std::mt19937::result_type> words;
	unsigned hw = std::thread::hardware_concurrency();
	const unsigned nthreads = std::max(1u, std::min(hw == 0 ? 1u : hw, 16u));
	for(StateIndex s0 = 0; s0 < N; s0 += chunk) {
		const StateIndex cnt = std::min<StateIndex>(chunk, N - s0);
		words.resize(static_cast<size_t>(2 * cnt));
		for(auto &w : words) w = gen();
		auto work = [&](StateIndex lo, StateIndex hiIdx) {
			std::uniform_real_distribution<double> dist(0.0, hi);
			Replay r{words.data() + 2 * lo};
			for(StateIndex i = lo; i < hiIdx; ++i) {
				const double ph = dist(r);
				double sn, cs;
				__builtin_sincos(ph, &sn, &cs);  // one sincos call, as the reference's build emits
				a[static_cast<size_t>(2 * (s0 + i))] = norm * cs;
				a[static_cast<size_t>(2 * (s0 + i) + 1)] = norm * sn;
			}
		};
		if(nthreads == 1 || cnt < 65536) {
			work(0, cnt);
		} else {
			std::vector<std::thread> ts;
			const StateIndex per = (cnt + nthreads - 1) / nthreads;
			for(unsigned t = 0; t < nthreads; ++t) {
				const StateIndex lo = t * per, hiI = std::min<StateIndex>(cnt, lo + per);
				if(lo < hiI) ts.emplace_back(work, lo, hiI);
			}
			for(auto &t : ts) t.join();
		}
	}
	return fromAmplitudes(qubits, std::move(a));
}

QuantumState QuantumState::sparsePattern(unsigned int qubits) {
	const StateIndex N = checkedStateCount(qubits);
	std::vector<StateIndex> sel;
	for(StateIndex x = 0;; ++x) {
		const StateIndex s = 8u * x + (x % 2u);
		if(s >= N) break;
		sel.push_back(s);
	}
	return uniformSubset(qubits, BasisStateList(std::move(sel)));
}

// reference :290-307
QuantumRegister QuantumState::materialize(const RegisterConfig &cfg) const {
	switch(impl_->kind) {
		case StateKind::Basis: return QuantumRegister(impl_->qubits, impl_->basisState, cfg);
		case StateKind::BasisAmplitude: return QuantumRegister(impl_->qubits, impl_->basisState, impl_->basisAmplitude, cfg);
		case StateKind::UniformSubset: {
			QuantumRegister reg(impl_->qubits, cfg);
			reg.initUniformSuperposition(impl_->basisStates);
			return reg;
		}
		case StateKind::Amplitudes: return QuantumRegister(impl_->qubits, impl_->amplitudes, cfg);
		case StateKind::Register: return *impl_->reg;
	}
	return QuantumRegister(impl_->qubits, cfg);
}

QuantumState QuantumState::fromRegister(QuantumRegister reg) {
	auto impl = std::make_unique<Impl>();
	impl->qubits = reg.qubitCount();
	impl->kind = StateKind::Register;
	impl->reg = std::make_unique<QuantumRegister>(std::move(reg));
	return QuantumState(std::move(impl));
}

}  // namespace state
}  // namespace tmfqs
