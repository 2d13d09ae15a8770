// QuantumState as a recipe: a value says how its amplitudes are made (a basis
// state, a uniform subset, an explicit vector, a seeded random-phase vector,
// or the register a simulation left behind) and only builds them where they
// are needed.  Queries on host-described states are answered on the host;
// materialize() builds the register on the requested backend, writing a
// vector straight from the recipe into device memory (no intermediate copy)
// when the backend is the Hip one; a random-phase state is generated there
// (tmfqs_hip_init_random_phase) without a host vector at all.
//
// Semantics follow src/tmfqs/state/quantum_state.cpp:44-315: normalization
// policies and their messages, measure() of a basis state draws nothing,
// uniform subsets behave as the dense backend's initUniformSuperposition
// (sorted, duplicates dropped, 1/sqrt(count)), and randomPhase draws its
// phases from std::mt19937(seed) through uniform_real_distribution(0,
// 2 acos(-1)) in index order.
#include "tmfqs/state/quantum_state.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <stdexcept>
#include <thread>
#include <type_traits>
#include <utility>
#include <variant>

#include "tmfqs/config/register_config.h"
#include "tmfqs/core/state_space.h"
#include "tmfqs/core/validation.h"
#include "tmfqs/register/quantum_register.h"
#include "tmfqs/storage/hip/hip_state_backend.h"
#include "tmfqs_hip.h"

namespace tmfqs {
namespace state {
namespace {

struct BasisRecipe {
	StateIndex state = 0;
	Amplitude amp{1.0, 0.0};
};
struct SubsetRecipe {
	BasisStateList states;
};
struct VectorRecipe {
	AmplitudesVector amps;
};
struct PhaseRecipe {
	unsigned int seed = 0;
};
// A simulation result: registers are never changed through a QuantumState,
// so copies of the value share one register (and its device memory).
struct ResultRecipe {
	std::shared_ptr<const QuantumRegister> reg;
};
using Recipe = std::variant<BasisRecipe, SubsetRecipe, VectorRecipe, PhaseRecipe, ResultRecipe>;

template <class... F> struct Visitor : F... { using F::operator()...; };
template <class... F> Visitor(F...) -> Visitor<F...>;

// Sequential sum of |a|^2 in index order (the reference's probability mass).
double massOf(const AmplitudesVector &a) {
	double m = 0.0;
	for(size_t e = 0; e + 1u < a.size(); e += 2u) m += a[e] * a[e] + a[e + 1u] * a[e + 1u];
	return m;
}

void applyValidation(AmplitudesVector &a, const StateValidation &v) {
	if(a.size() % 2u) throw std::invalid_argument("Amplitude vector length must be even");
	switch(v.normalization) {
		case NormalizationPolicy::AllowUnnormalized: return;
		case NormalizationPolicy::Normalize: {
			const double m = massOf(a);
			if(m <= 0.0) throw std::invalid_argument("Cannot normalize an all-zero amplitude vector");
			const double k = 1.0 / std::sqrt(m);
			for(double &x : a) x *= k;
			return;
		}
		case NormalizationPolicy::Validate:
			if(std::abs(massOf(a) - 1.0) > v.probabilityTolerance) throw std::invalid_argument("Amplitude vector is not normalized");
			return;
	}
}

// The states of a uniform subset as the dense backend keeps them.
std::vector<StateIndex> subsetMembers(unsigned int qubits, const BasisStateList &states) {
	const StateIndex N = checkedStateCount(qubits);
	if(states.empty()) throw std::invalid_argument("QuantumRegister::initUniformSuperposition requires at least one basis state");
	std::vector<StateIndex> m(states.begin(), states.end());
	for(StateIndex s : m) validateStateIndex("QuantumRegister::initUniformSuperposition", s, N);
	std::sort(m.begin(), m.end());
	m.erase(std::unique(m.begin(), m.end()), m.end());
	return m;
}

AmplitudesVector subsetVector(unsigned int qubits, const BasisStateList &states) {
	const std::vector<StateIndex> m = subsetMembers(qubits, states);
	AmplitudesVector a(checkedAmplitudeElementCount(qubits), 0.0);
	const double v = 1.0 / std::sqrt(static_cast<double>(m.size()));
	for(StateIndex s : m) a[static_cast<size_t>(2 * s)] = v;
	return a;
}

// A uniform random bit generator that hands out 32-bit words drawn earlier
// from std::mt19937, so that uniform_real_distribution maps them exactly as
// it maps the engine's own output.
struct WordReplay {
	using result_type = std::mt19937::result_type;
	const result_type *next;
	static constexpr result_type min() { return std::mt19937::min(); }
	static constexpr result_type max() { return std::mt19937::max(); }
	result_type operator()() { return *next++; }
};

double phaseRange() { return 2.0 * std::acos(-1.0); }

// randomPhase amplitudes on the host.  The engine's words are drawn in order,
// a block at a time; the distribution and one sincos per state (the call the
// reference's compiler emits for its cos/sin pair) run on worker threads
// over a replay of the block.
AmplitudesVector phaseVector(unsigned int qubits, unsigned int seed) {
	const StateIndex N = checkedStateCount(qubits);
	const double norm = 1.0 / std::sqrt(static_cast<double>(N));
	AmplitudesVector a(checkedAmplitudeElementCount(qubits), 0.0);
	std::mt19937 gen(seed);
	const StateIndex block = std::min<StateIndex>(N, StateIndex{1} << 22);
	std::vector<std::mt19937::result_type> words;
	const unsigned hw = std::thread::hardware_concurrency();
	const unsigned workers = std::max(1u, std::min(hw ? hw : 1u, 16u));
	for(StateIndex first = 0; first < N; first += block) {
		const StateIndex count = std::min<StateIndex>(block, N - first);
		words.resize(static_cast<size_t>(2 * count));  // generate_canonical<double, 53>: two words per value
		for(auto &w : words) w = gen();
		auto fill = [&](StateIndex lo, StateIndex hi) {
			std::uniform_real_distribution<double> phase(0.0, phaseRange());
			WordReplay src{words.data() + 2 * lo};
			for(StateIndex i = lo; i < hi; ++i) {
				double s, c;
				__builtin_sincos(phase(src), &s, &c);
				a[static_cast<size_t>(2 * (first + i))] = norm * c;
				a[static_cast<size_t>(2 * (first + i) + 1)] = norm * s;
			}
		};
		if(workers == 1 || count < 65536) {
			fill(0, count);
			continue;
		}
		std::vector<std::thread> pool;
		const StateIndex per = (count + workers - 1) / workers;
		for(StateIndex lo = 0; lo < count; lo += per) pool.emplace_back(fill, lo, std::min<StateIndex>(count, lo + per));
		for(auto &t : pool) t.join();
	}
	return a;
}

// One randomPhase amplitude without building the vector.
Amplitude phaseAmplitude(unsigned int qubits, unsigned int seed, StateIndex s) {
	std::mt19937 gen(seed);
	gen.discard(2 * s);
	std::uniform_real_distribution<double> phase(0.0, phaseRange());
	double sn, cs;
	__builtin_sincos(phase(gen), &sn, &cs);
	const double norm = 1.0 / std::sqrt(static_cast<double>(checkedStateCount(qubits)));
	return {norm * cs, norm * sn};
}

AmplitudesVector expandRecipe(unsigned int qubits, const Recipe &r) {
	return std::visit(Visitor{
	                      [&](const BasisRecipe &b) {
		                      AmplitudesVector a(checkedAmplitudeElementCount(qubits), 0.0);
		                      a[static_cast<size_t>(2 * b.state)] = b.amp.real;
		                      a[static_cast<size_t>(2 * b.state + 1)] = b.amp.imag;
		                      return a;
	                      },
	                      [&](const SubsetRecipe &u) { return subsetVector(qubits, u.states); },
	                      [&](const VectorRecipe &v) { return v.amps; },
	                      [&](const PhaseRecipe &p) { return phaseVector(qubits, p.seed); },
	                      [&](const ResultRecipe &g) { return g.reg->amplitudes(); },
	                  },
	                  r);
}

// Sequential cumulative scan of a host vector (the measure of a stored vector).
StateIndex sampleVector(const AmplitudesVector &a, double rnd, const char *notReached) {
	double cum = 0.0;
	for(size_t s = 0; 2 * s + 1 < a.size(); ++s) {
		cum += a[2 * s] * a[2 * s] + a[2 * s + 1] * a[2 * s + 1];
		if(rnd <= cum) return static_cast<StateIndex>(s);
	}
	throw std::runtime_error(notReached);
}

tmfqs_hip_state *deviceHandle(const QuantumRegister &reg) { return hipBackendHandle(reg.backend()); }

}  // namespace

struct QuantumState::Impl {
	unsigned int qubits = 0;
	Recipe recipe = BasisRecipe{};
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

const QuantumRegister *QuantumState::registerPtr() const {
	const auto *g = std::get_if<ResultRecipe>(&impl_->recipe);
	return g ? g->reg.get() : nullptr;
}

Amplitude QuantumState::amplitude(StateIndex s) const {
	if(s >= stateCount()) throw std::invalid_argument("Amplitude state index is out of range");
	const unsigned int q = impl_->qubits;
	return std::visit(Visitor{
	                      [&](const BasisRecipe &b) { return s == b.state ? b.amp : Amplitude{0.0, 0.0}; },
	                      [&](const SubsetRecipe &u) {
		                      const std::vector<StateIndex> m = subsetMembers(q, u.states);
		                      return std::binary_search(m.begin(), m.end(), s)
		                                 ? Amplitude{1.0 / std::sqrt(static_cast<double>(m.size())), 0.0}
		                                 : Amplitude{0.0, 0.0};
	                      },
	                      [&](const VectorRecipe &v) {
		                      return Amplitude{v.amps[static_cast<size_t>(2 * s)], v.amps[static_cast<size_t>(2 * s + 1)]};
	                      },
	                      [&](const PhaseRecipe &p) { return phaseAmplitude(q, p.seed, s); },
	                      [&](const ResultRecipe &g) { return g.reg->amplitude(s); },
	                  },
	                  impl_->recipe);
}

double QuantumState::totalProbability() const {
	if(const auto *b = std::get_if<BasisRecipe>(&impl_->recipe)) return b->amp.real * b->amp.real + b->amp.imag * b->amp.imag;
	if(const auto *g = std::get_if<ResultRecipe>(&impl_->recipe)) return g->reg->totalProbability();
	if(const auto *v = std::get_if<VectorRecipe>(&impl_->recipe)) return massOf(v->amps);
	return massOf(expandRecipe(impl_->qubits, impl_->recipe));
}

AmplitudesVector QuantumState::amplitudes() const { return expandRecipe(impl_->qubits, impl_->recipe); }

StateIndex QuantumState::measure(IRandomSource &randomSource) const {
	if(const auto *b = std::get_if<BasisRecipe>(&impl_->recipe)) return b->state;
	if(const auto *g = std::get_if<ResultRecipe>(&impl_->recipe)) return g->reg->measure(randomSource);
	if(std::holds_alternative<SubsetRecipe>(impl_->recipe)) {
		// materialized as a dense register, whose backend checks the draw
		const AmplitudesVector a = expandRecipe(impl_->qubits, impl_->recipe);
		const double rnd = randomSource.nextUnitDouble();
		if(rnd < 0.0 || rnd >= 1.0) throw std::invalid_argument("DenseStateBackend::sampleMeasurement requires rnd in [0,1)");
		return sampleVector(a, rnd, "DenseStateBackend::sampleMeasurement cumulative probability did not reach sample");
	}
	const double rnd = randomSource.nextUnitDouble();
	const auto *v = std::get_if<VectorRecipe>(&impl_->recipe);
	const char *msg = "QuantumState::measure cumulative probability did not reach sample";
	return v ? sampleVector(v->amps, rnd, msg) : sampleVector(expandRecipe(impl_->qubits, impl_->recipe), rnd, msg);
}

void QuantumState::setAmplitude(StateIndex s, Amplitude amplitude, const StateValidation &validation) {
	AmplitudesVector values = amplitudes();
	if(s >= stateCount()) throw std::invalid_argument("Amplitude state index is out of range");
	values[static_cast<size_t>(2 * s)] = amplitude.real;
	values[static_cast<size_t>(2 * s + 1)] = amplitude.imag;
	loadAmplitudes(std::move(values), validation);
}

void QuantumState::loadAmplitudes(AmplitudesVector amplitudes, const StateValidation &validation) {
	if(amplitudes.size() != checkedAmplitudeElementCount(impl_->qubits))
		throw std::invalid_argument("Amplitude vector size does not match state qubit count");
	applyValidation(amplitudes, validation);
	impl_->recipe = VectorRecipe{std::move(amplitudes)};
}

QuantumState QuantumState::basis(unsigned int qubits, StateIndex s) { return basisAmplitude(qubits, s, {1.0, 0.0}); }

QuantumState QuantumState::basisAmplitude(unsigned int qubits, StateIndex s, Amplitude amplitude) {
	if(s >= checkedStateCount(qubits)) throw std::invalid_argument("Basis state is out of range");
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->recipe = BasisRecipe{s, amplitude};
	return QuantumState(std::move(impl));
}

QuantumState QuantumState::uniformSubset(unsigned int qubits, BasisStateList states) {
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->recipe = SubsetRecipe{std::move(states)};
	return QuantumState(std::move(impl));
}

QuantumState QuantumState::fromAmplitudes(unsigned int qubits, AmplitudesVector amplitudes, const StateValidation &validation) {
	if(amplitudes.size() != checkedAmplitudeElementCount(qubits))
		throw std::invalid_argument("Amplitude vector size does not match qubit count");
	applyValidation(amplitudes, validation);
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->recipe = VectorRecipe{std::move(amplitudes)};
	return QuantumState(std::move(impl));
}

// The phases are drawn where the state is built (materialize, or a host
// query).  Its probability mass is checked there against the default 1e-9
// tolerance, as fromAmplitudes checks the vector the reference builds.
QuantumState QuantumState::randomPhase(unsigned int qubits, unsigned int seed) {
	(void)checkedStateCount(qubits);
	auto impl = std::make_unique<Impl>();
	impl->qubits = qubits;
	impl->recipe = PhaseRecipe{seed};
	return QuantumState(std::move(impl));
}

QuantumState QuantumState::sparsePattern(unsigned int qubits) {
	const StateIndex N = checkedStateCount(qubits);
	BasisStateList members;
	for(StateIndex x = 0; 8u * x + (x % 2u) < N; ++x) members.push_back(8u * x + (x % 2u));
	return uniformSubset(qubits, std::move(members));
}

QuantumRegister QuantumState::materialize(const RegisterConfig &config) const {
	const unsigned int q = impl_->qubits;
	return std::visit(
	    Visitor{
	        [&](const BasisRecipe &b) { return QuantumRegister(q, b.state, b.amp, config); },
	        [&](const SubsetRecipe &u) {
		        QuantumRegister reg(q, config);
		        reg.initUniformSuperposition(u.states);
		        return reg;
	        },
	        [&](const VectorRecipe &v) {
		        QuantumRegister reg(q, config);
		        if(tmfqs_hip_state *h = deviceHandle(reg))
			        raiseHipStatus(tmfqs_hip_load(h, q, v.amps.data(), v.amps.size()));  // straight from the recipe's vector
		        else
			        reg.loadAmplitudes(v.amps);
		        return reg;
	        },
	        [&](const PhaseRecipe &p) {
		        QuantumRegister reg(q, config);
		        if(tmfqs_hip_state *h = deviceHandle(reg)) {
			        // drawn, validated and stored in device memory; no host vector
			        raiseHipStatus(tmfqs_hip_init_random_phase(h, q, p.seed));
			        return reg;
		        }
		        AmplitudesVector a = phaseVector(q, p.seed);
		        applyValidation(a, StateValidation{});
		        reg.loadAmplitudes(std::move(a));
		        return reg;
	        },
	        [&](const ResultRecipe &g) { return QuantumRegister(*g.reg); },
	    },
	    impl_->recipe);
}

QuantumState QuantumState::fromRegister(QuantumRegister reg) {
	auto impl = std::make_unique<Impl>();
	impl->qubits = reg.qubitCount();
	impl->recipe = ResultRecipe{std::make_shared<const QuantumRegister>(std::move(reg))};
	return QuantumState(std::move(impl));
}

}  // namespace state
}  // namespace tmfqs
