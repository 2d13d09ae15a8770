// Register facade over a storage backend, with the reference's semantics
// (src/tmfqs/register/quantum_register.cpp): validation order, swap as three
// controlled-NOTs, the cached logical amplitude sum with its algebraic
// recurrence (:176-189, :405-420, :433-467) and the lazy affine overlay
// (:152-231).  Where the backend offers the optional accelerated reads
// (amplitudeSum, totalProbabilityAffine, sampleMeasurementAffine, applySwap)
// they replace the host loops that would stream the state to the host.
#include "tmfqs/register/quantum_register.h"

#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tmfqs/core/random.h"
#include "tmfqs/core/state_space.h"
#include "tmfqs/core/validation.h"
#include "tmfqs/storage/factory/state_backend_factory.h"

namespace tmfqs {
namespace {

constexpr double kDefaultPrintEpsilon = 1e-12;

Amplitude addA(Amplitude a, Amplitude b) { return {a.real + b.real, a.imag + b.imag}; }
Amplitude subA(Amplitude a, Amplitude b) { return {a.real - b.real, a.imag - b.imag}; }
Amplitude negA(Amplitude a) { return {-a.real, -a.imag}; }
Amplitude scaleA(Amplitude a, double s) { return {a.real * s, a.imag * s}; }
Amplitude mulA(Amplitude a, Amplitude b) { return {a.real * b.real - a.imag * b.imag, a.real * b.imag + a.imag * b.real}; }
Amplitude divA(Amplitude n, Amplitude d) {
	const double dd = d.real * d.real + d.imag * d.imag;
	if(dd == 0.0) throw std::logic_error("QuantumRegister affine overlay scale is zero");
	return {(n.real * d.real + n.imag * d.imag) / dd, (n.imag * d.real - n.real * d.imag) / dd};
}
double probOf(Amplitude a) { return a.real * a.real + a.imag * a.imag; }

}  // namespace

QuantumRegister::QuantumRegister() {
	initializeBackend(0);
	backend_->initBasis(numQubits_, 0, {1.0, 0.0});
}

void QuantumRegister::initializeBackend(unsigned int qubits) {
	numQubits_ = qubits;
	numStates_ = checkedStateCount(numQubits_);
	BackendSelection sel = StateBackendFactory::createSelection(numQubits_, config_);
	config_ = sel.config;
	resolvedStrategy_ = sel.strategy;
	backend_ = std::move(sel.backend);
}

QuantumRegister::QuantumRegister(unsigned int qubits, const RegisterConfig &cfg) {
	config_ = cfg;
	initializeBackend(qubits);
	backend_->initBasis(numQubits_, 0, {1.0, 0.0});
}

QuantumRegister::QuantumRegister(unsigned int qubits, StateIndex initState, const RegisterConfig &cfg) {
	config_ = cfg;
	initializeBackend(qubits);
	backend_->initBasis(numQubits_, initState, {1.0, 0.0});
}

QuantumRegister::QuantumRegister(unsigned int qubits, StateIndex initState, Amplitude amp, const RegisterConfig &cfg) {
	config_ = cfg;
	initializeBackend(qubits);
	backend_->initBasis(numQubits_, initState, amp);
}

QuantumRegister::QuantumRegister(unsigned int qubits, AmplitudesVector amplitudes, const RegisterConfig &cfg) {
	config_ = cfg;
	initializeBackend(qubits);
	backend_->loadAmplitudes(numQubits_, std::move(amplitudes));
}

QuantumRegister::QuantumRegister(const QuantumRegister &o)
    : numQubits_(o.numQubits_),
      numStates_(o.numStates_),
      config_(o.config_),
      resolvedStrategy_(o.resolvedStrategy_),
      cachedLogicalAmplitudeSum_(o.cachedLogicalAmplitudeSum_),
      affineScale_(o.affineScale_),
      affineBias_(o.affineBias_),
      affineOverlayActive_(o.affineOverlayActive_) {
	if(o.backend_) backend_ = o.backend_->clone();
}

QuantumRegister &QuantumRegister::operator=(const QuantumRegister &o) {
	if(this == &o) return *this;
	numQubits_ = o.numQubits_;
	numStates_ = o.numStates_;
	config_ = o.config_;
	resolvedStrategy_ = o.resolvedStrategy_;
	if(!o.backend_) {
		backend_.reset();
		resetAffineOverlay();
		return *this;
	}
	backend_ = o.backend_->clone();
	affineScale_ = o.affineScale_;
	affineBias_ = o.affineBias_;
	affineOverlayActive_ = o.affineOverlayActive_;
	cachedLogicalAmplitudeSum_ = o.cachedLogicalAmplitudeSum_;
	return *this;
}

QuantumRegister::~QuantumRegister() = default;

void QuantumRegister::requireInitialized(const char *operation) const {
	if(!backend_) throw std::logic_error(std::string("QuantumRegister is not initialized for ") + operation);
}

void QuantumRegister::resetAffineOverlay() {
	affineScale_ = {1.0, 0.0};
	affineBias_ = {0.0, 0.0};
	affineOverlayActive_ = false;
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

Amplitude QuantumRegister::applyAffineOverlay(Amplitude b) const {
	if(!affineOverlayActive_) return b;
	return addA(mulA(affineScale_, b), affineBias_);
}

Amplitude QuantumRegister::removeAffineOverlay(Amplitude l) const {
	if(!affineOverlayActive_) return l;
	return divA(subA(l, affineBias_), affineScale_);
}

Amplitude QuantumRegister::logicalAmplitudeSum() const {
	if(cachedLogicalAmplitudeSum_.has_value()) return *cachedLogicalAmplitudeSum_;
	Amplitude sum{0.0, 0.0};
	if(affineOverlayActive_ || !backend_->amplitudeSum(sum)) {
		sum = {0.0, 0.0};
		visitLogicalAmplitudeChunks([&](StateIndex, const double *d, size_t elems) {
			for(size_t e = 0; e + 1u < elems; e += 2u) {
				sum.real += d[e];
				sum.imag += d[e + 1u];
			}
		});
	}
	cachedLogicalAmplitudeSum_ = sum;
	return sum;
}

void QuantumRegister::visitLogicalAmplitudeChunks(const std::function<void(StateIndex, const double *, size_t)> &visitor) const {
	requireInitialized("amplitude chunk iteration");
	if(!affineOverlayActive_) {
		backend_->forEachAmplitudeChunk(visitor);
		return;
	}
	std::vector<double> logical;
	backend_->forEachAmplitudeChunk([&](StateIndex base, const double *d, size_t elems) {
		logical.resize(elems);
		for(size_t e = 0; e + 1u < elems; e += 2u) {
			const Amplitude l = applyAffineOverlay({d[e], d[e + 1u]});
			logical[e] = l.real;
			logical[e + 1u] = l.imag;
		}
		visitor(base, logical.data(), elems);
	});
}

AmplitudesVector QuantumRegister::snapshotLogicalAmplitudes() const {
	AmplitudesVector out;
	backend_->exportAmplitudes(out);
	if(affineOverlayActive_) {
		for(size_t e = 0; e + 1u < out.size(); e += 2u) {
			const Amplitude l = applyAffineOverlay({out[e], out[e + 1u]});
			out[e] = l.real;
			out[e + 1u] = l.imag;
		}
	}
	return out;
}

void QuantumRegister::flushAffineOverlay() {
	if(!affineOverlayActive_) return;
	const std::optional<Amplitude> saved = cachedLogicalAmplitudeSum_;
	backend_->applyAffineTransform(affineScale_, affineBias_);
	resetAffineOverlay();
	cachedLogicalAmplitudeSum_ = saved;
}

IStateBackend *QuantumRegister::beginExternalUpdate() {
	requireInitialized("external update");
	flushAffineOverlay();
	cachedLogicalAmplitudeSum_.reset();
	return backend_.get();
}

unsigned int QuantumRegister::qubitCount() const { return numQubits_; }
StateIndex QuantumRegister::stateCount() const { return numStates_; }
size_t QuantumRegister::amplitudeElementCount() const { return static_cast<size_t>(numStates_) * 2u; }

Amplitude QuantumRegister::amplitude(StateIndex s) const {
	requireInitialized("amplitude query");
	validateStateIndex("QuantumRegister::amplitude", s, numStates_);
	return applyAffineOverlay(backend_->amplitude(s));
}

AmplitudesVector QuantumRegister::amplitudes() const {
	requireInitialized("amplitude export");
	return snapshotLogicalAmplitudes();
}

double QuantumRegister::probability(StateIndex s) const {
	requireInitialized("probability query");
	validateStateIndex("QuantumRegister::probability", s, numStates_);
	return probOf(amplitude(s));
}

double QuantumRegister::totalProbability() const {
	requireInitialized("probability summation");
	if(!affineOverlayActive_) return backend_->totalProbability();
	double total = 0.0;
	if(backend_->totalProbabilityAffine(affineScale_, affineBias_, total)) return total;
	total = 0.0;
	visitLogicalAmplitudeChunks([&](StateIndex, const double *d, size_t elems) {
		for(size_t e = 0; e + 1u < elems; e += 2u) total += d[e] * d[e] + d[e + 1u] * d[e + 1u];
	});
	return total;
}

StateIndex QuantumRegister::measure(IRandomSource &src) const {
	requireInitialized("measurement");
	if(!affineOverlayActive_) return backend_->sampleMeasurement(src.nextUnitDouble());
	const double rnd = src.nextUnitDouble();
	StateIndex out = 0;
	if(backend_->sampleMeasurementAffine(rnd, affineScale_, affineBias_, out)) return out;
	double cum = 0.0;
	StateIndex found = numStates_;
	visitLogicalAmplitudeChunks([&](StateIndex base, const double *d, size_t elems) {
		if(found != numStates_) return;
		for(size_t i = 0; i < elems / 2u; ++i) {
			cum += d[2 * i] * d[2 * i] + d[2 * i + 1] * d[2 * i + 1];
			if(rnd <= cum) {
				found = base + i;
				return;
			}
		}
	});
	if(found != numStates_) return found;
	throw std::runtime_error("QuantumRegister::measure cumulative probability did not reach sample");
}

void QuantumRegister::setAmplitude(StateIndex s, Amplitude a) {
	requireInitialized("state update");
	validateStateIndex("QuantumRegister::setAmplitude", s, numStates_);
	backend_->setAmplitude(s, removeAffineOverlay(a));
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

void QuantumRegister::loadAmplitudes(AmplitudesVector a) {
	requireInitialized("amplitude loading");
	if(a.size() != amplitudeElementCount()) throw std::invalid_argument("QuantumRegister::loadAmplitudes amplitudes size mismatch");
	resetAffineOverlay();
	backend_->loadAmplitudes(numQubits_, std::move(a));
}

void QuantumRegister::initUniformSuperposition(const BasisStateList &states) {
	requireInitialized("uniform superposition initialization");
	if(states.empty()) throw std::invalid_argument("QuantumRegister::initUniformSuperposition requires at least one basis state");
	for(StateIndex s : states) validateStateIndex("QuantumRegister::initUniformSuperposition", s, numStates_);
	resetAffineOverlay();
	backend_->initUniformSuperposition(numQubits_, states);
}

StorageStrategyKind QuantumRegister::storageStrategy() const { return resolvedStrategy_; }

std::ostream &operator<<(std::ostream &os, const QuantumRegister &reg) {
	reg.requireInitialized("state printing");
	if(!reg.affineOverlayActive_) {
		reg.backend_->printNonZeroStates(os, kDefaultPrintEpsilon);
		return os;
	}
	reg.visitLogicalAmplitudeChunks([&](StateIndex base, const double *d, size_t elems) {
		for(size_t i = 0; i < elems / 2u; ++i)
			if(std::abs(d[2 * i]) > kDefaultPrintEpsilon || std::abs(d[2 * i + 1]) > kDefaultPrintEpsilon)
				os << (base + i) << ": " << d[2 * i] << " + " << d[2 * i + 1] << "i\n";
	});
	return os;
}

void QuantumRegister::printStatesVector(double epsilon) const {
	requireInitialized("state printing");
	if(epsilon < 0.0) throw std::invalid_argument("QuantumRegister::printStatesVector epsilon must be non-negative");
	if(!affineOverlayActive_) {
		backend_->printNonZeroStates(std::cout, epsilon);
		return;
	}
	visitLogicalAmplitudeChunks([&](StateIndex base, const double *d, size_t elems) {
		for(size_t i = 0; i < elems / 2u; ++i)
			if(std::abs(d[2 * i]) > epsilon || std::abs(d[2 * i + 1]) > epsilon)
				std::cout << (base + i) << ": " << d[2 * i] << " + " << d[2 * i + 1] << "i\n";
	});
}

void QuantumRegister::applyGate(const QuantumGate &gate, const QubitList &qubits) {
	requireInitialized("gate application");
	validateGateTargets("QuantumRegister::applyGate", qubits, numQubits_, gate.dimension());
	if(affineOverlayActive_) flushAffineOverlay();
	backend_->applyGate(gate, qubits);
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

void QuantumRegister::applyPhaseFlipBasisState(StateIndex s) {
	requireInitialized("basis-state phase flip");
	validateStateIndex("QuantumRegister::applyPhaseFlipBasisState", s, numStates_);
	if(cachedLogicalAmplitudeSum_.has_value() || affineOverlayActive_) {
		const Amplitude l = amplitude(s);
		if(cachedLogicalAmplitudeSum_.has_value()) {
			cachedLogicalAmplitudeSum_->real -= 2.0 * l.real;
			cachedLogicalAmplitudeSum_->imag -= 2.0 * l.imag;
		}
		if(affineOverlayActive_) {
			backend_->setAmplitude(s, removeAffineOverlay(negA(l)));
			return;
		}
	}
	backend_->phaseFlipBasisState(s);
}

void QuantumRegister::applyInversionAboutMean() {
	requireInitialized("inversion about mean");
	const Amplitude sum = logicalAmplitudeSum();
	applyInversionAboutMean({sum.real / static_cast<double>(numStates_), sum.imag / static_cast<double>(numStates_)});
}

void QuantumRegister::applyInversionAboutMean(Amplitude mean) {
	requireInitialized("inversion about mean");
	if(cachedLogicalAmplitudeSum_.has_value()) {
		cachedLogicalAmplitudeSum_->real = 2.0 * mean.real * static_cast<double>(numStates_) - cachedLogicalAmplitudeSum_->real;
		cachedLogicalAmplitudeSum_->imag = 2.0 * mean.imag * static_cast<double>(numStates_) - cachedLogicalAmplitudeSum_->imag;
	}
	const Amplitude scale = affineOverlayActive_ ? affineScale_ : Amplitude{1.0, 0.0};
	const Amplitude bias = affineOverlayActive_ ? affineBias_ : Amplitude{0.0, 0.0};
	affineScale_ = negA(scale);
	affineBias_ = subA(scaleA(mean, 2.0), bias);
	affineOverlayActive_ = true;
}

void QuantumRegister::applyInversionAboutMeanMaterialized() {
	requireInitialized("inversion about mean");
	const Amplitude sum = logicalAmplitudeSum();
	applyInversionAboutMeanMaterialized({sum.real / static_cast<double>(numStates_), sum.imag / static_cast<double>(numStates_)});
}

void QuantumRegister::applyInversionAboutMeanMaterialized(Amplitude mean) {
	requireInitialized("inversion about mean");
	if(cachedLogicalAmplitudeSum_.has_value()) {
		cachedLogicalAmplitudeSum_->real = 2.0 * mean.real * static_cast<double>(numStates_) - cachedLogicalAmplitudeSum_->real;
		cachedLogicalAmplitudeSum_->imag = 2.0 * mean.imag * static_cast<double>(numStates_) - cachedLogicalAmplitudeSum_->imag;
	}
	if(affineOverlayActive_) flushAffineOverlay();
	backend_->inversionAboutMean(mean);
}

void QuantumRegister::applyHadamard(QubitIndex q) {
	requireInitialized("Hadamard");
	validateQubitIndex("QuantumRegister::applyHadamard", q, numQubits_);
	if(affineOverlayActive_) flushAffineOverlay();
	backend_->applyHadamard(q);
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

void QuantumRegister::applyPauliX(QubitIndex q) {
	requireInitialized("PauliX");
	validateQubitIndex("QuantumRegister::applyPauliX", q, numQubits_);
	if(affineOverlayActive_) flushAffineOverlay();
	backend_->applyPauliX(q);
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

void QuantumRegister::applyControlledPhaseShift(QubitIndex c, QubitIndex t, double theta) {
	requireInitialized("controlled phase shift");
	validateQubitIndex("QuantumRegister::applyControlledPhaseShift", c, numQubits_);
	validateQubitIndex("QuantumRegister::applyControlledPhaseShift", t, numQubits_);
	validateDistinctQubits("QuantumRegister::applyControlledPhaseShift", c, t);
	if(affineOverlayActive_) flushAffineOverlay();
	backend_->applyControlledPhaseShift(c, t, theta);
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

void QuantumRegister::applyControlledNot(QubitIndex c, QubitIndex t) {
	requireInitialized("controlled not");
	validateQubitIndex("QuantumRegister::applyControlledNot", c, numQubits_);
	validateQubitIndex("QuantumRegister::applyControlledNot", t, numQubits_);
	validateDistinctQubits("QuantumRegister::applyControlledNot", c, t);
	if(affineOverlayActive_) flushAffineOverlay();
	backend_->applyControlledNot(c, t);
	cachedLogicalAmplitudeSum_ = std::nullopt;
}

// reference :518-527: three controlled-NOTs; a backend that supports the
// equivalent single permutation runs that instead.
void QuantumRegister::applySwap(QubitIndex a, QubitIndex b) {
	requireInitialized("swap");
	validateQubitIndex("QuantumRegister::applySwap", a, numQubits_);
	validateQubitIndex("QuantumRegister::applySwap", b, numQubits_);
	validateDistinctQubits("QuantumRegister::applySwap", a, b);
	if(affineOverlayActive_) flushAffineOverlay();
	if(backend_->applySwap(a, b)) {
		cachedLogicalAmplitudeSum_ = std::nullopt;
		return;
	}
	applyControlledNot(a, b);
	applyControlledNot(b, a);
	applyControlledNot(a, b);
}

void QuantumRegister::beginOperationBatch() {
	requireInitialized("operation batch begin");
	backend_->beginOperationBatch();
}

void QuantumRegister::endOperationBatch() {
	requireInitialized("operation batch end");
	backend_->endOperationBatch();
}

}  // namespace tmfqs
