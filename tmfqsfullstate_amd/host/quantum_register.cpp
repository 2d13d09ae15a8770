// QuantumRegister on top of an IStateBackend.
//
// The register adds three things to the backend's storage, each with the
// reference's semantics (src/tmfqs/register/quantum_register.cpp:61-539):
//   * argument validation, with its error classes and scopes;
//   * an optional affine map between stored and logical amplitudes (the lazy
//     diffusion, :152-231, :423-444), collapsed into the storage before any
//     gate;
//   * the exact sequential sum of the logical amplitudes once it has been
//     computed, kept up to date through phase flips and reflections about
//     the mean (:176-189, :405-420, :447-467), so that Grover's mean costs
//     no sweep after the first.
// Every arithmetic step on amplitudes and sums is written in the reference's
// operation order, which keeps results bit-identical.  Where the backend
// offers accelerated reads (amplitudeSum, totalProbabilityAffine,
// sampleMeasurementAffine) or a one-pass swap, they replace host scans.
#include "tmfqs/register/quantum_register.h"

#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>

#include "tmfqs/core/random.h"
#include "tmfqs/core/state_space.h"
#include "tmfqs/core/validation.h"
#include "tmfqs/storage/factory/state_backend_factory.h"

namespace tmfqs {
namespace {

// Cutoff of operator<< (the reference prints amplitudes above 1e-12).
constexpr double kStreamPrintCutoff = 1e-12;

void requireDistinctPair(const char *scope, QubitIndex a, QubitIndex b, unsigned int qubits) {
	validateQubitIndex(scope, a, qubits);
	validateQubitIndex(scope, b, qubits);
	validateDistinctQubits(scope, a, b);
}

void printAmplitude(std::ostream &os, StateIndex s, Amplitude a) { os << s << ": " << a.real << " + " << a.imag << "i\n"; }

}  // namespace

// ------------------------------------------------------------ affine map
Amplitude QuantumRegister::AffineMap::toLogical(Amplitude x) const {
	const double re = scale.real * x.real - scale.imag * x.imag;
	const double im = scale.real * x.imag + scale.imag * x.real;
	return {re + bias.real, im + bias.imag};
}

Amplitude QuantumRegister::AffineMap::toStored(Amplitude l) const {
	const double norm = scale.real * scale.real + scale.imag * scale.imag;
	if(norm == 0.0) throw std::logic_error("QuantumRegister affine overlay scale is zero");
	const double dr = l.real - bias.real, di = l.imag - bias.imag;
	return {(dr * scale.real + di * scale.imag) / norm, (di * scale.real - dr * scale.imag) / norm};
}

// Reflection about the mean composed after this map: 2m - (s x + b) = (-s) x + (2m - b).
QuantumRegister::AffineMap QuantumRegister::AffineMap::reflectedAbout(Amplitude mean) const {
	AffineMap r;
	r.scale = {-scale.real, -scale.imag};
	r.bias = {mean.real * 2.0 - bias.real, mean.imag * 2.0 - bias.imag};
	return r;
}

// ------------------------------------------------------------ construction
QuantumRegister::QuantumRegister(unsigned int qubits, const RegisterConfig &cfg,
                                 const std::function<void(IStateBackend &, unsigned int)> &init)
    : numQubits_(qubits), numStates_(checkedStateCount(qubits)), config_(cfg) {
	BackendSelection sel = StateBackendFactory::createSelection(numQubits_, config_);
	config_ = sel.config;
	resolvedStrategy_ = sel.strategy;
	backend_ = std::move(sel.backend);
	init(*backend_, numQubits_);
}

QuantumRegister::QuantumRegister() : QuantumRegister(0u, RegisterConfig{}) {}

QuantumRegister::QuantumRegister(unsigned int qubits, const RegisterConfig &cfg)
    : QuantumRegister(qubits, StateIndex{0}, Amplitude{1.0, 0.0}, cfg) {}

QuantumRegister::QuantumRegister(unsigned int qubits, StateIndex initState, const RegisterConfig &cfg)
    : QuantumRegister(qubits, initState, Amplitude{1.0, 0.0}, cfg) {}

QuantumRegister::QuantumRegister(unsigned int qubits, StateIndex initState, Amplitude amp, const RegisterConfig &cfg)
    : QuantumRegister(qubits, cfg, [&](IStateBackend &b, unsigned int n) { b.initBasis(n, initState, amp); }) {}

QuantumRegister::QuantumRegister(unsigned int qubits, AmplitudesVector amplitudes, const RegisterConfig &cfg)
    : QuantumRegister(qubits, cfg, [&](IStateBackend &b, unsigned int n) { b.loadAmplitudes(n, std::move(amplitudes)); }) {}

QuantumRegister::QuantumRegister(const QuantumRegister &o)
    : numQubits_(o.numQubits_),
      numStates_(o.numStates_),
      config_(o.config_),
      resolvedStrategy_(o.resolvedStrategy_),
      backend_(o.backend_ ? o.backend_->clone() : nullptr),
      overlay_(o.backend_ ? o.overlay_ : std::nullopt),
      logicalSum_(o.backend_ ? o.logicalSum_ : std::nullopt) {}

QuantumRegister &QuantumRegister::operator=(const QuantumRegister &o) {
	if(this != &o) {
		QuantumRegister copy(o);
		*this = std::move(copy);
	}
	return *this;
}

QuantumRegister::~QuantumRegister() = default;

// ------------------------------------------------------------ internals
IStateBackend &QuantumRegister::backendFor(const char *operation) const {
	if(!backend_) throw std::logic_error(std::string("QuantumRegister is not initialized for ") + operation);
	return *backend_;
}

// Writes the pending affine map into the storage; the logical state, and so
// its known sum, do not change.
void QuantumRegister::collapseOverlay() {
	if(!overlay_) return;
	backend_->applyAffineTransform(overlay_->scale, overlay_->bias);
	overlay_.reset();
}

void QuantumRegister::gateUpdate(const char *operation, const std::function<void()> &validate,
                                 const std::function<void(IStateBackend &)> &apply) {
	IStateBackend &b = backendFor(operation);
	validate();
	collapseOverlay();
	apply(b);
	logicalSum_.reset();
}

// Streams the logical amplitudes in index order until visit returns false;
// returns whether it stopped early.
bool QuantumRegister::scanLogical(const std::function<bool(StateIndex, Amplitude)> &visit) const {
	IStateBackend &b = backendFor("amplitude chunk iteration");
	bool stopped = false;
	b.forEachAmplitudeChunk([&](StateIndex base, const double *d, size_t elems) {
		for(size_t e = 0; !stopped && e + 1u < elems; e += 2u) {
			Amplitude a{d[e], d[e + 1u]};
			if(overlay_) a = overlay_->toLogical(a);
			stopped = !visit(base + static_cast<StateIndex>(e / 2u), a);
		}
	});
	return stopped;
}

Amplitude QuantumRegister::logicalSum() const {
	if(logicalSum_) return *logicalSum_;
	Amplitude sum{0.0, 0.0};
	if(overlay_ || !backendFor("amplitude chunk iteration").amplitudeSum(sum)) {
		sum = {0.0, 0.0};
		scanLogical([&](StateIndex, Amplitude a) {
			sum.real += a.real;
			sum.imag += a.imag;
			return true;
		});
	}
	logicalSum_ = sum;
	return sum;
}

// The known sum after a reflection about `mean`: sum' = 2 m N - sum.
void QuantumRegister::reflectSum(Amplitude mean) const {
	if(!logicalSum_) return;
	const double N = static_cast<double>(numStates_);
	logicalSum_->real = 2.0 * mean.real * N - logicalSum_->real;
	logicalSum_->imag = 2.0 * mean.imag * N - logicalSum_->imag;
}

void QuantumRegister::printLogical(std::ostream &os, double epsilon) const {
	if(!overlay_) {
		backend_->printNonZeroStates(os, epsilon);
		return;
	}
	scanLogical([&](StateIndex s, Amplitude a) {
		if(std::abs(a.real) > epsilon || std::abs(a.imag) > epsilon) printAmplitude(os, s, a);
		return true;
	});
}

IStateBackend *QuantumRegister::beginExternalUpdate() {
	IStateBackend &b = backendFor("external update");
	collapseOverlay();
	logicalSum_.reset();
	return &b;
}

// ------------------------------------------------------------ queries
unsigned int QuantumRegister::qubitCount() const { return numQubits_; }
StateIndex QuantumRegister::stateCount() const { return numStates_; }
size_t QuantumRegister::amplitudeElementCount() const { return static_cast<size_t>(numStates_) * 2u; }
StorageStrategyKind QuantumRegister::storageStrategy() const { return resolvedStrategy_; }

Amplitude QuantumRegister::amplitude(StateIndex state) const {
	IStateBackend &b = backendFor("amplitude query");
	validateStateIndex("QuantumRegister::amplitude", state, numStates_);
	const Amplitude stored = b.amplitude(state);
	return overlay_ ? overlay_->toLogical(stored) : stored;
}

double QuantumRegister::probability(StateIndex state) const {
	backendFor("probability query");
	validateStateIndex("QuantumRegister::probability", state, numStates_);
	const Amplitude a = amplitude(state);
	return a.real * a.real + a.imag * a.imag;
}

AmplitudesVector QuantumRegister::amplitudes() const {
	IStateBackend &b = backendFor("amplitude export");
	AmplitudesVector out;
	b.exportAmplitudes(out);
	if(overlay_)
		for(size_t e = 0; e + 1u < out.size(); e += 2u) {
			const Amplitude l = overlay_->toLogical({out[e], out[e + 1u]});
			out[e] = l.real;
			out[e + 1u] = l.imag;
		}
	return out;
}

double QuantumRegister::totalProbability() const {
	IStateBackend &b = backendFor("probability summation");
	if(!overlay_) return b.totalProbability();
	double total = 0.0;
	if(b.totalProbabilityAffine(overlay_->scale, overlay_->bias, total)) return total;
	total = 0.0;
	scanLogical([&](StateIndex, Amplitude a) {
		total += a.real * a.real + a.imag * a.imag;
		return true;
	});
	return total;
}

StateIndex QuantumRegister::measure(IRandomSource &randomSource) const {
	IStateBackend &b = backendFor("measurement");
	if(!overlay_) return b.sampleMeasurement(randomSource.nextUnitDouble());
	const double rnd = randomSource.nextUnitDouble();
	StateIndex hit = 0;
	if(b.sampleMeasurementAffine(rnd, overlay_->scale, overlay_->bias, hit)) return hit;
	double cumulative = 0.0;
	const bool found = scanLogical([&](StateIndex s, Amplitude a) {
		cumulative += a.real * a.real + a.imag * a.imag;
		if(rnd <= cumulative) {
			hit = s;
			return false;
		}
		return true;
	});
	if(!found) throw std::runtime_error("QuantumRegister::measure cumulative probability did not reach sample");
	return hit;
}

void QuantumRegister::printStatesVector(double epsilon) const {
	backendFor("state printing");
	if(epsilon < 0.0) throw std::invalid_argument("QuantumRegister::printStatesVector epsilon must be non-negative");
	printLogical(std::cout, epsilon);
}

std::ostream &operator<<(std::ostream &os, const QuantumRegister &reg) {
	reg.backendFor("state printing");
	reg.printLogical(os, kStreamPrintCutoff);
	return os;
}

// ------------------------------------------------------------ state updates
void QuantumRegister::setAmplitude(StateIndex state, Amplitude amp) {
	IStateBackend &b = backendFor("state update");
	validateStateIndex("QuantumRegister::setAmplitude", state, numStates_);
	b.setAmplitude(state, overlay_ ? overlay_->toStored(amp) : amp);
	logicalSum_.reset();
}

void QuantumRegister::loadAmplitudes(AmplitudesVector amplitudes) {
	IStateBackend &b = backendFor("amplitude loading");
	if(amplitudes.size() != amplitudeElementCount())
		throw std::invalid_argument("QuantumRegister::loadAmplitudes amplitudes size mismatch");
	overlay_.reset();
	logicalSum_.reset();
	b.loadAmplitudes(numQubits_, std::move(amplitudes));
}

void QuantumRegister::initUniformSuperposition(const BasisStateList &basisStates) {
	IStateBackend &b = backendFor("uniform superposition initialization");
	if(basisStates.empty())
		throw std::invalid_argument("QuantumRegister::initUniformSuperposition requires at least one basis state");
	for(StateIndex s : basisStates) validateStateIndex("QuantumRegister::initUniformSuperposition", s, numStates_);
	overlay_.reset();
	logicalSum_.reset();
	b.initUniformSuperposition(numQubits_, basisStates);
}

// ------------------------------------------------------------ Grover steps
void QuantumRegister::applyPhaseFlipBasisState(StateIndex state) {
	IStateBackend &b = backendFor("basis-state phase flip");
	validateStateIndex("QuantumRegister::applyPhaseFlipBasisState", state, numStates_);
	if(!logicalSum_ && !overlay_) {
		b.phaseFlipBasisState(state);
		return;
	}
	const Amplitude before = amplitude(state);
	if(logicalSum_) {
		logicalSum_->real -= 2.0 * before.real;
		logicalSum_->imag -= 2.0 * before.imag;
	}
	if(overlay_)
		b.setAmplitude(state, overlay_->toStored({-before.real, -before.imag}));
	else
		b.phaseFlipBasisState(state);
}

void QuantumRegister::applyInversionAboutMean() {
	backendFor("inversion about mean");
	const Amplitude s = logicalSum();
	const double N = static_cast<double>(numStates_);
	applyInversionAboutMean({s.real / N, s.imag / N});
}

void QuantumRegister::applyInversionAboutMean(Amplitude mean) {
	backendFor("inversion about mean");
	reflectSum(mean);
	overlay_ = overlay_.value_or(AffineMap{}).reflectedAbout(mean);
}

void QuantumRegister::applyInversionAboutMeanMaterialized() {
	backendFor("inversion about mean");
	const Amplitude s = logicalSum();
	const double N = static_cast<double>(numStates_);
	applyInversionAboutMeanMaterialized({s.real / N, s.imag / N});
}

void QuantumRegister::applyInversionAboutMeanMaterialized(Amplitude mean) {
	IStateBackend &b = backendFor("inversion about mean");
	reflectSum(mean);
	collapseOverlay();
	b.inversionAboutMean(mean);
}

// ------------------------------------------------------------ gates
void QuantumRegister::applyGate(const QuantumGate &gate, const QubitList &qubits) {
	gateUpdate(
	    "gate application", [&] { validateGateTargets("QuantumRegister::applyGate", qubits, numQubits_, gate.dimension()); },
	    [&](IStateBackend &b) { b.applyGate(gate, qubits); });
}

void QuantumRegister::applyHadamard(QubitIndex qubit) {
	gateUpdate(
	    "Hadamard", [&] { validateQubitIndex("QuantumRegister::applyHadamard", qubit, numQubits_); },
	    [&](IStateBackend &b) { b.applyHadamard(qubit); });
}

void QuantumRegister::applyPauliX(QubitIndex qubit) {
	gateUpdate(
	    "PauliX", [&] { validateQubitIndex("QuantumRegister::applyPauliX", qubit, numQubits_); },
	    [&](IStateBackend &b) { b.applyPauliX(qubit); });
}

void QuantumRegister::applyControlledPhaseShift(QubitIndex controlQubit, QubitIndex targetQubit, double theta) {
	gateUpdate(
	    "controlled phase shift",
	    [&] { requireDistinctPair("QuantumRegister::applyControlledPhaseShift", controlQubit, targetQubit, numQubits_); },
	    [&](IStateBackend &b) { b.applyControlledPhaseShift(controlQubit, targetQubit, theta); });
}

void QuantumRegister::applyControlledNot(QubitIndex controlQubit, QubitIndex targetQubit) {
	gateUpdate(
	    "controlled not", [&] { requireDistinctPair("QuantumRegister::applyControlledNot", controlQubit, targetQubit, numQubits_); },
	    [&](IStateBackend &b) { b.applyControlledNot(controlQubit, targetQubit); });
}

// A swap is three controlled-NOTs (reference :518-527); a backend with a
// one-pass swap (a bit permutation, bit-identical) runs that instead.
void QuantumRegister::applySwap(QubitIndex qubit1, QubitIndex qubit2) {
	gateUpdate(
	    "swap", [&] { requireDistinctPair("QuantumRegister::applySwap", qubit1, qubit2, numQubits_); },
	    [&](IStateBackend &b) {
		    if(b.applySwap(qubit1, qubit2)) return;
		    b.applyControlledNot(qubit1, qubit2);
		    b.applyControlledNot(qubit2, qubit1);
		    b.applyControlledNot(qubit1, qubit2);
	    });
}

void QuantumRegister::beginOperationBatch() { backendFor("operation batch begin").beginOperationBatch(); }

void QuantumRegister::endOperationBatch() { backendFor("operation batch end").endOperationBatch(); }

}  // namespace tmfqs
