// Circuit-level behaviour on the Hip strategy: the checks of the reference's
// tests/integration/circuit/test_circuit.cpp (known answers, seeded Grover
// outcomes, error cases, operation sequences), run through the public C++ API.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <vector>

#include "test_helpers.h"

using namespace tmfqs;

namespace {

StateIndex grover(unsigned n, BasisStateList marked, IRandomSource &rng, bool materialized = false, RegisterConfig cfg = hipcheck::deviceConfig()) {
	circuit::GroverCircuitOptions o;
	o.markedStates = std::move(marked);
	o.materializedDiffusion = materialized;
	sim::ExecutionConfig ex;
	ex.backend = cfg;
	return sim::Simulator(ex).run(circuit::makeGrover(n, o), state::QuantumState::basis(n)).state.measure(rng);
}

state::QuantumState qft(unsigned n, StateIndex input) {
	sim::ExecutionConfig ex;
	ex.backend = hipcheck::deviceConfig();
	return sim::Simulator(ex).run(circuit::makeQft(n), state::QuantumState::basis(n, input)).state;
}

bool in(const std::vector<StateIndex> &v, StateIndex x) { return std::find(v.begin(), v.end(), x) != v.end(); }

void complexExpValues() {
	const Amplitude a = complexExp({0.0, 0.0});
	CHECK(hipcheck::near(a.real, 1.0) && hipcheck::near(a.imag, 0.0));
	const Amplitude b = complexExp({0.0, kPi});
	CHECK(hipcheck::near(b.real, -1.0) && hipcheck::near(b.imag, 0.0));
	CHECK(hipcheck::near(complexExp({1.0, 0.0}).real, std::exp(1.0)));
}

void groverFindsMark() {
	int hits = 0;
	for(uint32_t s = 100; s < 120; ++s) {
		Mt19937RandomSource rng(s);
		hits += grover(5, {3}, rng) == 3 ? 1 : 0;
	}
	CHECK(hits > 0);
	for(uint32_t s = 200; s < 232; ++s) {
		Mt19937RandomSource rng(s);
		CHECK(grover(2, {2}, rng) == 2);
	}
}

void gateScalar() {
	const QuantumGate g = QuantumGate::Hadamard() * Amplitude{2.0, 0.0};
	CHECK(hipcheck::near(g[0][0].real, 2.0 / std::sqrt(2.0)));
	CHECK(hipcheck::near(g[1][1].real, -2.0 / std::sqrt(2.0)));
}

void qftKnownAnswers() {
	const state::QuantumState z = qft(3, 0);
	for(StateIndex y = 0; y < 8; ++y) {
		CHECK(hipcheck::near(z.amplitude(y).real, 1.0 / std::sqrt(8.0), 1e-12));
		CHECK(hipcheck::near(z.amplitude(y).imag, 0.0, 1e-12));
	}
	CHECK(hipcheck::near(z.totalProbability(), 1.0, 1e-12));
	const state::QuantumState f = qft(3, 5);
	for(StateIndex y = 0; y < 8; ++y) {
		const double ang = 2.0 * kPi * 5.0 * static_cast<double>(y) / 8.0;
		CHECK(hipcheck::near(f.amplitude(y).real, std::cos(ang) / std::sqrt(8.0), 1e-12));
		CHECK(hipcheck::near(f.amplitude(y).imag, std::sin(ang) / std::sqrt(8.0), 1e-12));
	}
}

void registerPrimitives() {
	QuantumRegister r(2, StateIndex{0}, hipcheck::deviceConfig());
	r.applyHadamard(0);
	r.applyHadamard(1);
	const Amplitude before = r.amplitude(3);
	r.applyPhaseFlipBasisState(3);
	CHECK(hipcheck::near(r.amplitude(3).real, -before.real, 1e-15));
	r.applyInversionAboutMean();
	CHECK(hipcheck::near(r.totalProbability(), 1.0, 1e-12));
}

void meanAwareInversion() {
	const AmplitudesVector v = {0.20, 0.10, -0.15, 0.05, 0.30, -0.20, 0.00, 0.10, -0.25, -0.15, 0.18, 0.12, -0.08, 0.04, 0.22, -0.06};
	QuantumRegister a(3, v, hipcheck::deviceConfig()), b(3, v, hipcheck::deviceConfig());
	Amplitude mean{0, 0};
	for(StateIndex s = 0; s < 8; ++s) {
		mean.real += b.amplitude(s).real;
		mean.imag += b.amplitude(s).imag;
	}
	mean.real /= 8.0;
	mean.imag /= 8.0;
	a.applyInversionAboutMean();
	b.applyInversionAboutMean(mean);
	CHECK(hipcheck::sameState(a, b, 1e-12));
	QuantumRegister c(3, v, hipcheck::deviceConfig());
	c.applyInversionAboutMeanMaterialized(mean);
	CHECK(hipcheck::sameState(a, c, 1e-12));
}

void groverErrorsAndVariants() {
	Mt19937RandomSource r1(42);
	CHECK(hipcheck::raises<std::invalid_argument>([&] { grover(3, {8}, r1); }));
	Mt19937RandomSource r2(700);
	CHECK(hipcheck::raises<std::invalid_argument>([&] { grover(3, {1, 8}, r2); }));
	for(uint32_t s = 300; s < 332; ++s) {
		Mt19937RandomSource rng(s);
		CHECK(in({1, 6}, grover(3, {1, 6}, rng)));
	}
	for(uint32_t s = 400; s < 416; ++s) {
		Mt19937RandomSource rng(s);
		CHECK(in({2, 5}, grover(3, {2, 5, 2}, rng)));
	}
	Mt19937RandomSource r3(500);
	CHECK(grover(2, {0, 1, 2, 3}, r3) < 4);
	for(bool materialized : {false, true}) {
		Mt19937RandomSource rng(600 + (materialized ? 10 : 0));
		CHECK(in({1, 6}, grover(3, {1, 6}, rng, materialized)));
	}
}

void operationSequences() {
	sim::ExecutionConfig ex;
	ex.backend = hipcheck::deviceConfig();
	circuit::Circuit c(2);
	c.h(0).swap(0, 1).swap(0, 1);
	const sim::RunResult r = sim::Simulator(ex).run(c, state::QuantumState::basis(2));
	CHECK(hipcheck::near(r.state.amplitude(0).real, 1.0 / std::sqrt(2.0), 1e-15));
	CHECK(hipcheck::near(r.state.amplitude(2).real, 1.0 / std::sqrt(2.0), 1e-15));
	CHECK(r.report.operationCount == 3 && r.report.strategy == StorageStrategyKind::Hip);
	sim::ExecutionConfig tr = ex;
	tr.observability.traceOperations = true;
	const sim::RunResult t = sim::Simulator(tr).run(c, state::QuantumState::basis(2));
	CHECK(t.report.operations.size() == 3 && t.report.operations[1].name == "swap");
	CHECK(experiment::toJson(t.report).find("\"strategy\": \"hip\"") != std::string::npos);
	Mt19937RandomSource a(12345), b(12345);
	CHECK(grover(5, {3}, a) == grover(5, {3}, b));
}

void defaultsRunOnTheDevice() {
	// default RegisterConfig / Simulator select the device backend in this build
	const sim::RunResult r = sim::Simulator().run(circuit::makeQft(4), state::QuantumState::basis(4, 3));
	CHECK(r.report.strategy == StorageStrategyKind::Hip);
	const state::QuantumState ref = qft(4, 3);
	const experiment::StateComparison cmp = experiment::compareStates(ref, r.state);
	CHECK(cmp.maxAbsAmplitudeError == 0.0 && cmp.bitwiseEqual);
}

}  // namespace

int main() {
	hipcheck::Suite suite("Hip circuit tests");
	suite.add("complexExpValues", complexExpValues);
	suite.add("groverFindsMark", groverFindsMark);
	suite.add("gateScalar", gateScalar);
	suite.add("qftKnownAnswers", qftKnownAnswers);
	suite.add("registerPrimitives", registerPrimitives);
	suite.add("meanAwareInversion", meanAwareInversion);
	suite.add("groverErrorsAndVariants", groverErrorsAndVariants);
	suite.add("operationSequences", operationSequences);
	suite.add("defaultsRunOnTheDevice", defaultsRunOnTheDevice);
	return suite.run();
}
