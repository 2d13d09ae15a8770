// Register error contract and storage behaviour on the Hip strategy: the
// checks of the reference's tests/integration/register/test_register_validation.cpp
// and tests/integration/storage/test_storage_parity.cpp (fixed-sample overlay
// measurement, copy/assignment independence, registry availability), through
// the public C++ API.
#include <iostream>
#include <vector>

#include "test_helpers.h"

using namespace tmfqs;

namespace {

class FixedRandomSource final : public IRandomSource {
	public:
		explicit FixedRandomSource(double v) : v_(v) {}
		double nextUnitDouble() override { return v_; }

	private:
		double v_;
};

void defaultRegister() {
	QuantumRegister r;
	CHECK(r.qubitCount() == 0u && r.stateCount() == 1u && r.amplitudeElementCount() == 2u);
	CHECK(hipcheck::near(r.amplitude(0).real, 1.0, 1e-15));
	CHECK(hipcheck::near(r.totalProbability(), 1.0, 1e-15));
	Mt19937RandomSource rng(1);
	CHECK(r.measure(rng) == 0u);
	CHECK(hipcheck::raises<std::out_of_range>([&] { r.applyHadamard(0); }));
}

void rangeErrors() {
	QuantumRegister r(2, StateIndex{0}, hipcheck::deviceConfig());
	CHECK(hipcheck::raises<std::out_of_range>([&] { (void)r.amplitude(4); }));
	CHECK(hipcheck::raises<std::out_of_range>([&] { (void)r.probability(4); }));
	CHECK(hipcheck::raises<std::out_of_range>([&] { r.applyHadamard(2); }));
	CHECK(hipcheck::raises<std::out_of_range>([&] { r.applyPauliX(2); }));
	CHECK(hipcheck::raises<std::out_of_range>([&] { r.applyControlledNot(0, 2); }));
	CHECK(hipcheck::raises<std::invalid_argument>([&] { r.applyControlledNot(1, 1); }));
	CHECK(hipcheck::raises<std::invalid_argument>([&] { r.applySwap(0, 0); }));
	CHECK(hipcheck::raises<std::invalid_argument>([&] { r.applyGate(QuantumGate::ControlledNot(), QubitList{1, 1}); }));
	CHECK(hipcheck::raises<std::out_of_range>([&] { r.applyGate(QuantumGate::Hadamard(), QubitList{2}); }));
	CHECK(hipcheck::raises<std::invalid_argument>([&] { r.applyGate(QuantumGate::Hadamard(), QubitList{0, 1}); }));
	QuantumRegister u(3, StateIndex{0}, hipcheck::deviceConfig());
	CHECK(hipcheck::raises<std::invalid_argument>([&] { u.initUniformSuperposition(BasisStateList{}); }));
	CHECK(hipcheck::raises<std::out_of_range>([&] { u.initUniformSuperposition(BasisStateList{0, 8}); }));
	FixedRandomSource bad(1.0);
	CHECK(hipcheck::raises<std::invalid_argument>([&] { (void)r.measure(bad); }));
}

void registryAvailability() {
	CHECK(StorageStrategyRegistry::isAvailable(StorageStrategyKind::Hip));
	CHECK(StorageStrategyRegistry::toString(StorageStrategyKind::Hip) == "hip");
	CHECK(StorageStrategyRegistry::toString(StorageStrategyKind::Dense) == "dense");
	const std::vector<std::string> names = StorageStrategyRegistry::listAvailable();
	CHECK(!names.empty() && names.front() == "hip");
	for(StorageStrategyKind k : {StorageStrategyKind::Dense, StorageStrategyKind::Blosc, StorageStrategyKind::Zfp}) {
		RegisterConfig cfg;
		cfg.strategy = k;
		CHECK(hipcheck::raises<std::runtime_error>([&] { (void)StateBackendFactory::create(4, cfg); }));
	}
	RegisterConfig autoCfg;
	autoCfg.strategy = StorageStrategyKind::Auto;
	autoCfg.autoThresholdBytes = 1u;
	CHECK(StorageStrategyRegistry::resolve(3, autoCfg) == StorageStrategyKind::Hip);
	{
		QuantumRegister viaAuto(3, 5, autoCfg);
		CHECK(viaAuto.amplitude(5).real == 1.0);
	}
	RegisterConfig keep = hipcheck::deviceConfig();
	keep.hip.tileQubits = 9;
	keep.hipOverrides.tileQubits = true;
	const RegisterConfig tuned = StorageStrategyRegistry::tuneConfig(20, keep, StorageStrategyKind::Hip);
	CHECK(tuned.hip.tileQubits == 9u && tuned.hipOverrides.tileQubits);
}

void overlayMeasurement() {
	const AmplitudesVector v = {0.20, 0.10, -0.15, 0.05, 0.30, -0.20, 0.00, 0.10, -0.25, -0.15, 0.18, 0.12, -0.08, 0.04, 0.22, -0.06};
	QuantumRegister lazy(3, v, hipcheck::deviceConfig());
	Amplitude mean{0, 0};
	for(StateIndex s = 0; s < 8; ++s) {
		mean.real += v[2 * s];
		mean.imag += v[2 * s + 1];
	}
	mean.real /= 8.0;
	mean.imag /= 8.0;
	lazy.applyInversionAboutMean(mean);
	QuantumRegister mat(3, v, hipcheck::deviceConfig());
	mat.applyInversionAboutMeanMaterialized(mean);
	const double tp = mat.totalProbability();
	CHECK(hipcheck::near(lazy.totalProbability(), tp, 1e-12));
	FixedRandomSource a(tp * 0.5), b(tp * 0.5);
	CHECK(lazy.measure(a) == mat.measure(b));
}

void copySemantics() {
	QuantumRegister r(5, StateIndex{0}, hipcheck::deviceConfig());
	r.beginOperationBatch();
	r.applyHadamard(0);
	r.applyHadamard(1);
	r.applyControlledNot(0, 2);
	r.applyControlledPhaseShift(1, 4, kPi / 9.0);
	r.applyPhaseFlipBasisState(9);
	r.applyInversionAboutMean();
	r.endOperationBatch();
	QuantumRegister copied(r);
	CHECK(hipcheck::sameState(r, copied, 0.0 + 1e-300));
	QuantumRegister assigned(5, StateIndex{0}, hipcheck::deviceConfig());
	assigned = r;
	CHECK(hipcheck::sameState(r, assigned, 1e-300));
	const Amplitude before = copied.amplitude(9);
	r.applyPhaseFlipBasisState(9);
	const Amplitude after = copied.amplitude(9);
	CHECK(before.real == after.real && before.imag == after.imag);
	CHECK(r.amplitude(9).real == -after.real || after.real == 0.0);
}

void builtinGateDispatch() {
	// applyGate with the built-in matrices equals the direct operations
	QuantumRegister a(4, StateIndex{0}, hipcheck::deviceConfig()), b(4, StateIndex{0}, hipcheck::deviceConfig());
	a.applyGate(QuantumGate::Hadamard(), QubitList{0});
	a.applyGate(QuantumGate::PauliX(), QubitList{3});
	a.applyGate(QuantumGate::ControlledNot(), QubitList{0, 2});
	a.applyGate(QuantumGate::ControlledPhaseShift(kPi / 7.0), QubitList{2, 1});
	a.applyGate(QuantumGate::Swap(), QubitList{1, 3});
	b.applyHadamard(0);
	b.applyPauliX(3);
	b.applyControlledNot(0, 2);
	b.applyControlledPhaseShift(2, 1, kPi / 7.0);
	b.applySwap(1, 3);
	CHECK(hipcheck::sameState(a, b, 1e-15));
}

}  // namespace

int main() {
	hipcheck::Suite suite("Hip register tests");
	suite.add("defaultRegister", defaultRegister);
	suite.add("rangeErrors", rangeErrors);
	suite.add("registryAvailability", registryAvailability);
	suite.add("overlayMeasurement", overlayMeasurement);
	suite.add("copySemantics", copySemantics);
	suite.add("builtinGateDispatch", builtinGateDispatch);
	return suite.run();
}
