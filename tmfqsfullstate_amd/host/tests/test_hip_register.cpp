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
	HIP_TEST_REQUIRE(r.qubitCount() == 0u && r.stateCount() == 1u && r.amplitudeElementCount() == 2u);
	HIP_TEST_REQUIRE(hip_test::approxEqual(r.amplitude(0).real, 1.0, 1e-15));
	HIP_TEST_REQUIRE(hip_test::approxEqual(r.totalProbability(), 1.0, 1e-15));
	Mt19937RandomSource rng(1);
	HIP_TEST_REQUIRE(r.measure(rng) == 0u);
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { r.applyHadamard(0); }));
}

void rangeErrors() {
	QuantumRegister r(2, StateIndex{0}, hip_test::hipConfig());
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { (void)r.amplitude(4); }));
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { (void)r.probability(4); }));
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { r.applyHadamard(2); }));
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { r.applyPauliX(2); }));
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { r.applyControlledNot(0, 2); }));
	HIP_TEST_REQUIRE(throwsAs<std::invalid_argument>([&] { r.applyControlledNot(1, 1); }));
	HIP_TEST_REQUIRE(throwsAs<std::invalid_argument>([&] { r.applySwap(0, 0); }));
	HIP_TEST_REQUIRE(throwsAs<std::invalid_argument>([&] { r.applyGate(QuantumGate::ControlledNot(), QubitList{1, 1}); }));
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { r.applyGate(QuantumGate::Hadamard(), QubitList{2}); }));
	HIP_TEST_REQUIRE(throwsAs<std::invalid_argument>([&] { r.applyGate(QuantumGate::Hadamard(), QubitList{0, 1}); }));
	QuantumRegister u(3, StateIndex{0}, hip_test::hipConfig());
	HIP_TEST_REQUIRE(throwsAs<std::invalid_argument>([&] { u.initUniformSuperposition(BasisStateList{}); }));
	HIP_TEST_REQUIRE(throwsAs<std::out_of_range>([&] { u.initUniformSuperposition(BasisStateList{0, 8}); }));
	FixedRandomSource bad(1.0);
	HIP_TEST_REQUIRE(throwsAs<std::invalid_argument>([&] { (void)r.measure(bad); }));
}

void registryAvailability() {
	HIP_TEST_REQUIRE(StorageStrategyRegistry::isAvailable(StorageStrategyKind::Hip));
	HIP_TEST_REQUIRE(StorageStrategyRegistry::toString(StorageStrategyKind::Hip) == "hip");
	HIP_TEST_REQUIRE(StorageStrategyRegistry::toString(StorageStrategyKind::Dense) == "dense");
	const std::vector<std::string> names = StorageStrategyRegistry::listAvailable();
	HIP_TEST_REQUIRE(!names.empty() && names.front() == "hip");
	for(StorageStrategyKind k : {StorageStrategyKind::Dense, StorageStrategyKind::Blosc, StorageStrategyKind::Zfp}) {
		RegisterConfig cfg;
		cfg.strategy = k;
		HIP_TEST_REQUIRE(throwsAs<std::runtime_error>([&] { (void)StateBackendFactory::create(4, cfg); }));
	}
	RegisterConfig autoCfg;
	autoCfg.strategy = StorageStrategyKind::Auto;
	autoCfg.autoThresholdBytes = 1u;
	HIP_TEST_REQUIRE(StorageStrategyRegistry::resolve(3, autoCfg) == StorageStrategyKind::Hip);
	{
		QuantumRegister viaAuto(3, 5, autoCfg);
		HIP_TEST_REQUIRE(viaAuto.amplitude(5).real == 1.0);
	}
	RegisterConfig keep = hip_test::hipConfig();
	keep.hip.tileQubits = 9;
	keep.hipOverrides.tileQubits = true;
	const RegisterConfig tuned = StorageStrategyRegistry::tuneConfig(20, keep, StorageStrategyKind::Hip);
	HIP_TEST_REQUIRE(tuned.hip.tileQubits == 9u && tuned.hipOverrides.tileQubits);
}

void overlayMeasurement() {
	const AmplitudesVector v = {0.20, 0.10, -0.15, 0.05, 0.30, -0.20, 0.00, 0.10, -0.25, -0.15, 0.18, 0.12, -0.08, 0.04, 0.22, -0.06};
	QuantumRegister lazy(3, v, hip_test::hipConfig());
	Amplitude mean{0, 0};
	for(StateIndex s = 0; s < 8; ++s) {
		mean.real += v[2 * s];
		mean.imag += v[2 * s + 1];
	}
	mean.real /= 8.0;
	mean.imag /= 8.0;
	lazy.applyInversionAboutMean(mean);
	QuantumRegister mat(3, v, hip_test::hipConfig());
	mat.applyInversionAboutMeanMaterialized(mean);
	const double tp = mat.totalProbability();
	HIP_TEST_REQUIRE(hip_test::approxEqual(lazy.totalProbability(), tp, 1e-12));
	FixedRandomSource a(tp * 0.5), b(tp * 0.5);
	HIP_TEST_REQUIRE(lazy.measure(a) == mat.measure(b));
}

void copySemantics() {
	QuantumRegister r(5, StateIndex{0}, hip_test::hipConfig());
	r.beginOperationBatch();
	r.applyHadamard(0);
	r.applyHadamard(1);
	r.applyControlledNot(0, 2);
	r.applyControlledPhaseShift(1, 4, kPi / 9.0);
	r.applyPhaseFlipBasisState(9);
	r.applyInversionAboutMean();
	r.endOperationBatch();
	QuantumRegister copied(r);
	hip_test::assertRegistersClose(r, copied, 0.0 + 1e-300);
	QuantumRegister assigned(5, StateIndex{0}, hip_test::hipConfig());
	assigned = r;
	hip_test::assertRegistersClose(r, assigned, 1e-300);
	const Amplitude before = copied.amplitude(9);
	r.applyPhaseFlipBasisState(9);
	const Amplitude after = copied.amplitude(9);
	HIP_TEST_REQUIRE(before.real == after.real && before.imag == after.imag);
	HIP_TEST_REQUIRE(r.amplitude(9).real == -after.real || after.real == 0.0);
}

void builtinGateDispatch() {
	// applyGate with the built-in matrices equals the direct operations
	QuantumRegister a(4, StateIndex{0}, hip_test::hipConfig()), b(4, StateIndex{0}, hip_test::hipConfig());
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
	hip_test::assertRegistersClose(a, b, 1e-15);
}

}  // namespace

int main() {
	try {
		defaultRegister();
		rangeErrors();
		registryAvailability();
		overlayMeasurement();
		copySemantics();
		builtinGateDispatch();
	} catch(const std::exception &e) {
		std::cerr << "FAILED: " << e.what() << "\n";
		return 1;
	}
	std::cout << "Hip register tests passed.\n";
	return 0;
}
