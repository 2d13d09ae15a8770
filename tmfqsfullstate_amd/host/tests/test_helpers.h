// Minimal assertion helpers for the C++ test drivers (the reference's tests
// use the same style: tests/integration/common/test_helpers.h).
#pragma once

#include <cmath>
#include <sstream>
#include <stdexcept>

#include "tmfqsfs.h"

namespace hip_test {

[[noreturn]] inline void fail(const char *expr, const char *file, int line) {
	std::ostringstream m;
	m << file << ":" << line << ": test assertion failed: " << expr;
	throw std::runtime_error(m.str());
}

inline void require(bool c, const char *expr, const char *file, int line) {
	if(!c) fail(expr, file, line);
}

inline bool approxEqual(double a, double b, double tol = 1e-10) { return std::fabs(a - b) < tol; }

inline tmfqs::RegisterConfig hipConfig() {
	tmfqs::RegisterConfig cfg;
	cfg.strategy = tmfqs::StorageStrategyKind::Hip;
	return cfg;
}

inline void assertRegistersClose(const tmfqs::QuantumRegister &a, const tmfqs::QuantumRegister &b, double tol) {
	require(a.stateCount() == b.stateCount(), "stateCount", __FILE__, __LINE__);
	const tmfqs::AmplitudesVector x = a.amplitudes(), y = b.amplitudes();
	for(size_t i = 0; i < x.size(); ++i) require(approxEqual(x[i], y[i], tol), "amplitude", __FILE__, __LINE__);
	require(approxEqual(a.totalProbability(), b.totalProbability(), tol), "totalProbability", __FILE__, __LINE__);
}

}  // namespace hip_test

#define HIP_TEST_REQUIRE(expr) ::hip_test::require((expr), #expr, __FILE__, __LINE__)

template <typename Exc, typename Fn>
bool throwsAs(Fn fn) {
	try {
		fn();
	} catch(const Exc &) {
		return true;
	} catch(...) {
		return false;
	}
	return false;
}
