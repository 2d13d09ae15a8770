// Complex helpers and the measurement RNG.
#include <cmath>

#include "tmfqs/core/math.h"
#include "tmfqs/core/random.h"

namespace tmfqs {

// reference: src/tmfqs/core/math.cpp:8-29 (same operation order).
Amplitude amplitudeMultiply(Amplitude a, Amplitude b) {
	return {a.real * b.real - a.imag * b.imag, a.real * b.imag + a.imag * b.real};
}

Amplitude amplitudeAdd(Amplitude a, Amplitude b) { return {a.real + b.real, a.imag + b.imag}; }

// cos and sin of one angle come from a single sincos call: that is what the
// reference's g++ build emits for the pair, and glibc's sincos can differ from
// separate cos/sin calls in the last bit.
Amplitude complexExp(Amplitude e) {
	const double r = std::exp(e.real);
	double s, c;
	__builtin_sincos(e.imag, &s, &c);
	return {r * c, r * s};
}

// reference: src/tmfqs/core/random.cpp:6-19.
Mt19937RandomSource::Mt19937RandomSource() : engine_(std::random_device{}()) {}
Mt19937RandomSource::Mt19937RandomSource(uint32_t seed) : engine_(seed) {}
double Mt19937RandomSource::nextUnitDouble() { return distribution_(engine_); }
void Mt19937RandomSource::reseed(uint32_t seed) { engine_.seed(seed); }

}  // namespace tmfqs
