// randomPhase arithmetic shared by the device kernels and the host checks:
// the mt19937 word recurrence and tempering, std::generate_canonical<double,
// 53> over two words, and cos/sin of the phase in double-double with a flag
// for results too close to a rounding midpoint to be trusted.
//
// The reference draws phase = uniform_real_distribution(0, 2 acos(-1)) from
// std::mt19937(seed) and stores norm * cos(phase), norm * sin(phase)
// (src/tmfqs/state/quantum_state.cpp:264-277); its compiler turns the cos/sin
// pair into one glibc sincos call.  glibc's sincos is not correctly rounded:
// measured against 113-bit references on 4e8 uniform phases its results lie
// up to 0.006 ulp (|v| in [1/2, 1)), 0.011 (|v| in [1/4, 1/2)), 0.023
// ([1/8, 1/4)), 0.009 ([1/16, 1/8)), 0.0024, 0.0004 and below 1e-4 ulp
// (smaller binades) beyond the rounding midpoint, and 0.06 ulp for cos at
// the doubles next to pi/2 and 3 pi/2 (|v| near 2^-52; 4e7 arguments within
// 2^-8 of multiples of pi/2 showed no other case).  The device therefore
// rounds the double-double value (within 2^-76 of the exact one, relative)
// to nearest and flags it when it lies within midpoint_band(binade) ulp of a
// midpoint -- at least five times the measured excess in every binade, and
// every value below 2^-30 -- so that the host recomputes exactly those with
// glibc.  Outside the band the correctly rounded value and glibc's agree.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define TMFQS_PM_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define TMFQS_PM_HD inline
#endif

#include <math.h>

namespace tmfqs_phase {

// ------------------------------------------------------------ mt19937
// The published MT19937 parameters (Matsumoto and Nishimura 1998; the values
// std::mt19937 fixes in [rand.predef]).
constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr uint32_t kMtMatrixA = 0x9908b0dfu;
constexpr uint32_t kMtUpper = 0x80000000u;
constexpr uint32_t kMtLower = 0x7fffffffu;
constexpr int kMtDegree = 19937;  // degree of the characteristic polynomial
constexpr int kPolyWords = (kMtDegree + 1 + 63) / 64;  // 312 words hold a polynomial of degree <= 19937

// The next raw word from x_k (upper bit), x_{k+1} (lower 31 bits) and x_{k+397}.
TMFQS_PM_HD uint32_t mt_next(uint32_t xk, uint32_t xk1, uint32_t xkm) {
	const uint32_t y = (xk & kMtUpper) | (xk1 & kMtLower);
	return xkm ^ (y >> 1) ^ ((y & 1u) ? kMtMatrixA : 0u);
}

TMFQS_PM_HD uint32_t mt_temper(uint32_t y) {
	y ^= y >> 11;
	y ^= (y << 7) & 0x9d2c5680u;
	y ^= (y << 15) & 0xefc60000u;
	y ^= y >> 18;
	return y;
}

// libstdc++'s generate_canonical<double, 53> on a 32-bit engine: two words,
// sum = w0 + w1 * 2^32 rounded to double, divided by 2^64, kept below 1.
TMFQS_PM_HD double canonical53(uint32_t w0, uint32_t w1) {
	double s = (double)w0;
	s = s + (double)w1 * 4294967296.0;
	double r = s * 5.421010862427522170037264e-20;  // 2^-64, exact
	if(r >= 1.0) r = 0x1.fffffffffffffp-1;
	return r;
}

// ------------------------------------------------------------ double-double
struct DD {
	double hi, lo;
};

TMFQS_PM_HD DD two_sum(double a, double b) {
	const double s = a + b;
	const double bb = s - a;
	const double e = (a - (s - bb)) + (b - bb);
	return {s, e};
}

TMFQS_PM_HD DD fast_two_sum(double a, double b) {  // |a| >= |b|
	const double s = a + b;
	return {s, b - (s - a)};
}

TMFQS_PM_HD DD two_prod(double a, double b) {
	const double p = a * b;
	return {p, fma(a, b, -p)};
}

TMFQS_PM_HD DD dd_add(DD a, DD b) {
	DD s = two_sum(a.hi, b.hi);
	const DD t = two_sum(a.lo, b.lo);
	s.lo += t.hi;
	s = fast_two_sum(s.hi, s.lo);
	s.lo += t.lo;
	return fast_two_sum(s.hi, s.lo);
}

TMFQS_PM_HD DD dd_add_d(DD a, double b) {
	DD s = two_sum(a.hi, b);
	s.lo += a.lo;
	return fast_two_sum(s.hi, s.lo);
}

TMFQS_PM_HD DD dd_mul(DD a, DD b) {
	DD p = two_prod(a.hi, b.hi);
	p.lo += a.hi * b.lo + a.lo * b.hi;
	return fast_two_sum(p.hi, p.lo);
}

TMFQS_PM_HD DD dd_mul_d(DD a, double b) {
	DD p = two_prod(a.hi, b);
	p.lo += a.lo * b;
	return fast_two_sum(p.hi, p.lo);
}

// 1/m as a double-double (m an integer below 2^53): the remainder of the
// rounded reciprocal is exact through fma.
TMFQS_PM_HD DD dd_recip(double m) {
	const double h = 1.0 / m;
	return {h, fma(-h, m, 1.0) / m};
}

// pi/2 to 159 bits as three doubles (hi + mid + lo).
constexpr double kPio2Hi = 0x1.921fb54442d18p0;
constexpr double kPio2Mid = 0x1.1a62633145c07p-54;
constexpr double kPio2Lo = -0x1.f1976b7ed8fbcp-110;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

// cos and sin of t, |t| <= pi/4 (slightly beyond is fine): Taylor series,
// the first five terms in double-double, the rest (below 2^-25 of the value)
// in double.
TMFQS_PM_HD void dd_sincos_reduced(DD t, DD &s_out, DD &c_out) {
	const DD u = dd_add_d(two_prod(t.hi, t.hi), 2.0 * t.hi * t.lo);
	const double ud = u.hi;
	// cos: sum_k (-1)^k u^k / (2k)!, tail k = 5..11 in double.
	double tc = -1.0 / 1124000727777607680000.0;     // 1/22!
	tc = tc * ud + 1.0 / 2432902008176640000.0;      // 1/20!
	tc = tc * ud - 1.0 / 6402373705728000.0;         // 1/18!
	tc = tc * ud + 1.0 / 20922789888000.0;           // 1/16!
	tc = tc * ud - 1.0 / 87178291200.0;              // 1/14!
	tc = tc * ud + 1.0 / 479001600.0;                // 1/12!
	tc = tc * ud - 1.0 / 3628800.0;                  // 1/10!
	DD c = {tc, 0.0};
	c = dd_add(dd_mul(c, u), dd_recip(40320.0));  // + 1/8!
	c = dd_add(dd_mul(c, u), dd_mul_d(dd_recip(720.0), -1.0));
	c = dd_add(dd_mul(c, u), dd_recip(24.0));
	c = dd_add(dd_mul(c, u), DD{-0.5, 0.0});
	c = dd_add(dd_mul(c, u), DD{1.0, 0.0});
	// sin / t: sum_k (-1)^k u^k / (2k+1)!, tail k = 5..11 in double.
	double ts = -1.0 / 25852016738884976640000.0;  // 1/23!
	ts = ts * ud + 1.0 / 51090942171709440000.0;    // 1/21!
	ts = ts * ud - 1.0 / 121645100408832000.0;      // 1/19!
	ts = ts * ud + 1.0 / 355687428096000.0;         // 1/17!
	ts = ts * ud - 1.0 / 1307674368000.0;           // 1/15!
	ts = ts * ud + 1.0 / 6227020800.0;              // 1/13!
	ts = ts * ud - 1.0 / 39916800.0;                // 1/11!
	DD s = {ts, 0.0};
	s = dd_add(dd_mul(s, u), dd_recip(362880.0));  // + 1/9!
	s = dd_add(dd_mul(s, u), dd_mul_d(dd_recip(5040.0), -1.0));
	s = dd_add(dd_mul(s, u), dd_recip(120.0));
	s = dd_add(dd_mul(s, u), dd_mul_d(dd_recip(6.0), -1.0));
	s = dd_add(dd_mul(s, u), DD{1.0, 0.0});
	s_out = dd_mul(s, t);
	c_out = c;
}

// cos and sin of x in [0, 2 pi) as double-doubles.
TMFQS_PM_HD void dd_sincos(double x, DD &s, DD &c) {
	const double k = rint(x * kTwoOverPi);
	const DD p = two_prod(k, kPio2Hi);
	DD t = two_sum(x, -p.hi);
	t = dd_add_d(t, -p.lo);
	t = dd_add(t, dd_mul_d(DD{kPio2Mid, 0.0}, -k));
	t = dd_add_d(t, -k * kPio2Lo);
	DD sr, cr;
	dd_sincos_reduced(t, sr, cr);
	switch((int)k & 3) {
		case 0: s = sr; c = cr; break;
		case 1: s = cr; c = DD{-sr.hi, -sr.lo}; break;
		case 2: s = DD{-sr.hi, -sr.lo}; c = DD{-cr.hi, -cr.lo}; break;
		default: s = DD{-cr.hi, -cr.lo}; c = sr; break;
	}
}

// Width of the flag band around rounding midpoints, in ulps of the result,
// by binade b (|v| in [2^-(b+1), 2^-b)); see the header comment.
TMFQS_PM_HD double midpoint_band(int b) {
	switch(b) {
		case 0: return 0x1p-5;
		case 1: return 0x1p-4;
		case 2: return 0x1p-3;
		case 3: return 0x1p-4;
		case 4: return 0x1p-6;
		case 5: return 0x1p-8;
		default: return b < 30 ? 0x1p-10 : 1.0;  // 1: every value (distance <= 1/2)
	}
}

TMFQS_PM_HD uint64_t bits_of(double v) {
	union {
		double d;
		uint64_t u;
	} c{v};
	return c.u;
}

TMFQS_PM_HD double double_of(uint64_t u) {
	union {
		uint64_t u;
		double d;
	} c{u};
	return c.d;
}

// Rounds the normalized double-double v (|v.hi| <= 1) to nearest; *near is
// set when the exact value may lie within the band of a midpoint, i.e. where
// glibc's rounding can differ.
TMFQS_PM_HD double round_flagged(DD v, bool *near) {
	*near = false;
	if(v.hi == 0.0) return v.hi;  // sin(0): glibc's value is exactly 0 too
	const uint64_t b = bits_of(v.hi);
	const int e = (int)((b >> 52) & 0x7ff) - 1023;  // |v.hi| in [2^e, 2^(e+1))
	const int binade = e >= 0 ? 0 : -e - 1;
	const double band = midpoint_band(binade);
	if(band >= 0.5) {
		*near = true;
		return v.hi;
	}
	if(v.lo == 0.0) return v.hi;  // exactly v.hi: half an ulp from either midpoint
	const bool away = (v.lo > 0.0) == (v.hi > 0.0);  // neighbour towards v.lo has the larger magnitude
	const double gap = fabs(double_of(away ? b + 1 : b - 1) - v.hi);  // a power of two
	// Distance to the midpoint 0.5 - |lo| / gap below the band, without a
	// division: (0.5 - band) * gap is exact.
	*near = fabs(v.lo) > (0.5 - band) * gap;
	return v.hi;
}

}  // namespace tmfqs_phase
