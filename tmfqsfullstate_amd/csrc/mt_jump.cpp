// mt19937 jump-ahead polynomials (see mt_jump.h).
#include "mt_jump.h"

#include <mutex>
#include <stdexcept>

namespace tmfqs_mt {
namespace {

using namespace tmfqs_phase;
constexpr int W = kPolyWords;

inline int bit(const std::vector<uint64_t> &v, size_t i) { return (int)((v[i >> 6] >> (i & 63)) & 1u); }
inline void flip(std::vector<uint64_t> &v, size_t i) { v[i >> 6] ^= uint64_t{1} << (i & 63); }

// v ^= a << s (bit shift), a's words [0, na).
void xor_shifted(std::vector<uint64_t> &v, const uint64_t *a, size_t na, size_t s) {
	const size_t q = s >> 6, r = s & 63;
	for(size_t w = 0; w < na; ++w) {
		if(!a[w]) continue;
		v[q + w] ^= a[w] << r;
		if(r) v[q + w + 1] ^= a[w] >> (64 - r);
	}
}

// The 64 sequence bits starting at bit o of v (zero beyond its end).
inline uint64_t window(const std::vector<uint64_t> &v, size_t o) {
	const size_t q = o >> 6, r = o & 63;
	uint64_t x = q < v.size() ? v[q] >> r : 0;
	if(r && q + 1 < v.size()) x |= v[q + 1] << (64 - r);
	return x;
}

Poly berlekamp_massey_phi() {
	// Bit 0 of the raw words x_624, x_625, ... of the seed-5489 stream.
	const size_t nbits = 2 * (size_t)kMtDegree + 64;
	std::vector<uint32_t> x(kMtN + nbits);
	seed_state(5489u, x.data());
	for(size_t k = 0; k < nbits; ++k) x[k + kMtN] = mt_next(x[k], x[k + 1], x[k + kMtM]);
	// Reversed bit sequence: rev bit (nbits - 1 - i) = s_i, so that the
	// discrepancy sum_i c_i s_{n-i} is a word-wise AND with a window of rev.
	const size_t nw = (nbits + 63) / 64 + 2;
	std::vector<uint64_t> rev(nw, 0);
	for(size_t i = 0; i < nbits; ++i)
		if(x[kMtN + i] & 1u) flip(rev, nbits - 1 - i);
	std::vector<uint64_t> C(nw, 0), B(nw, 0), T;
	C[0] = B[0] = 1;
	size_t L = 0, m = 1;
	for(size_t n = 0; n < nbits; ++n) {
		const size_t off = nbits - 1 - n;  // rev bit off + i = s_{n-i}
		uint64_t acc = 0;
		for(size_t w = 0; w <= L / 64; ++w) acc ^= C[w] & window(rev, off + 64 * w);
		if(!(__builtin_popcountll(acc) & 1)) {
			++m;
			continue;
		}
		// C += x^m B (B's words past its degree are zero; the shifted words stay
		// inside the buffer because deg(x^m B) never exceeds the final L).
		if(m / 64 + 2 > nw) throw std::logic_error("mt19937 characteristic polynomial: shift out of range");
		if(2 * L <= n) {
			T = C;
			xor_shifted(C, B.data(), nw - m / 64 - 1, m);
			L = n + 1 - L;
			B.swap(T);
			m = 1;
		} else {
			xor_shifted(C, B.data(), nw - m / 64 - 1, m);
			++m;
		}
	}
	if(L != (size_t)kMtDegree || !bit(C, L)) throw std::logic_error("mt19937 characteristic polynomial: unexpected degree");
	// phi_j = C_{L-j}.
	Poly phi(W, 0);
	for(size_t j = 0; j <= L; ++j)
		if(bit(C, L - j)) flip(phi, j);
	return phi;
}

// a^2 mod phi for deg a < 19937.
Poly sqr_mod(const Poly &a, const Poly &phi) {
	std::vector<uint64_t> t(2 * W + 2, 0);
	for(size_t i = 0; i < (size_t)kMtDegree; ++i)
		if(bit(a, i)) flip(t, 2 * i);
	for(size_t pos = 2 * (size_t)kMtDegree - 2; pos >= (size_t)kMtDegree; --pos)
		if(bit(t, pos)) xor_shifted(t, phi.data(), W, pos - kMtDegree);
	return Poly(t.begin(), t.begin() + W);
}

std::mutex g_mu;

}  // namespace

void seed_state(uint32_t seed, uint32_t st[kMtN]) {
	st[0] = seed;
	for(int i = 1; i < kMtN; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
}

const Poly &characteristic() {
	static const Poly phi = berlekamp_massey_phi();
	return phi;
}

const Poly &pow2_mod(unsigned m) {
	if(m >= 64) throw std::out_of_range("mt19937 jump exponent");
	static std::vector<Poly> cache;  // cache[i] = x^(2^i) mod phi
	std::lock_guard<std::mutex> lk(g_mu);
	const Poly &phi = characteristic();
	while(cache.size() <= m) {
		const unsigned i = (unsigned)cache.size();
		if((uint64_t{1} << i) < (uint64_t)kMtDegree) {
			Poly p(W, 0);
			flip(p, (size_t)1 << i);
			cache.push_back(std::move(p));
		} else {
			cache.push_back(sqr_mod(cache.back(), phi));
		}
	}
	return cache[m];
}

void jump_host(uint32_t st[kMtN], const Poly &g) {
	std::vector<uint32_t> x(kJumpSeqWords);
	for(int i = 0; i < kMtN; ++i) x[i] = st[i];
	for(int k = 0; k + kMtN < kJumpSeqWords; ++k) x[k + kMtN] = mt_next(x[k], x[k + 1], x[k + kMtM]);
	uint32_t acc[kMtN] = {};
	for(int w = 0; w < W; ++w)
		for(uint64_t b = g[w]; b; b &= b - 1) {
			const int j = 64 * w + __builtin_ctzll(b);
			for(int m = 0; m < kMtN; ++m) acc[m] ^= x[j + m];
		}
	for(int m = 0; m < kMtN; ++m) st[m] = acc[m];
}

}  // namespace tmfqs_mt
