// Jump-ahead for std::mt19937 over GF(2), so that independent workgroups can
// each generate a contiguous slice of one seed's word stream.
//
// A generator state is the buffer of the last 624 raw (untempered) words
// x_k .. x_{k+623}, with the next output temper(x_{k+624}) -- the form
// std::mt19937 holds right after seeding and after every twist.  The word
// recurrence is linear over GF(2); every bit sequence it produces satisfies
// the characteristic polynomial phi (degree 19937) of the transition T, so
// for g = x^J mod phi the state J words on is
//   x_{k+J+m} = XOR over the set coefficients j of g of x_{k+j+m},  m < 624,
// which needs only the first 19937 + 623 words from the state (the low 31
// bits of the oldest word, which never reach an output, may differ).
// phi is found once with Berlekamp-Massey on 2 * 19937 output bits; the jump
// polynomials x^(2^m) mod phi come from repeated squaring and are cached.
#pragma once

#include <cstdint>
#include <vector>

#include "phase_math.h"

namespace tmfqs_mt {

using Poly = std::vector<uint64_t>;  // tmfqs_phase::kPolyWords words; bit i = coefficient of x^i

// phi (monic, degree 19937).
const Poly &characteristic();
// x^(2^m) mod phi, m < 64.
const Poly &pow2_mod(unsigned m);
// std::mt19937(seed)'s state right after seeding.
void seed_state(uint32_t seed, uint32_t st[tmfqs_phase::kMtN]);
// st <- the state g's jump reaches (host evaluation, for checks and small jumps).
void jump_host(uint32_t st[tmfqs_phase::kMtN], const Poly &g);
// Words of the raw sequence a jump reads: x_k .. x_{k+19936+623}.
constexpr int kJumpSeqWords = tmfqs_phase::kMtDegree - 1 + tmfqs_phase::kMtN;

}  // namespace tmfqs_mt
