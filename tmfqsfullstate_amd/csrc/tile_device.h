// Tiled multi-operation passes (register stages over an LDS tile) shared by
// the single-state, dual (adjoint) and batched launches.  Internal.
#pragma once

#include "device_common.h"

namespace tmfqs_hip {

template <typename R, int K>
__device__ __forceinline__ void tile_uk(typename V2T<R>::t *tile, u64 nloc, const TileOp &op, const double *__restrict__ mats) {
	constexpr int D = 1 << K;
	unsigned sorted[K];
	for(int i = 0; i < K; ++i) sorted[i] = op.jl[i];
	for(int i = 0; i < K; ++i)
		for(int j = i + 1; j < K; ++j)
			if(sorted[j] < sorted[i]) {
				const unsigned t = sorted[i];
				sorted[i] = sorted[j];
				sorted[j] = t;
			}
	const double *g = mats + op.mat;
	for(u64 b = threadIdx.x; b < (nloc >> K); b += blockDim.x) {
		u64 base = b;
		for(int i = 0; i < K; ++i) base = ins0(base, sorted[i]);
		R xr[D], xi[D];
		u64 idx[D];
		bool nz = false;
#pragma unroll
		for(int r = 0; r < D; ++r) {
			u64 off = 0;
#pragma unroll
			for(int i = 0; i < K; ++i)
				if((r >> (K - 1 - i)) & 1) off |= 1ull << op.jl[i];
			idx[r] = base | off;
			const auto v = tile[idx[r]];
			xr[r] = v.x;
			xi[r] = v.y;
			nz = nz || (v.x != 0 || v.y != 0);
		}
		if(!nz) continue;
#pragma unroll
		for(int r = 0; r < D; ++r) {
			R nr = 0, ni = 0;
#pragma unroll
			for(int c = 0; c < D; ++c) {
				const R gr = (R)g[2 * (r * D + c)], gi = (R)g[2 * (r * D + c) + 1];
				nr += gr * xr[c] - gi * xi[c];
				ni += gr * xi[c] + gi * xr[c];
			}
			typename V2T<R>::t o;
			o.x = nr;
			o.y = ni;
			tile[idx[r]] = o;
		}
	}
}

// Applies one tile op to a tile of 2^T amplitudes held in LDS.  `base` is the
// global index of the tile's element 0; local index l maps to global
// base | (l & lowmask) | hitab[l >> C].
template <typename R, int MAXK = 5>
__device__ __forceinline__ void tile_apply(typename V2T<R>::t *tile, const TileOp &op, u64 nloc, u64 base, unsigned C,
                                           const u64 *hitab, const double *__restrict__ mats,
                                           const TwiddleTerm *__restrict__ terms, const double2 *__restrict__ ptab) {
	using V = typename V2T<R>::t;
	const u64 lowmask = (1ull << C) - 1ull;
	const R hc = (R)(1.0 / sqrt(2.0));
	switch(op.kind) {
		case TOP_H: {
			const u64 m = 1ull << op.j0;
			for(u64 q = threadIdx.x; q < (nloc >> 1); q += blockDim.x) {
				const u64 l0 = ins0(q, op.j0), l1 = l0 | m;
				const V x = tile[l0], y = tile[l1];
				if(x.x == 0 && x.y == 0 && y.x == 0 && y.y == 0) continue;
				V o0, o1;
				o0.x = (x.x + y.x) * hc;
				o0.y = (x.y + y.y) * hc;
				o1.x = (x.x - y.x) * hc;
				o1.y = (x.y - y.y) * hc;
				tile[l0] = o0;
				tile[l1] = o1;
			}
			break;
		}
		case TOP_X: {
			const u64 m = 1ull << op.j0;
			for(u64 q = threadIdx.x; q < (nloc >> 1); q += blockDim.x) {
				const u64 l0 = ins0(q, op.j0), l1 = l0 | m;
				if((l0 & op.lcond) != op.lcond) continue;
				const V x = tile[l0], y = tile[l1];
				if(x.x == y.x && x.y == y.y) continue;
				tile[l0] = y;
				tile[l1] = x;
			}
			break;
		}
		case TOP_U1: {
			const u64 m = 1ull << op.j0;
			const double *g = mats + op.mat;
			const R g00r = (R)g[0], g00i = (R)g[1], g01r = (R)g[2], g01i = (R)g[3];
			const R g10r = (R)g[4], g10i = (R)g[5], g11r = (R)g[6], g11i = (R)g[7];
			for(u64 q = threadIdx.x; q < (nloc >> 1); q += blockDim.x) {
				const u64 l0 = ins0(q, op.j0), l1 = l0 | m;
				const V x = tile[l0], y = tile[l1];
				if((x.x == 0 && x.y == 0) && (y.x == 0 && y.y == 0)) continue;
				V o0, o1;
				o0.x = g00r * x.x - g00i * x.y + g01r * y.x - g01i * y.y;
				o0.y = g00r * x.y + g00i * x.x + g01r * y.y + g01i * y.x;
				o1.x = g10r * x.x - g10i * x.y + g11r * y.x - g11i * y.y;
				o1.y = g10r * x.y + g10i * x.x + g11r * y.y + g11i * y.x;
				tile[l0] = o0;
				tile[l1] = o1;
			}
			break;
		}
		case TOP_UK: {
			switch(op.k) {
				case 2: tile_uk<R, 2>(tile, nloc, op, mats); break;
				case 3: tile_uk<R, 3>(tile, nloc, op, mats); break;
				case 4: tile_uk<R, 4>(tile, nloc, op, mats); break;
				case 5:
					if constexpr(MAXK >= 5) tile_uk<R, 5>(tile, nloc, op, mats);
					break;
				default: break;
			}
			break;
		}
		case TOP_CP: {
			const R cr = (R)op.cr, ci = (R)op.ci;
			const unsigned nb = __popcll(op.lcond);
			for(u64 q = threadIdx.x; q < (nloc >> nb); q += blockDim.x) {
				u64 l = q;
				u64 cm = op.lcond;
				while(cm) {
					const unsigned b = __ffsll((long long)cm) - 1;
					l = ins0(l, b) | (1ull << b);
					cm &= cm - 1;
				}
				const V x = tile[l];
				if(x.x == 0 && x.y == 0) continue;
				V o;
				o.x = cr * x.x - ci * x.y;
				o.y = cr * x.y + ci * x.x;
				tile[l] = o;
			}
			break;
		}
		case TOP_LADDER: {
			const unsigned nb = __popcll(op.lcond);
			for(u64 q = threadIdx.x; q < (nloc >> nb); q += blockDim.x) {
				u64 l = q;
				u64 cm = op.lcond;
				while(cm) {
					const unsigned b = __ffsll((long long)cm) - 1;
					l = ins0(l, b) | (1ull << b);
					cm &= cm - 1;
				}
				const u64 g = base | (l & lowmask) | hitab[l >> C];
				const u64 F = (g & op.amask) << op.shift;
				if(F == 0) continue;
				tile[l] = rot<V>(tile[l], phase_of_range(F, ptab, op.cj0, op.cj1));
			}
			break;
		}
		case TOP_TWIDDLE: {
			const TwiddleTerm *tt = terms + op.twid;
			for(u64 l = threadIdx.x; l < nloc; l += blockDim.x) {
				const u64 g = base | (l & lowmask) | hitab[l >> C];
				u64 F = 0;
				for(unsigned t = 0; t < op.nterms; ++t)
					if(g & tt[t].cond) F += (g & tt[t].mask) << tt[t].shift;
				if(F == 0) continue;
				tile[l] = rot<V>(tile[l], phase_of_range(F, ptab, op.cj0, op.cj1));
			}
			break;
		}
		default: break;
	}
}

template <typename R>
__device__ __forceinline__ void tile_hitab(u64 *hitab, const TilePass &P) {
	for(unsigned j = threadIdx.x; j < (1u << P.nh); j += blockDim.x) {
		u64 v = 0;
		for(unsigned b = 0; b < P.nh; ++b)
			if((j >> b) & 1u) v |= 1ull << P.hbits[b];
		hitab[j] = v;
	}
}

// ------------------------------------------------------------------------
// Register-staged tiled pass.  The tile (2^T amplitudes, local index l maps to
// global base | (l & lowmask) | hitab[l >> C]) is loaded into LDS with
// coalesced reads; each register stage then moves 2^W amplitudes per thread
// into registers, applies its ops there and writes them back, so a run of
// gates costs one LDS round trip and one barrier instead of one per gate.
template <typename V> __device__ __forceinline__ bool is_zero(V x) { return x.x == 0 && x.y == 0; }

// The pair ops skip all-zero pairs as the reference does
// (dense_state_backend.cpp:263, gate_block_apply.h:117-143).  When no active
// lane of the wavefront holds such a pair, the unchecked form runs: the same
// arithmetic for every pair, in one block the compiler schedules across pairs
// (the checked form's per-pair branches serialise them).
template <typename V, int W, int J>
__device__ __forceinline__ bool any_zero_pair(const V *v) {
	bool z = false;
#pragma unroll
	for(int r = 0; r < (1 << W); ++r)
		if(!(r & (1 << J))) z = z | (is_zero(v[r]) & is_zero(v[r | (1 << J)]));
	return z;
}

template <typename V, typename R, int W, int J, bool CHECKED>
__device__ __forceinline__ void reg_h_pairs(V *v, R hc) {
#pragma unroll
	for(int r = 0; r < (1 << W); ++r) {
		if(r & (1 << J)) continue;
		const V x = v[r], y = v[r | (1 << J)];
		if(CHECKED && is_zero(x) && is_zero(y)) continue;
		V o0, o1;
		o0.x = (x.x + y.x) * hc;
		o0.y = (x.y + y.y) * hc;
		o1.x = (x.x - y.x) * hc;
		o1.y = (x.y - y.y) * hc;
		v[r] = o0;
		v[r | (1 << J)] = o1;
	}
}

template <typename V, typename R, int W, int J>
__device__ __forceinline__ void reg_h(V *v, R hc) {
	if(__any(any_zero_pair<V, W, J>(v))) reg_h_pairs<V, R, W, J, true>(v, hc);
	else reg_h_pairs<V, R, W, J, false>(v, hc);
}

template <typename V, int W, int J>
__device__ __forceinline__ void reg_x(V *v, const unsigned *lidx, u64 lcond) {
#pragma unroll
	for(int r = 0; r < (1 << W); ++r) {
		if(r & (1 << J)) continue;
		if((lidx[r] & lcond) != lcond) continue;
		const V x = v[r], y = v[r | (1 << J)];
		if(x.x == y.x && x.y == y.y) continue;
		v[r] = y;
		v[r | (1 << J)] = x;
	}
}

template <typename V, typename R, int W, int J, bool CHECKED>
__device__ __forceinline__ void reg_u1_pairs(V *v, R g00r, R g00i, R g01r, R g01i, R g10r, R g10i, R g11r, R g11i) {
#pragma unroll
	for(int r = 0; r < (1 << W); ++r) {
		if(r & (1 << J)) continue;
		const V x = v[r], y = v[r | (1 << J)];
		if(CHECKED && is_zero(x) && is_zero(y)) continue;
		V o0, o1;
		o0.x = g00r * x.x - g00i * x.y + g01r * y.x - g01i * y.y;
		o0.y = g00r * x.y + g00i * x.x + g01r * y.y + g01i * y.x;
		o1.x = g10r * x.x - g10i * x.y + g11r * y.x - g11i * y.y;
		o1.y = g10r * x.y + g10i * x.x + g11r * y.y + g11i * y.x;
		v[r] = o0;
		v[r | (1 << J)] = o1;
	}
}

template <typename V, typename R, int W, int J>
__device__ __forceinline__ void reg_u1(V *v, const double *g) {
	const R g00r = (R)g[0], g00i = (R)g[1], g01r = (R)g[2], g01i = (R)g[3];
	const R g10r = (R)g[4], g10i = (R)g[5], g11r = (R)g[6], g11i = (R)g[7];
	if(__any(any_zero_pair<V, W, J>(v))) reg_u1_pairs<V, R, W, J, true>(v, g00r, g00i, g01r, g01i, g10r, g10i, g11r, g11i);
	else reg_u1_pairs<V, R, W, J, false>(v, g00r, g00i, g01r, g01i, g10r, g10i, g11r, g11i);
}

// Dispatch on the runtime stage position sj (< W).
template <typename V, typename R, int W>
__device__ __forceinline__ void reg_nondiag(V *v, const TileOp &op, const unsigned *lidx, const double *__restrict__ mats, R hc) {
#define TMFQS_REG_CASE(J)                                                      \
	case J:                                                                    \
		if constexpr(J < W) {                                                  \
			if(op.kind == TOP_H)                                               \
				reg_h<V, R, W, J>(v, hc);                                      \
			else if(op.kind == TOP_X)                                          \
				reg_x<V, W, J>(v, lidx, op.lcond);                             \
			else                                                               \
				reg_u1<V, R, W, J>(v, mats + op.mat);                          \
		}                                                                      \
		break;
	switch(op.sj) {
		TMFQS_REG_CASE(0)
		TMFQS_REG_CASE(1)
		TMFQS_REG_CASE(2)
		TMFQS_REG_CASE(3)
		TMFQS_REG_CASE(4)
		default: break;
	}
#undef TMFQS_REG_CASE
}

// Global index bit of local position b.
__device__ __forceinline__ u64 local_to_global_bit(unsigned b, unsigned C, const u64 *hitab) {
	return b < C ? (1ull << b) : hitab[1u << (b - C)];
}

template <typename V, typename R, int W>
__device__ __forceinline__ void reg_diag(V *v, const TileOp &op, const unsigned *lidx, unsigned lb, const unsigned *sbits, u64 tbase,
                                         u64 lowmask, unsigned C, const u64 *hitab, const TwiddleTerm *__restrict__ terms,
                                         const double2 *__restrict__ ptab) {
	if(op.kind == TOP_CP) {
		// states with every condition bit set, zero amplitudes untouched (a
		// branch per amplitude: where the condition bits are stage bits the
		// whole wavefront skips the states outside it)
		const R cr = (R)op.cr, ci = (R)op.ci;
#pragma unroll
		for(int r = 0; r < (1 << W); ++r) {
			if((lidx[r] & op.lcond) != op.lcond) continue;
			const V x = v[r];
			if(is_zero(x)) continue;
			V o;
			o.x = cr * x.x - ci * x.y;
			o.y = cr * x.y + ci * x.x;
			v[r] = o;
		}
	} else if(op.kind == TOP_LADDER) {
		// F is linear in the stage bits (disjoint bits of g & amask), so the
		// phase is e(F0) times e(F_j) for every set stage bit j.
		const u64 g0 = tbase | (lb & lowmask) | hitab[lb >> C];
		const u64 F0 = (g0 & op.amask) << op.shift;
		if(F0) {
			const double2 p0 = phase_of_range(F0, ptab, op.cj0, op.cj1);
#pragma unroll
			for(int r = 0; r < (1 << W); ++r)
				if((lidx[r] & op.lcond) == op.lcond) v[r] = rot<V>(v[r], p0);
		}
#pragma unroll
		for(int j = 0; j < W; ++j) {
			const u64 Fj = (local_to_global_bit(sbits[j], C, hitab) & op.amask) << op.shift;
			if(!Fj) continue;  // uniform
			const double2 pj = phase_of_range(Fj, ptab, op.cj0, op.cj1);
#pragma unroll
			for(int r = 0; r < (1 << W); ++r)
				if(((r >> j) & 1) && (lidx[r] & op.lcond) == op.lcond) v[r] = rot<V>(v[r], pj);
		}
	} else if(op.kind == TOP_TWIDDLE) {
		const TwiddleTerm *tt = terms + op.twid;
#pragma unroll
		for(int r = 0; r < (1 << W); ++r) {
			const u64 g = tbase | (lidx[r] & lowmask) | hitab[lidx[r] >> C];
			u64 F = 0;
			for(unsigned t = 0; t < op.nterms; ++t)
				if(g & tt[t].cond) F += (g & tt[t].mask) << tt[t].shift;
			if(F == 0) continue;
			v[r] = rot<V>(v[r], phase_of_range(F, ptab, op.cj0, op.cj1));
		}
	}
}

// ------------------------------------------------------------------------
// complex64 five-qubit matrix on an LDS tile with FP32 MFMA.  The tile's
// 2^(T-5) blocks of 32 amplitudes (the matrix's local bits jl, jl[0] its
// index MSB) are the columns of X, and Y = [[Mr, -Mi], [Mi, Mr]] [Xr; Xi]
// is a real 64 x 64 by 64 x 2^(T-5) product in 32 x 32 x 2 MFMA steps with
// FP32 accumulation.  A wavefront owns one half of the output rows (real or
// imaginary parts) and up to four 32-block column tiles; the matrix half it
// multiplies by stays in its registers.  Needs T >= 10 and four wavefronts.
typedef float mfma_f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned block_offset(unsigned c, const unsigned *jl) {
	unsigned off = 0;
#pragma unroll
	for(int i = 0; i < 5; ++i)
		if((c >> (4 - i)) & 1u) off |= 1u << jl[i];
	return off;
}

__device__ void tile_mfma5(float2 *tile, const TileOp &op, unsigned T, const double *__restrict__ mats) {
	const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	const unsigned half = wave & 1u;  // output rows 32 half .. 32 half + 31
	unsigned jl[5], sorted[5];
#pragma unroll
	for(int i = 0; i < 5; ++i) jl[i] = sorted[i] = op.jl[i];
	for(int i = 0; i < 5; ++i)
		for(int j = i + 1; j < 5; ++j)
			if(sorted[j] < sorted[i]) {
				const unsigned t = sorted[i];
				sorted[i] = sorted[j];
				sorted[j] = t;
			}
	// this lane's A operands Mbig[32 half + lane % 32][2 s + lane / 32], s = 0..31,
	// prepared by the planner in FP32 after the op's matrix (at op.twid)
	const unsigned rr = lane & 31u, kh = lane >> 5;
	const float4 *a4 = (const float4 *)(mats + op.twid) + (half * 64u + lane) * 8u;
	float A[32];
#pragma unroll
	for(int i = 0; i < 8; ++i) {
		const float4 v = a4[i];
		A[4 * i] = v.x;
		A[4 * i + 1] = v.y;
		A[4 * i + 2] = v.z;
		A[4 * i + 3] = v.w;
	}
	const unsigned ncol = 1u << (T - 10);  // column tiles of 32 blocks
	mfma_f16v acc[4];
#pragma unroll
	for(int u = 0; u < 4; ++u) {
		const unsigned ct = (wave >> 1) + 2u * (unsigned)u;
		acc[u] = mfma_f16v{0};
		if(ct >= ncol) continue;  // uniform per wavefront
		unsigned base = ct * 32u + rr;  // block index of this lane's B column
#pragma unroll
		for(int i = 0; i < 5; ++i) base = ((base >> sorted[i]) << (sorted[i] + 1)) | (base & ((1u << sorted[i]) - 1u));
		// steps s and s + 16 take the real and the imaginary part of the same
		// amplitude (k and k + 32): one LDS read serves both
#pragma unroll
		for(int s2 = 0; s2 < 16; ++s2) {
			const float2 x = tile[base | block_offset(2u * (unsigned)s2 + kh, jl)];
			acc[u] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s2], x.x, acc[u], 0, 0, 0);
			acc[u] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s2 + 16], x.y, acc[u], 0, 0, 0);
		}
	}
	__syncthreads();  // every wavefront has read its columns
#pragma unroll
	for(int u = 0; u < 4; ++u) {
		const unsigned ct = (wave >> 1) + 2u * (unsigned)u;
		if(ct >= ncol) continue;
		unsigned base = ct * 32u + rr;  // D column = block ct * 32 + lane % 32
#pragma unroll
		for(int i = 0; i < 5; ++i) base = ((base >> sorted[i]) << (sorted[i] + 1)) | (base & ((1u << sorted[i]) - 1u));
#pragma unroll
		for(int v = 0; v < 16; ++v) {
			const unsigned r = 8u * ((unsigned)v >> 2) + 4u * kh + ((unsigned)v & 3u);  // D row (32x32x2 layout)
			float *dst = (float *)&tile[base | block_offset(r, jl)];
			dst[half] = acc[u][v];
		}
	}
}

// complex128 five-qubit matrix on an LDS tile with FP64 MFMA (16 x 16 x 4,
// FP64 accumulation).  As in tile_mfma5, Y = [[Mr, -Mi], [Mi, Mr]] [Xr; Xi]
// is a real 64 x 64 by 64 x 2^(T-5) product; wavefront w owns output rows
// 16 w .. 16 w + 15 (w = 0, 1: real parts of amplitude rows 0..15, 16..31;
// w = 2, 3: imaginary parts) and keeps its 16 A operands (one per K step of
// four rows) in registers.  Columns are blocks of 32 amplitudes taken sixteen
// at a time, four such groups in flight per wavefront so consecutive MFMAs
// do not wait on one accumulator.  K steps s and s + 8 take the real and
// imaginary part of the same amplitude, so one LDS read serves both.
// Operand layout (v_mfma_f64_16x16x4f64): lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and receives D[l / 16 + 4 v][l % 16], v = 0..3.
// Needs T >= 11 (at least four groups of sixteen blocks) and four wavefronts.
// Not the dense backend's accumulation order: results agree with it to
// rounding (the exact mode keeps tile_uk instead).
typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));

__device__ void tile_mfma5_f64(double2 *tile, const TileOp &op, unsigned T, const double *__restrict__ mats) {
	const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	unsigned jl[5], sorted[5];
#pragma unroll
	for(int i = 0; i < 5; ++i) jl[i] = sorted[i] = op.jl[i];
	for(int i = 0; i < 5; ++i)
		for(int j = i + 1; j < 5; ++j)
			if(sorted[j] < sorted[i]) {
				const unsigned t = sorted[i];
				sorted[i] = sorted[j];
				sorted[j] = t;
			}
	// A[s] = Mbig[16 wave + lane % 16][4 s + lane / 16], prepared by the planner
	const double2 *a2 = (const double2 *)(mats + op.twid) + (wave * 64u + lane) * 8u;
	double A[16];
#pragma unroll
	for(int i = 0; i < 8; ++i) {
		const double2 v = a2[i];
		A[2 * i] = v.x;
		A[2 * i + 1] = v.y;
	}
	const unsigned col = lane & 15u, kq = lane >> 4;
	const unsigned part = wave >> 1, rbase = 16u * (wave & 1u);
	const unsigned ngroups = 1u << (T - 9);  // groups of 16 blocks of 32 amplitudes
	unsigned roff[8], woff[4];
#pragma unroll
	for(int s2 = 0; s2 < 8; ++s2) roff[s2] = block_offset(4u * (unsigned)s2 + kq, jl);
#pragma unroll
	for(int v = 0; v < 4; ++v) woff[v] = block_offset(rbase + kq + 4u * (unsigned)v, jl);
	for(unsigned g0 = 0; g0 < ngroups; g0 += 4) {
		unsigned base[4];
		mfma_f64x4 acc[4];
#pragma unroll
		for(int u = 0; u < 4; ++u) {
			unsigned b = (g0 + (unsigned)u) * 16u + col;  // this lane's B column (block index)
#pragma unroll
			for(int i = 0; i < 5; ++i) b = ((b >> sorted[i]) << (sorted[i] + 1)) | (b & ((1u << sorted[i]) - 1u));
			base[u] = b;
			acc[u] = mfma_f64x4{0.0, 0.0, 0.0, 0.0};
		}
#pragma unroll
		for(int s2 = 0; s2 < 8; ++s2) {
			double2 x[4];
#pragma unroll
			for(int u = 0; u < 4; ++u) x[u] = tile[base[u] | roff[s2]];
#pragma unroll
			for(int u = 0; u < 4; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[s2], x[u].x, acc[u], 0, 0, 0);
#pragma unroll
			for(int u = 0; u < 4; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[s2 + 8], x[u].y, acc[u], 0, 0, 0);
		}
		__syncthreads();  // every wavefront has read these columns
#pragma unroll
		for(int u = 0; u < 4; ++u)
#pragma unroll
			for(int v = 0; v < 4; ++v) ((double *)&tile[base[u] | woff[v]])[part] = acc[u][v];
		__syncthreads();
	}
}

// One tile of a staged pass: loads it into LDS, runs every stage, stores it.
// ta/tb/hitab/gacc are the workgroup's LDS buffers; hitab must hold P's table.
template <typename R, int W, bool DUAL, int LDS>
__device__ __forceinline__ void staged_tile(typename V2T<R>::t *__restrict__ a, typename V2T<R>::t *__restrict__ b,
                                            const TilePass &P, u64 tid, const TileOp *__restrict__ ops,
                                            const TileStage *__restrict__ stages, unsigned nstages,
                                            const double *__restrict__ mats, const TwiddleTerm *__restrict__ terms,
                                            const double2 *__restrict__ ptab, typename V2T<R>::t *ta,
                                            typename V2T<R>::t *tb, const u64 *hitab, double *gacc) {
	using V = typename V2T<R>::t;
	constexpr int NR = 1 << W;
	constexpr bool PAD = LDS == 0;
	const unsigned nloc = 1u << P.T;
	const unsigned C = P.C;
	const u64 lowmask = (1ull << C) - 1ull;
	const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	const R hc = (R)(1.0 / sqrt(2.0));
	const unsigned ngrp = nloc >> W;
	const unsigned iters = (ngrp + blockDim.x - 1) / blockDim.x;
	u64 tbase = tid << C;
	for(unsigned bb = 0; bb < P.nh; ++bb) tbase = ins0(tbase, P.hbits[bb]);
	__syncthreads();
	// Loads in groups of TU per thread, all issued before the first is used,
	// so a thread has TU (dual: 2 TU) loads in flight instead of one.
	constexpr unsigned TU = DUAL ? 8u : 16u;
	for(unsigned l0 = threadIdx.x; l0 < nloc; l0 += blockDim.x * TU) {
		V xa[TU], xb[TU];
#pragma unroll
		for(unsigned u = 0; u < TU; ++u) {
			const unsigned l = l0 + u * blockDim.x;
			if(l < nloc) {
				const u64 g = tbase | (l & lowmask) | hitab[l >> C];
				xa[u] = nt_load(&a[g]);
				if(DUAL) xb[u] = nt_load(&b[g]);
			}
		}
#pragma unroll
		for(unsigned u = 0; u < TU; ++u) {
			const unsigned l = l0 + u * blockDim.x;
			if(l < nloc) {
				const unsigned pl = PAD ? l + (l >> 4) : l;
				ta[pl] = xa[u];
				if(DUAL) tb[pl] = xb[u];
			}
		}
	}
	__syncthreads();
	for(unsigned s = 0; s < nstages; ++s) {
		const TileStage st = stages[s];
		if(LDS && st.lds) {  // k-qubit matrices: their registers would cap every other stage's occupancy
			for(unsigned o = 0; o < st.nops; ++o) {
				const TileOp op = ops[st.op0 + o];
				if((tbase & op.gcond) != op.gcond) continue;  // uniform across the block
				if constexpr(sizeof(R) == 4) {
					if(op.kind == TOP_UK && op.k == 5 && P.T >= 10 && blockDim.x == 256) {
						tile_mfma5((float2 *)ta, op, P.T, mats);
						if(DUAL) {
							__syncthreads();
							tile_mfma5((float2 *)tb, op, P.T, mats);
						}
						__syncthreads();
						continue;
					}
				} else if constexpr(LDS == 2) {
					// the planner picks this build only for tiles of >= 11 bits whose
					// five-qubit blocks all carry MFMA operands
					if(op.kind == TOP_UK && op.k == 5) {
						tile_mfma5_f64((double2 *)ta, op, P.T, mats);
						if(DUAL) tile_mfma5_f64((double2 *)tb, op, P.T, mats);
						continue;
					}
				}
				tile_apply<R, LDS == 2 ? 4 : 5>(ta, op, nloc, tbase, C, hitab, mats, terms, ptab);
				if(DUAL) tile_apply<R, LDS == 2 ? 4 : 5>(tb, op, nloc, tbase, C, hitab, mats, terms, ptab);
				__syncthreads();
			}
			continue;
		}
		unsigned sbits[W];
#pragma unroll
		for(int j = 0; j < W; ++j) sbits[j] = st.bits[j];
		for(unsigned it = 0; it < iters; ++it) {
			const unsigned grp = threadIdx.x + it * blockDim.x;
			const bool act = grp < ngrp;
			unsigned lb = act ? grp : 0u;
#pragma unroll
			for(int j = 0; j < W; ++j) lb = ((lb >> sbits[j]) << (sbits[j] + 1)) | (lb & ((1u << sbits[j]) - 1u));
			unsigned lidx[NR];
			V va[NR], vb[NR];
#pragma unroll
			for(int r = 0; r < NR; ++r) {
				unsigned off = 0;
#pragma unroll
				for(int j = 0; j < W; ++j)
					if((r >> j) & 1) off |= 1u << sbits[j];
				lidx[r] = lb | off;
				if(act) {
					const unsigned pl = PAD ? lidx[r] + (lidx[r] >> 4) : lidx[r];
					va[r] = ta[pl];
					if(DUAL) vb[r] = tb[pl];
				}
			}
			// complex128: Im(conj(b) a) of each register amplitude, shared by
			// consecutive gradient ops (a run's reductions come together) until
			// an op changes the amplitudes (complex64 recomputes it per op: the
			// cache would cost its three-bit dual build an occupancy step)
			constexpr bool GCACHE = DUAL && sizeof(R) == 8;
			double gt[GCACHE ? NR : 1];
			bool gvalid = false;
			for(unsigned o = 0; o < st.nops; ++o) {
				const TileOp op = ops[st.op0 + o];
				if((tbase & op.gcond) != op.gcond) continue;  // uniform across the block
				if(op.kind == TOP_GRAD) {
					if constexpr(DUAL) {
						double acc = 0.0;
						if(act) {
							if constexpr(GCACHE) {
								if(!gvalid) {
#pragma unroll
									for(int r = 0; r < NR; ++r)
										gt[r] = (double)vb[r].x * (double)va[r].y - (double)vb[r].y * (double)va[r].x;
									gvalid = true;
								}
#pragma unroll
								for(int r = 0; r < NR; ++r)
									if((lidx[r] & op.lcond) == op.lcond) acc += gt[r];
							} else {
#pragma unroll
								for(int r = 0; r < NR; ++r)
									if((lidx[r] & op.lcond) == op.lcond)
										acc += (double)vb[r].x * (double)va[r].y - (double)vb[r].y * (double)va[r].x;
							}
						}
						acc = wave_sum_dpp(acc);  // one per gradient op and tile: DPP, not LDS permutes
						if(lane == 0) gacc[wave * kMaxGradSlots + op.mat] += acc;
					}
					continue;
				}
				if(!act) continue;
				gvalid = false;
				if constexpr(DUAL && sizeof(R) == 8)  // the planner emits phase runs for complex128 dual passes only
					if(op.kind == TOP_PHASES) {
					// one angle sum and one sincos per amplitude, shared by both states
#pragma unroll
					for(int r = 0; r < NR; ++r) {
						const u64 g = tbase | (lidx[r] & lowmask) | hitab[lidx[r] >> C];
						double th = 0.0;  // never -0, so adding +0 leaves it unchanged
						for(unsigned t = 0; t < op.nterms; ++t) {
							const TwiddleTerm w = terms[op.twid + t];
							th += ((g & w.cond) == w.cond) ? __longlong_as_double((long long)w.mask) : 0.0;
						}
						double2 f;
						sincos(th, &f.y, &f.x);
						va[r] = rot<V>(va[r], f);
						vb[r] = rot<V>(vb[r], f);
					}
					continue;
					}
				if(op.kind == TOP_H || op.kind == TOP_X || op.kind == TOP_U1) {
					reg_nondiag<V, R, W>(va, op, lidx, mats, hc);
					if(DUAL) reg_nondiag<V, R, W>(vb, op, lidx, mats, hc);
				} else {
					reg_diag<V, R, W>(va, op, lidx, lb, sbits, tbase, lowmask, C, hitab, terms, ptab);
					if(DUAL) reg_diag<V, R, W>(vb, op, lidx, lb, sbits, tbase, lowmask, C, hitab, terms, ptab);
				}
			}
			if(act) {
#pragma unroll
				for(int r = 0; r < NR; ++r) {
					const unsigned pl = PAD ? lidx[r] + (lidx[r] >> 4) : lidx[r];
					ta[pl] = va[r];
					if(DUAL) tb[pl] = vb[r];
				}
			}
		}
		__syncthreads();
	}
	for(unsigned l0 = threadIdx.x; l0 < nloc; l0 += blockDim.x * TU) {
#pragma unroll
		for(unsigned u = 0; u < TU; ++u) {
			const unsigned l = l0 + u * blockDim.x;
			if(l < nloc) {
				const u64 g = tbase | (l & lowmask) | hitab[l >> C];
				const unsigned pl = PAD ? l + (l >> 4) : l;
				nt_store(ta[pl], &a[g]);
				if(DUAL) nt_store(tb[pl], &b[g]);
			}
		}
	}
}

// LDS elements of one staged tile of 2^T amplitudes (one pad slot per 16
// amplitudes keeps register-stage gathers over low bit positions, strided
// across lanes, off a single bank; LDS-op stages index the tile directly).
__host__ __device__ __forceinline__ unsigned staged_tile_elems(unsigned T, bool ldsops) {
	return ldsops ? (1u << T) : (1u << T) + ((1u << T) >> 4);
}

// TH lanes per workgroup: 256, or 512 for register-stage builds whose stages
// hold 2^(T-9) amplitudes per lane (twice the wavefronts per CU at the same LDS
// tile, with one-bit-narrower stages).
template <typename R, int W, bool DUAL, int LDS, int TH>
__global__ void __launch_bounds__(TH) k_tile_staged(typename V2T<R>::t *__restrict__ a, typename V2T<R>::t *__restrict__ b,
                                                    TilePass P, const TileOp *__restrict__ ops,
                                                    const TileStage *__restrict__ stages, unsigned nstages,
                                                    const double *__restrict__ mats, const TwiddleTerm *__restrict__ terms,
                                                    const double2 *__restrict__ ptab, double *__restrict__ gpart,
                                                    unsigned nslots) {
	using V = typename V2T<R>::t;
	constexpr unsigned NW = TH / 64;
	extern __shared__ char smem[];
	const unsigned nlocp = staged_tile_elems(P.T, LDS != 0);
	V *ta = (V *)smem;
	V *tb = ta + nlocp;  // second tile (dual passes only)
	u64 *hitab = (u64 *)(ta + (DUAL ? 2u : 1u) * nlocp);
	double *gacc = (double *)(hitab + (1u << P.nh));  // dual: [NW wavefronts][kMaxGradSlots]
	tile_hitab<R>(hitab, P);
	if(DUAL)
		for(unsigned i = threadIdx.x; i < NW * kMaxGradSlots; i += blockDim.x) gacc[i] = 0.0;
	for(u64 tid = blockIdx.x; tid < P.ntiles; tid += gridDim.x)
		staged_tile<R, W, DUAL, LDS>(a, b, P, tid, ops, stages, nstages, mats, terms, ptab, ta, tb, hitab, gacc);
	if(DUAL) {
		__syncthreads();
		for(unsigned s = threadIdx.x; s < nslots; s += blockDim.x) {
			double sum = gacc[s];
#pragma unroll
			for(unsigned w = 1; w < NW; ++w) sum += gacc[w * kMaxGradSlots + s];  // fixed order
			gpart[(u64)s * gridDim.x + blockIdx.x] = sum;
		}
	}
}

template <typename R, int W, bool DUAL, int LDS, int TH>
static hipError_t launch_staged_th(const Launch &L, void *a, void *b, const TilePass &P, const TileOp *dops,
                                   const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                                   const double *dphase, double *gpart, unsigned nslots) {
	using V = typename V2T<R>::t;
	const size_t nlocp = staged_tile_elems(P.T, LDS != 0);
	const size_t shmem =
	    (DUAL ? 2 : 1) * sizeof(V) * nlocp + (8ull << P.nh) + (DUAL ? (TH / 64) * kMaxGradSlots * sizeof(double) : 0);
	if(shmem > 160 * 1024) return hipErrorInvalidValue;
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_tile_staged<R, W, DUAL, LDS, TH>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                          160 * 1024);
		attr = true;
	}
	TMFQS_LAUNCH((k_tile_staged<R, W, DUAL, LDS, TH>), tile_grid(P, DUAL, L.prec), TH, shmem, L.stream, (V *)a, (V *)b, P, dops,
	             dstages, nstages, dmats, dterms, (const double2 *)dphase, gpart, nslots);
	return hipGetLastError();
}

template <typename R, int W, bool DUAL, int LDS>
static hipError_t launch_staged_t(const Launch &L, void *a, void *b, const TilePass &P, const TileOp *dops, const TileStage *dstages,
                                  unsigned nstages, const double *dmats, const TwiddleTerm *dterms, const double *dphase,
                                  double *gpart, unsigned nslots) {
	if constexpr(LDS == 0 && W <= 3)
		if(P.T >= 9 + (unsigned)W)  // 512 lanes x 2^W amplitudes cover the tile
			return launch_staged_th<R, W, DUAL, LDS, 512>(L, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
	return launch_staged_th<R, W, DUAL, LDS, 256>(L, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

template <typename R, bool DUAL, int LDS>
static hipError_t launch_staged_w(const Launch &L, unsigned W, void *a, void *b, const TilePass &P, const TileOp *dops,
                                  const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                                  const double *dphase, double *gpart, unsigned nslots) {
#define TMFQS_STAGED(WW) \
	launch_staged_t<R, WW, DUAL, LDS>(L, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots)
	switch(W) {
		case 1: return TMFQS_STAGED(1);
		case 2: return TMFQS_STAGED(2);
		case 3: return TMFQS_STAGED(3);
		case 4: return TMFQS_STAGED(4);
		default: return hipErrorInvalidValue;
	}
#undef TMFQS_STAGED
}

// One translation unit per (precision, dual, LDS ops) instantiates the tiled
// kernels, so the library builds in parallel.
#define TMFQS_TILE_ENTRY(NAME)                                                                                              \
	hipError_t NAME(const Launch &L, unsigned W, void *a, void *b, const TilePass &P, const TileOp *dops,                   \
	                const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,             \
	                const double *dphase, double *gpart, unsigned nslots)
TMFQS_TILE_ENTRY(launch_tile_c128);
TMFQS_TILE_ENTRY(launch_tile_c128_lds);
TMFQS_TILE_ENTRY(launch_tile_c128_dual);
TMFQS_TILE_ENTRY(launch_tile_c128_dual_lds);
TMFQS_TILE_ENTRY(launch_tile_c128_mfma);
TMFQS_TILE_ENTRY(launch_tile_c128_dual_mfma);
TMFQS_TILE_ENTRY(launch_tile_c64);
TMFQS_TILE_ENTRY(launch_tile_c64_lds);
TMFQS_TILE_ENTRY(launch_tile_c64_dual);
TMFQS_TILE_ENTRY(launch_tile_c64_dual_lds);

}  // namespace tmfqs_hip
