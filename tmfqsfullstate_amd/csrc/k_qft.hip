// gfx950 QFT kernels: eight-target block passes and the six-target tail fused
// with the bit reversal.
#include "device_common.h"

#include <algorithm>
#include <type_traits>

namespace tmfqs_hip {

// ------------------------------------------------------------------------
// QFT block pass.  Targets are the 8 bit positions pbot+7 .. pbot (MSB first,
// as makeQft orders its target qubits); for each target t it applies H on t
// and then every controlled phase pi/2^(t-c) from the lower bits c < t, i.e.
// the factor e^{i pi (x mod 2^t) / 2^t} on states with bit t set.  Splitting
// x mod 2^t into the part inside the block and the part r below it, the inner
// factors are constants c[j][v] = e^{i pi v / 2^j} and the outer ones commute
// with the remaining H gates of the block, so they are applied last as one
// twiddle per state: block bit j set contributes e^{i pi r / 2^(pb+j)}, so a
// state with block bits y takes w^(r rev8(y)), w = e^{2 pi i / 2^(pb+8)}.
//
// A workgroup owns 2^(8+C) states: the 8 block bits times the C contiguous
// low bits (coalesced 2^C-state runs).  Each thread holds 16 states in
// registers: first the 16 values of block bits 7..4 (stage A), then, after an
// LDS transpose, the 16 values of block bits 3..0 (stage B).  Butterflies are
// unscaled; the eight Hadamard factors 1/sqrt(2) are one exact scale by 1/16,
// folded into the twiddles.  The stage-A loads of the workgroup's next tile
// are issued as soon as the current tile is in LDS, so they are in flight
// during stage B and the stores.
struct QftBlockParams {
	u64 ntiles;  // 2^(n - 8 - C)
	uint32_t n, pbot;
	uint32_t K;    // targets: the top K block bits (8, or fewer in a partial pass)
	uint32_t pad;
	double scale;  // (1/sqrt(2))^K
};


// Waves per SIMD the QFT block kernel is compiled for, and whether a
// workgroup prefetches its next tile during stage B.  complex128 (two waves
// per SIMD) prefetches; complex64 keeps to 128 registers (four waves per
// SIMD) by not prefetching, and other workgroups' loads overlap instead.
template <typename R, int C> struct QftBlockWaves {
	static constexpr int value = sizeof(R) == 8 ? 2 : 4;
	static constexpr bool prefetch = sizeof(R) == 8;
};

// Stage-A loads of tile `tid`: register ra holds block bits 7..4 = ra.
template <typename V, int C>
__device__ __forceinline__ void qft_block_load(V *v, const V *__restrict__ a, u64 tid, unsigned pb, unsigned f4, unsigned lo) {
	const u64 x = tid << C;
	const V *src = a + ((((x >> pb) << (pb + 8)) | (x & ((1ull << pb) - 1ull))) + ((u64)f4 << pb) + lo);
#pragma unroll
	for(int ra = 0; ra < 16; ++ra) v[ra] = nt_load(src + ((u64)ra << (pb + 4)));
}

// PART: a partial pass whose targets are only the top K < 8 block bits (the
// leading (n - 6) mod 8 targets of a QFT); the other block bits are ladder
// controls only.
template <typename R, int C, bool PART>
__global__ void __launch_bounds__(16 << C) __attribute__((amdgpu_waves_per_eu(QftBlockWaves<R, C>::value)))
k_qft_block(typename V2T<R>::t *__restrict__ a, QftBlockParams P, const double2 *__restrict__ ptab,
            const double2 *__restrict__ cin, const typename LanePhase<typename V2T<R>::t>::t *__restrict__ cinp) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	V *tile = (V *)smem;
	const unsigned t = threadIdx.x;
	const unsigned lo = t & ((1u << C) - 1u);
	const unsigned f4 = t >> C;  // stage A: block bits 3..0; stage B: block bits 7..4
	const unsigned pb = P.pbot;
	const u64 lowmask = (1ull << pb) - 1ull;
	// Per-thread constants: the stage-A ladder factors of target 4 + b on
	// register bits low below it, c[b][low] e^{i pi f4 / 2^(4+b)} with the
	// in-block constants c[b][v] = cin[2^b - 1 + v] = e^{i pi v / 2^b}, and the
	// lane's part w^lo of the twiddle base.  Stage B's in-block constants are
	// read where they are used (uniform addresses: scalar loads), from cinp:
	// the same table in the amplitudes' precision.
	typename LanePhase<V>::t pa[15];
#pragma unroll
	for(int b = 0; b < 4; ++b) {
		const double2 tp = ptab[(f4 << (3 - b)) & 255u];  // e^{i pi f4 / 2^(4+b)}
#pragma unroll
		for(int low = 0; low < (1 << b); ++low) pa[(1 << b) - 1 + low] = LanePhase<V>::of(cmul(cin[(1 << b) - 1 + low], tp));
	}
	const double2 zlo = phase_of_range((u64)lo << (56 - pb), ptab, 0, 4);
	u64 tid = blockIdx.x;
	if(tid >= P.ntiles) return;  // uniform
	V v[16];
	qft_block_load<V, C>(v, a, tid, pb, f4, lo);
	for(;;) {
		// ---- stage A: registers index block bits 7..4
#pragma unroll
		for(int j = 7; j >= 4; --j) {
			const int b = j - 4;
			if constexpr(PART)
				if(j < 8 - (int)P.K) continue;  // uniform
#pragma unroll
			for(int ra = 0; ra < 16; ++ra)
				if(!(ra & (1 << b))) butterfly(v[ra], v[ra | (1 << b)]);
#pragma unroll
			for(int ra = 0; ra < 16; ++ra) {
				if(!(ra & (1 << b))) continue;
				v[ra] = rotp(v[ra], pa[(1 << b) - 1 + (ra & ((1 << b) - 1))]);
			}
		}
		__syncthreads();  // previous tile's stage-B reads are complete
#pragma unroll
		for(int ra = 0; ra < 16; ++ra) {
			const unsigned l = ((((unsigned)ra << 4) | f4) << C) | lo;
			tile[l + (l >> (4 + C))] = v[ra];
		}
		__syncthreads();
		const u64 nxt = tid + gridDim.x;
		const bool more = nxt < P.ntiles;  // uniform
		if constexpr(QftBlockWaves<R, C>::prefetch)
			if(more) qft_block_load<V, C>(v, a, nxt, pb, f4, lo);
		// ---- stage B: registers index block bits 3..0
		V w[16];
#pragma unroll
		for(int rb = 0; rb < 16; ++rb) {
			const unsigned l = (((f4 << 4) | (unsigned)rb) << C) | lo;
			w[rb] = tile[l + (l >> (4 + C))];
		}
#pragma unroll
		for(int j = 3; j >= 0; --j) {
			if constexpr(PART)
				if(j < 8 - (int)P.K) continue;  // uniform
#pragma unroll
			for(int rb = 0; rb < 16; ++rb)
				if(!(rb & (1 << j))) butterfly(w[rb], w[rb | (1 << j)]);
			if(j == 0) continue;  // c[0][0] = 1
#pragma unroll
			for(int rb = 0; rb < 16; ++rb) {
				if(!(rb & (1 << j))) continue;
				const int low = rb & ((1 << j) - 1);
				if(low == 0) continue;
				w[rb] = rotp(w[rb], cinp[(1 << j) - 1 + low]);
			}
		}
		// ---- twiddle w^(r rev8(y)) / 16 for y = (f4, rb), r = rhi + lo:
		// rev8(y) = 16 rev4(rb) + rev4(f4), so register rb takes
		// base u^rev4(rb) with base = w^(r rev4(f4)) / 16 and u = w^(16 r).
		const u64 x = tid << C;
		const u64 rhi = x & lowmask;
		double2 z[4];
		z[0] = cmul(phase_of_range(rhi << (56 - pb), ptab, 0, 4), zlo);  // w^r
#pragma unroll
		for(int k = 1; k < 4; ++k) z[k] = csqr(z[k - 1]);
		const double2 u = csqr(z[3]);
		// A partial pass keeps the target bits of rev8(y) only (its low K bits):
		// f4 bit b counts when 3 - b < K, and register rb takes u^(rev4(rb) & m2).
		double2 ph = make_double2(PART ? P.scale : 0.0625, 0.0);
#pragma unroll
		for(int b = 0; b < 4; ++b)
			if(((f4 >> b) & 1u) && (!PART || 3 - b < (int)P.K)) ph = cmul(ph, z[3 - b]);
		const unsigned m2 = PART ? (P.K > 4 ? (1u << (P.K - 4)) - 1u : 0u) : 15u;
		const double2 ph0 = ph;
#pragma unroll
		for(int k = 0; k < 16; ++k) {
			const int rb = (int)(__brev((unsigned)k) >> 28);
			if(PART && k && !(k & m2)) ph = ph0;
			w[rb] = rot<V>(w[rb], ph);
			if(k < 15) ph = cmul(ph, u);
		}
		const u64 base = ((x >> pb) << (pb + 8)) | rhi;
		V *dst = a + (base + ((u64)f4 << (pb + 4)) + lo);
#pragma unroll
		for(int rb = 0; rb < 16; ++rb) nt_store(w[rb], dst + ((u64)rb << pb));
		if(!more) break;
		if constexpr(!QftBlockWaves<R, C>::prefetch) qft_block_load<V, C>(v, a, nxt, pb, f4, lo);
		tid = nxt;
	}
}

template <typename R, int C, bool PART>
static hipError_t launch_qft_block_t(const Launch &L, void *a, unsigned n, unsigned pbot, unsigned K, const double *dphase,
                                     const double *dcin) {
	using V = typename V2T<R>::t;
	QftBlockParams P;
	P.n = n;
	P.pbot = pbot;
	P.K = K;
	P.pad = 0;
	P.scale = std::ldexp(K & 1u ? 1.0 / std::sqrt(2.0) : 1.0, -(int)(K / 2));  // (1/sqrt(2))^K, one rounding at most
	P.ntiles = 1ull << (n - 8 - C);
	// complex128: a grid of 1024 workgroups that loop over their tiles, so each
	// next tile's loads overlap stage B (3% faster at n = 30 than one tile per
	// workgroup); complex64 (no prefetch): at most 65536 workgroups, two tiles
	// each at n = 30 (3% faster than one tile per workgroup).
	const u64 cap = QftBlockWaves<R, C>::prefetch ? 1024 : 65536;
	const unsigned grid = (unsigned)(P.ntiles < cap ? P.ntiles : cap);
	const size_t shmem = (size_t)((256u << C) + 16u) * sizeof(V);
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_qft_block<R, C, PART>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	using PT = typename LanePhase<V>::t;
	const PT *cinp = sizeof(R) == 8 ? (const PT *)dcin : (const PT *)(dcin + 126);  // FP32 copy after the 63 FP64 constants
	hipLaunchKernelGGL((k_qft_block<R, C, PART>), dim3(grid), dim3(16 << C), shmem, L.stream, (V *)a, P,
	                   (const double2 *)dphase, (const double2 *)dcin, cinp);
	return hipGetLastError();
}

template <typename R, int C>
static hipError_t launch_qft_block_k(const Launch &L, void *a, unsigned n, unsigned pbot, unsigned K, const double *dphase,
                                     const double *dcin) {
	return K == 8 ? launch_qft_block_t<R, C, false>(L, a, n, pbot, 8, dphase, dcin)
	              : launch_qft_block_t<R, C, true>(L, a, n, pbot, K, dphase, dcin);
}

hipError_t launch_qft_block(const Launch &L, void *a, unsigned n, unsigned pbot, const double *dphase, const double *dcin,
                            unsigned C, unsigned K) {
	if(pbot < C || pbot + 8 > n || K < 1 || K > 8) return hipErrorInvalidValue;
	if(L.prec == 0) {
		if(C == 4) return launch_qft_block_k<double, 4>(L, a, n, pbot, K, dphase, dcin);
		if(C == 5) return launch_qft_block_k<double, 5>(L, a, n, pbot, K, dphase, dcin);
	} else {
		if(C == 5) return launch_qft_block_k<float, 5>(L, a, n, pbot, K, dphase, dcin);
		if(C == 6) return launch_qft_block_k<float, 6>(L, a, n, pbot, K, dphase, dcin);
	}
	return hipErrorInvalidValue;
}

unsigned qft_block_coalesce_bits(int prec, int requested) {
	if(prec == 0) return requested == 4 ? 4u : 5u;  // 5: 512 B runs per row (measured faster)
	return requested == 6 ? 6u : 5u;
}

// ------------------------------------------------------------------------
// QFT tail fused with the final bit reversal (makeQft's swap network).
// The last six targets 5..0 only mix the low six bits, so the tail is a
// 64-point transform of every contiguous run of 64 amplitudes; the reversal
// then sends index (H, M, Lo) -> (rev6(Lo), rev(M), rev6(H)) for the top six
// bits H, middle bits M and low six bits Lo.  Tile M = {*, M, *} (64 rows of
// 64 amplitudes) goes, transposed, into the place of tile rev(M) and tile
// rev(M) into the place of tile M: the work is a list of tile pairs
// {M, rev(M)} exchanged in place.
//
// Pairs are enumerated directly, with no index skipped.  Split the m = n - 12
// middle bits into a top half x (h = m / 2 bits), a middle bit c when m is
// odd, and a bottom half y.  With u = x and w = rev_h(y), M = (u, c, rev_h(w))
// and rev(M) = (w, c, rev_h(u)): the reversal swaps u and w, so the pairs are
// the entries u <= w of a K x K triangle (K = 2^h) per value of c, u == w
// being the palindromes.  Rows r and K-1-r of the triangle hold K+1 entries
// together, which folds it into a (K/2) x (K+1) rectangle.
struct QftTailParams {
	u64 npairs;
	uint32_t n, m, h, K;
};

__device__ __forceinline__ void tail_pair(const QftTailParams &P, u64 pair, u64 &M, u64 &Mr) {
	unsigned p = (unsigned)pair, c = 0;  // pair < 2^29 for n <= 40
	if(P.m & 1u) {
		c = p & 1u;
		p >>= 1;
	}
	unsigned u = 0, w = 0;
	if(P.K > 1) {
		const unsigned r = p / (P.K + 1), col = p % (P.K + 1);
		if(col < P.K - r) {
			u = r;
			w = r + col;
		} else {
			u = P.K - 1 - r;
			w = u + (col - (P.K - r));
		}
	}
	const unsigned top = P.m - P.h;
	const u64 revu = P.h ? (__brev(u) >> (32 - P.h)) : 0u, revw = P.h ? (__brev(w) >> (32 - P.h)) : 0u;
	M = ((u64)u << top) | ((u64)c << P.h) | revw;
	Mr = ((u64)w << top) | ((u64)c << P.h) | revu;
}

__device__ __forceinline__ unsigned rev6(unsigned x) { return __brev(x) >> 26; }
// Lane layout of a 64 x 64 tile: LB column bits in lanes (a row is one
// 128 B line per load instruction: 8 lanes of complex128, 16 of complex64),
// the other 6 - LB column bits in registers, and each lane holds 16 / 2^(6-LB)
// rows spaced 256 / 2^LB apart, 16 amplitudes in all.
template <typename V> struct TailLayout {
	static constexpr unsigned LB = sizeof(V) == 16 ? 3u : 4u;  // lane (column) bits
	static constexpr unsigned RB = 6u - LB;                    // register (column) bits
	static constexpr unsigned LPR = 1u << LB;                  // lanes per row
	static constexpr unsigned RPL = 16u >> RB;                 // rows per lane
	static constexpr unsigned RSTEP = 256u >> LB;              // row spacing
};

// LDS slot of (row, col) in a 64 x 64 tile: the column's lane bits are XORed
// with row bits 5..6-LB, which spreads the transposed reads (rows rev6(c) for
// 2^LB consecutive c: one 128 B bank sweep) over the banks; the row-wise
// writes (2^LB lanes per row) are spread by the column itself.  The
// register bits of the column are left alone, so a lane's slots in a row are
// constant offsets from one address.
template <typename V> __device__ __forceinline__ unsigned tail_slot(unsigned row, unsigned col) {
	constexpr unsigned LB = TailLayout<V>::LB;
	return row * 64u + (col ^ ((row >> (6u - LB)) & ((1u << LB) - 1u)));
}

// Value of lane (lane ^ MASK) within each row of 16 lanes, by DPP moves:
// quad_perm [1,0,3,2] and [2,3,0,1] for MASK 1 and 2; for MASK 4,
// row_half_mirror (lane ^ 7) followed by quad_perm [3,2,1,0] (lane ^ 3); for
// MASK 8, row_ror:8.
template <unsigned MASK> __device__ __forceinline__ int dpp_xor(int x) {
	if constexpr(MASK == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
	else if constexpr(MASK == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
	else if constexpr(MASK == 8) return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);  // row_ror:8
	else {
		static_assert(MASK == 4, "lane masks 1, 2, 4, 8");
		return __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
	}
}
template <unsigned MASK> __device__ __forceinline__ double dpp_xor(double x) {
	const long long b = __double_as_longlong(x);
	const int lo = dpp_xor<MASK>((int)(unsigned)b), hi = dpp_xor<MASK>((int)(unsigned)(b >> 32));
	return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <unsigned MASK> __device__ __forceinline__ float dpp_xor(float x) {
	return __int_as_float(dpp_xor<MASK>(__float_as_int(x)));
}

// Unscaled butterfly over the target bit MASK of the lane index: lanes with
// the bit clear keep x + y, their partners y - x, computed as y + (+-x), which
// rounds identically (the sign flip is exact).
__device__ __forceinline__ double flip_if(double x, unsigned f) {
	return __longlong_as_double(__double_as_longlong(x) ^ ((long long)f << 63));
}
__device__ __forceinline__ float flip_if(float x, unsigned f) { return __int_as_float(__float_as_int(x) ^ (int)(f << 31)); }

template <unsigned MASK, typename V> __device__ __forceinline__ V lane_bfly(V x, unsigned lane) {
	const unsigned f = (lane & MASK) ? 1u : 0u;
	return mk<V>(dpp_xor<MASK>(x.x) + flip_if(x.x, f), dpp_xor<MASK>(x.y) + flip_if(x.y, f));
}

// The 64-point tail transform of one row held by 2^LB lanes: registers are
// column bits 5..LB, the lane's position q in its row is column bits LB-1..0.
// The ladder factor e^{i pi low / 2^t} of a register target t splits into a
// constant of the register index (uniform: scalar loads) and the lane's
// qf[t - LB] = e^{i pi q / 2^t}; the lane targets LB-1..0 are DPP
// butterflies, and target t > 0 then rotates its partners by
// e^{i pi (q mod 2^t) / 2^t} (qf[RB + LB - 1 - t]; 1 on the other lanes, so
// every lane rotates and nothing is selected).  The six Hadamard factors
// 1/sqrt(2) are one exact scale by 1/8.
template <typename V, unsigned MASK> __device__ __forceinline__ V tail_lane_step(V x, unsigned lane, const double2 *qf) {
	constexpr unsigned LB = TailLayout<V>::LB, RB = TailLayout<V>::RB;
	x = lane_bfly<MASK>(x, lane);
	if constexpr(MASK > 1) x = rot<V>(x, qf[RB + LB - 1 - (31 - __builtin_clz(MASK))]);
	return x;
}

template <typename V>
__device__ __forceinline__ void tail_transform(V *v, unsigned lane, const double2 *qf, const double2 *__restrict__ cq) {
	constexpr int LB = (int)TailLayout<V>::LB, RB = (int)TailLayout<V>::RB, NR = 1 << RB;
#pragma unroll
	for(int tt = 5; tt >= LB; --tt) {
		const int b = tt - LB;
#pragma unroll
		for(int ra = 0; ra < NR; ++ra)
			if(!(ra & (1 << b))) butterfly(v[ra], v[ra | (1 << b)]);
#pragma unroll
		for(int ra = 0; ra < NR; ++ra) {
			if(!(ra & (1 << b))) continue;
			const int hi = ra & ((1 << b) - 1);  // register bits below the target
			v[ra] = rot<V>(v[ra], hi ? cmul(cq[(1 << tt) - 1 + (hi << LB)], qf[b]) : qf[b]);
		}
	}
#pragma unroll
	for(int ra = 0; ra < NR; ++ra) {
		V x = v[ra];
		if constexpr(LB == 4) x = tail_lane_step<V, 8>(x, lane, qf);
		x = tail_lane_step<V, 4>(x, lane, qf);
		x = tail_lane_step<V, 2>(x, lane, qf);
		x = tail_lane_step<V, 1>(x, lane, qf);
		v[ra] = mk<V>(x.x * 0.125f, x.y * 0.125f);
	}
}

// Transposed, bit-reversed store of the tile staged in LDS into tile `dst`
// (workgroups of 256 lanes, 16 amplitudes per lane).  Lane t writes column
// c = t mod 64 of output rows 4j + t/64; the output row r comes from LDS
// column rev6(r) of row rev6(c).
template <typename V>
__device__ __forceinline__ void tail_store(const V *tile, V *__restrict__ a, unsigned sh, u64 dst) {
	constexpr unsigned LB = TailLayout<V>::LB;
	const unsigned w = threadIdx.x >> 6, c = threadIdx.x & 63u, rowp = rev6(c);
	const V *src = tile + rowp * 64u + rev6(w);  // + (rev6(4j) ^ g): LDS column rev6(4j + w) ^ g
	const unsigned g = (rowp >> (6u - LB)) & ((1u << LB) - 1u);
	V *out = a + (((u64)w << sh) + (dst << 6) + c);
	const u64 step = (u64)4 << sh;
#pragma unroll
	for(unsigned j = 0; j < 16; ++j) {
		nt_store(src[(__brev(4u * j) >> 26) ^ g], out + j * step);
		if((j & 3u) == 3u) __builtin_amdgcn_sched_barrier(0);  // at most four reads in flight: registers
	}
}

// Row-wise writes of a lane's rows rp + RSTEP s into the LDS tile: register
// ra of row s is column (ra << LB) | q.
template <typename V>
__device__ __forceinline__ void tail_stage(V *tile, const V *v, unsigned rp, unsigned q) {
	using Lay = TailLayout<V>;
#pragma unroll
	for(unsigned s = 0; s < Lay::RPL; ++s) {
		V *dst = tile + tail_slot<V>(rp + Lay::RSTEP * s, q);
#pragma unroll
		for(unsigned ra = 0; ra < (1u << Lay::RB); ++ra) dst[ra * Lay::LPR] = v[(s << Lay::RB) + ra];
	}
}

template <typename V>
__device__ __forceinline__ void tail_load(V *v, const V *__restrict__ a, unsigned rp, unsigned q, unsigned sh, u64 mm) {
	using Lay = TailLayout<V>;
#pragma unroll
	for(unsigned s = 0; s < Lay::RPL; ++s) {
		const V *src = a + ((((u64)(rp + Lay::RSTEP * s)) << sh) + (mm << 6) + q);
#pragma unroll
		for(unsigned ra = 0; ra < (1u << Lay::RB); ++ra) v[(s << Lay::RB) + ra] = nt_load(src + ra * Lay::LPR);
	}
}

// Waves per SIMD the tail kernel is compiled for (one wave of each workgroup
// per SIMD): its LDS tile (64 KiB complex128, 32 KiB complex64) allows two
// workgroups per CU, or more for complex64, where registers allow three.
template <typename R> struct QftTailWaves {
	static constexpr int value = sizeof(R) == 8 ? 2 : 3;
};

// Workgroups of 256 lanes in the layout of TailLayout: every load instruction
// reads whole 128 B lines.  Both tiles of a pair are read before
// either is written; each is transformed in registers and staged through one
// 64 x 64 LDS tile for its transposed store.  When a workgroup has several
// pairs (complex64), the loads of its next pair are issued as soon as the
// registers they replace are staged, so they are in flight while the current
// pair is stored.
template <typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QftTailWaves<R>::value)))
k_qft_tail_reverse(typename V2T<R>::t *__restrict__ a, QftTailParams P, const double2 *__restrict__ cq) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	V *tile = (V *)smem;  // 64 x 64
	const unsigned t = threadIdx.x;
	using Lay = TailLayout<V>;
	constexpr unsigned LB = Lay::LB, RB = Lay::RB;
	const unsigned lane = t & 63u, q = t & (Lay::LPR - 1u), rp = t >> LB;
	const unsigned sh = P.n - 6;
	// qf[b] = e^{i pi q / 2^(LB + b)} for the register targets; then, for lane
	// targets t = LB-1 .. 1, e^{i pi (q mod 2^t) / 2^t} on lanes with bit t set.
	double2 qf[5];
#pragma unroll
	for(unsigned b = 0; b < RB; ++b) qf[b] = cq[(1u << (LB + b)) - 1u + q];
#pragma unroll
	for(unsigned tt = LB - 1; tt >= 1; --tt) {
		const unsigned low = q & ((1u << tt) - 1u);
		qf[RB + LB - 1 - tt] = ((q >> tt) & 1u) ? cq[(1u << tt) - 1u + low] : make_double2(1.0, 0.0);
	}
	u64 p = blockIdx.x;
	if(p >= P.npairs) return;  // uniform
	u64 M, Mr;
	tail_pair(P, p, M, Mr);
	V va[16], vb[16];
	tail_load(va, a, rp, q, sh, M);
	tail_load(vb, a, rp, q, sh, Mr);
	for(;;) {
		const u64 pn = p + gridDim.x;
		const bool more = pn < P.npairs;
		u64 Mn = 0, Mnr = 0;
		if(more) tail_pair(P, pn, Mn, Mnr);
#pragma unroll
		for(unsigned s = 0; s < Lay::RPL; ++s) {
			tail_transform(va + (s << RB), lane, qf, cq);
			tail_transform(vb + (s << RB), lane, qf, cq);
		}
		__syncthreads();  // the previous pair's transposed reads are done
		tail_stage(tile, va, rp, q);
		__syncthreads();
		if(more) tail_load(va, a, rp, q, sh, Mn);
		tail_store(tile, a, sh, Mr);  // tile M into the place of tile rev(M)
		__syncthreads();
		tail_stage(tile, vb, rp, q);
		__syncthreads();
		if(more) tail_load(vb, a, rp, q, sh, Mnr);
		tail_store(tile, a, sh, M);  // palindromes (M == rev(M)) write the same values twice
		if(!more) break;
		p = pn;
		M = Mn;
		Mr = Mnr;
	}
}

template <typename R>
static hipError_t launch_qft_tail_t(const Launch &L, void *a, unsigned n, const double *dcq) {
	using V = typename V2T<R>::t;
	QftTailParams P;
	P.n = n;
	P.m = n - 12;
	P.h = P.m / 2;
	P.K = 1u << P.h;
	P.npairs = (u64)((P.m & 1u) ? 2 : 1) * ((u64)P.K * (P.K + 1) / 2);
	const size_t shmem = 4096 * sizeof(V);
	// Measured at n = 30: complex128 runs fastest with one pair per workgroup
	// (6.50 ms against 6.80 ms at 16 pairs each), complex64 with about three
	// (3.34 ms at a grid of 49152 against 3.44 ms at one pair each).
	const u64 cap = sizeof(R) == 8 ? P.npairs : 49152;
	const unsigned grid = (unsigned)(P.npairs < cap ? P.npairs : cap);
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_qft_tail_reverse<R>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	hipLaunchKernelGGL(k_qft_tail_reverse<R>, dim3(grid), dim3(256), shmem, L.stream, (V *)a, P, (const double2 *)dcq);
	return hipGetLastError();
}

hipError_t launch_qft_tail_reverse(const Launch &L, void *a, unsigned n, const double *dcq) {
	if(n < 12 || n > 40) return hipErrorInvalidValue;
	return L.prec == 0 ? launch_qft_tail_t<double>(L, a, n, dcq) : launch_qft_tail_t<float>(L, a, n, dcq);
}

// ------------------------------------------------------------------------
// Ten-target QFT passes: (n - 10) / 10 rounded up block passes plus one tail
// pass fused with the bit reversal, three full-state passes at n = 30.
//
// A workgroup of 512 lanes owns a tile of 2^(10 + C) amplitudes: ten
// transform bits y and C passenger bits whose 2^C values make one 128 B run
// (C = 3 in complex128, 4 in complex64).  The ten-bit transform runs in two
// register stages joined by one LDS transpose:
//   complex128: 16 amplitudes per lane; stage A holds y bits 9..6 in
//   registers and bit 5 across lane bit 3 (DPP row rotation by 8), stage B
//   bits 4..1 in registers and bit 0 across lane bit 3;
//   complex64: 32 amplitudes per lane; stage A bits 9..5, stage B bits 4..0.
// The ladder factor e^{i pi (y mod 2^j) / 2^j} of a register target j is a
// constant of the register bits below j (a uniform operand) times a factor of
// the lane's other bits.  That lane factor multiplies only registers with bit
// j set, so it commutes with the stage's remaining register butterflies and
// is applied once per register after them: in stage A, register ra takes
// omega^(yA rev(ra)), omega = e^{2 pi i / 1024}, yA the lane's other y bits
// (one chain of products per lane and tile); in stage B the lane bit is y
// bit 0 and the factors are constants e^{i pi rev(rb) / 16}.  The phases of
// the bits r below the block become one twiddle per state, w^(r rev10(y)),
// w = e^{2 pi i / 2^(pb+10)}, as in the eight-target pass.  Butterflies are
// unscaled and (1/sqrt(2))^K goes into the twiddle.
template <typename R> struct WideQft;
template <> struct WideQft<double> {
	static constexpr unsigned C = 3, RB = 4, DL = 1;
};
template <> struct WideQft<float> {
	static constexpr unsigned C = 4, RB = 5, DL = 0;
};
template <typename R> using WV = typename V2T<R>::t;
template <typename R> using WP = typename LanePhase<WV<R>>::t;

__host__ __device__ constexpr unsigned rev_bits(unsigned x, unsigned b) {
	unsigned r = 0;
	for(unsigned i = 0; i < b; ++i) r |= ((x >> i) & 1u) << (b - 1u - i);
	return r;
}

// f(std::integral_constant<int, K>) for K = K0 .. N-1: register indices
// derived from K (bit reversals) are then constants, so register arrays stay
// in registers.
template <int K0, int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
	if constexpr(K0 < N) {
		f(std::integral_constant<int, K0>{});
		static_for<K0 + 1, N>(f);
	}
}

// Butterflies over register bit B, then the in-block constants of its
// target, c[B][low] = e^{i pi low / 2^B} (cst[2^B - 1 + low]), on the
// registers with bit B set (low = the register bits below B).
template <int B, int NR, typename V>
__device__ __forceinline__ void wide_reg_target(V *v, const typename LanePhase<V>::t *cst) {
#pragma unroll
	for(int r = 0; r < NR; ++r)
		if(!(r & (1 << B))) butterfly(v[r], v[r | (1 << B)]);
#pragma unroll
	for(int r = 0; r < NR; ++r) {
		if(!(r & (1 << B))) continue;
		const int low = r & ((1 << B) - 1);
		if(low) v[r] = rotp(v[r], cst[(1 << B) - 1 + low]);
	}
}

// Register targets B, B-1, ..., lowest (uniform; lowest <= 0: all).
template <int B, int NR, typename V>
__device__ __forceinline__ void wide_reg_targets(V *v, int lowest, const typename LanePhase<V>::t *cst) {
	if constexpr(B >= 0) {
		if(B >= lowest) {
			wide_reg_target<B, NR>(v, cst);
			wide_reg_targets<B - 1, NR>(v, lowest, cst);
		}
	}
}

// Stage A on a lane's registers (register ra = y bits 9 .. 10-RB), the top K
// y bits being targets (PART: K < 10).  stepA = omega^yA; lfd is the ladder
// factor of the lane target 9-RB, e^{i pi (yA mod 2^(9-RB)) / 2^(9-RB)} on
// lanes with that bit set, 1 elsewhere.
template <typename R, bool PART>
__device__ __forceinline__ void wide_stage_a(WV<R> *v, unsigned K, double2 stepA, double2 lfd, const WP<R> *cst) {
	using W = WideQft<R>;
	constexpr int RB = (int)W::RB, NA = 1 << RB;
	wide_reg_targets<RB - 1, NA>(v, PART ? RB - (int)K : 0, cst);
	const unsigned km = (!PART || K >= (unsigned)RB) ? (unsigned)NA - 1u : (1u << K) - 1u;
	double2 ph = make_double2(1.0, 0.0);
	static_for<1, NA>([&](auto kc) {
		constexpr int k = decltype(kc)::value;
		constexpr int ra = (int)rev_bits((unsigned)k, (unsigned)RB);
		if(!PART || (k & km)) {
			ph = cmul(ph, stepA);
			v[ra] = rot<WV<R>>(v[ra], ph);
		} else
			ph = make_double2(1.0, 0.0);
	});
	if constexpr(W::DL != 0) {
		if(!PART || K > (unsigned)RB) {
			const unsigned lane = threadIdx.x;
#pragma unroll
			for(int r = 0; r < NA; ++r) v[r] = rot<WV<R>>(lane_bfly<(1u << W::C)>(v[r], lane), lfd);
		}
	}
}

// Stage B (register rb = y bits RB+DL-1 .. DL; dB = y bit 0 across lane bit
// C when DL).  The lane factor of target j = b + DL is e^{i pi dB / 2^j};
// over the register bits it is e^{i pi dB k / 2^RB} on register rev(k)
// (cst[2^RB - 1 + k]), restricted to the target bits in a partial pass.
template <typename R, bool PART>
__device__ __forceinline__ void wide_stage_b(WV<R> *w, unsigned K, unsigned dB, const WP<R> *cst) {
	using W = WideQft<R>;
	constexpr int RB = (int)W::RB, NA = 1 << RB, DL = (int)W::DL;
	wide_reg_targets<RB - 1, NA>(w, PART ? 10 - (int)K - DL : 0, cst);
	if constexpr(DL != 0) {
		const int e = PART ? (int)K + DL + RB - 10 : RB;
		const unsigned mB = e <= 0 ? 0u : (e >= RB ? (unsigned)NA - 1u : (1u << e) - 1u);
		static_for<1, NA>([&](auto kc) {
			constexpr int k = decltype(kc)::value;
			constexpr int rb = (int)rev_bits((unsigned)k, (unsigned)RB);
			if(PART && !(k & mB)) return;
			const WV<R> t = rotp(w[rb], cst[(1 << RB) - 1 + (PART ? (k & (int)mB) : k)]);
			// component selects: a select of the whole vector struct becomes a
			// select of addresses, which puts the register array in scratch
			w[rb].x = dB ? t.x : w[rb].x;
			w[rb].y = dB ? t.y : w[rb].y;
		});
		if(!PART || K == 10u) {
			const unsigned lane = threadIdx.x;
#pragma unroll
			for(int r = 0; r < NA; ++r) w[r] = lane_bfly<(1u << W::C)>(w[r], lane);
		}
	}
}

struct WideBlockParams {
	u64 ntiles;  // 2^(n - 10 - C)
	uint32_t n, pbot, K, xcd;  // xcd: tiles in XCD-contiguous order (ntiles a multiple of 8)
	double scale;              // (1/sqrt(2))^K
};

// Workgroup v's tile when tiles are dealt in XCD-contiguous order: workgroups
// are dispatched round-robin over the eight XCDs, so v mod 8 names the XCD
// and each XCD walks its own contiguous eighth of the tiles.
__device__ __forceinline__ u64 xcd_tile(u64 v, u64 ntiles, unsigned xcd) {
	return xcd ? ((v & 7ull) * (ntiles >> 3) + (v >> 3)) : v;
}

// In-place pass over the block bits pbot+9 .. pbot, whose top K bits are
// targets (K = 10, or fewer in a leading pass); C passenger bits lo at the
// bottom of the index (pbot >= C).  A 128 KiB LDS tile per workgroup (one
// workgroup per CU), tiles dealt in XCD-contiguous order to a persistent
// grid.  The next tile is loaded straight into LDS while the current one
// finishes: once stage B has read its tile out of LDS, the next tile's rows
// are requested into it (16 LDS-direct loads per lane), so they are in flight
// during stage B, the twiddle and the stores; those stores drain during the
// next tile's stage A.  The LDS tile holds the rows as they sit in memory
// (slot (y << C) | lo), which is also the exchange layout.  Waiting for the
// next tile takes vmcnt(NA): the NA stores issued after its loads may stay
// outstanding.  Nothing else may sit on that counter between the loads and
// the wait: the stage constants come through scalar loads and the twiddle's
// phase tables are copied into LDS beside the tile.
// Measured at n = 30 against other forms: one tile per workgroup with
// register loads 8-16% slower (complex128 9.0 / 7.2 ms against 8.3 / 6.1 ms
// at pbot 20 / 10); prefetching into registers spilled; exchanging the tile
// through LDS in two 64 KiB halves (without a second workgroup per CU,
// which 16 complex128 amplitudes per lane cannot fit in 128 registers) was
// 4-8% slower.
template <typename R, bool PART>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
k_qft_wide_block(WV<R> *__restrict__ a, WideBlockParams P, const double2 *__restrict__ ptab, const WP<R> *__restrict__ cst) {
	using V = WV<R>;
	using W = WideQft<R>;
	constexpr unsigned C = W::C, RB = W::RB, DL = W::DL, NA = 1u << RB, CM = (1u << C) - 1u;
	constexpr unsigned SB = RB + DL;
	constexpr unsigned EPL = 16u / (unsigned)sizeof(V);            // slots per lane and load
	constexpr unsigned NDMA = (1u << (10u + C)) / (512u * EPL);   // loads per lane and tile
	static_assert(NDMA * 512u * EPL == (1u << (10u + C)), "tile");
	extern __shared__ char smem[];
	V *tile = (V *)smem;
	const unsigned t = threadIdx.x;
	const unsigned lo = t & CM, u = t >> C;
	const unsigned yA = DL ? (((u & 1u) << (9u - RB)) | (u >> 1)) : u;
	const unsigned dB = DL ? (u & 1u) : 0u, ub = u >> DL;
	const unsigned pb = P.pbot, K = P.K;
	const u64 lowmask = (1ull << pb) - 1ull;
	const double2 stepA = phase_of_range((u64)yA << 54, ptab, 0, 1);
	double2 lfd = make_double2(1.0, 0.0);
	if(DL && (yA >> (9u - RB))) lfd = phase_of_range((u64)(yA & ((1u << (9u - RB)) - 1u)) << (54u + RB), ptab, 0, 1);
	const unsigned kmask = (1u << K) - 1u;
	const unsigned mt = (rev_bits(ub, 10u - SB) | (dB << 9)) & kmask;
	const unsigned m2 = kmask >> (10u - SB);
	const unsigned wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63u;
	// the twiddle's phase tables in LDS after the tile: their reads do not
	// queue behind the next tile's loads on the vector-memory counter
	double2 *ltab = (double2 *)(smem + ((size_t)1 << (10u + C)) * sizeof(V));
	for(unsigned i = t; i < 5u * 256u; i += 512u) ltab[i] = ptab[i];
	// load i of lane ln fills slot s = (i * 8 + wv) * 64 EPL + ln EPL = (y << C) | lo
	// with y = 64 i + ((wv * 64 + ln) EPL >> C) (2^C = 8 EPL): one lane offset
	// plus i rows of 64 << pb (kept as one base, not sixteen offsets in registers)
	const unsigned s_lane = (wv * 64u + ln) * EPL;
	const u64 lane_off = ((u64)(s_lane >> C) << pb) + (s_lane & CM);
	auto request = [&](u64 vt) {
		const u64 x = xcd_tile(vt, P.ntiles, P.xcd) << C;
		const V *src = a + ((((x >> pb) << (pb + 10)) | (x & lowmask)) + lane_off);
#pragma unroll
		for(unsigned i = 0; i < NDMA; ++i)
			lds_direct_load16(src + ((u64)i << (pb + 6)), tile + (i * 8u + wv) * 64u * EPL);
	};
	V v[NA];
	u64 vt = blockIdx.x;
	if(vt < P.ntiles) request(vt);
	bool first = true;
	for(; vt < P.ntiles; vt += gridDim.x) {
		const u64 x = xcd_tile(vt, P.ntiles, P.xcd) << C;
		V *const base = a + ((((x >> pb) << (pb + 10)) | (x & lowmask)) + lo);
		if(first) wait_vmcnt<0>();
		else wait_vmcnt<NA>();
		first = false;
		__syncthreads();  // every wavefront's part of the tile is in LDS
#pragma unroll
		for(unsigned ra = 0; ra < NA; ++ra) v[ra] = tile[(((ra << (10u - RB)) | yA) << C) | lo];
		wide_stage_a<R, PART>(v, K, stepA, lfd, cst);
		__syncthreads();  // every lane has read its stage-A amplitudes
#pragma unroll
		for(unsigned ra = 0; ra < NA; ++ra) tile[(((ra << (10u - RB)) | yA) << C) | lo] = v[ra];
		__syncthreads();
#pragma unroll
		for(unsigned rb = 0; rb < NA; ++rb) v[rb] = tile[(((ub << SB) | (rb << DL) | dB) << C) | lo];
		__syncthreads();  // the tile is in registers: the LDS takes the next one
		if(vt + gridDim.x < P.ntiles) request(vt + gridDim.x);
		wide_stage_b<R, PART>(v, K, dB, cst);
		const u64 r = (x & lowmask) | lo;
		double2 bph = phase_of((r * mt) << (54u - pb), ltab);
		bph.x *= P.scale;
		bph.y *= P.scale;
		const double2 us = phase_of(r << (64u - pb - SB), ltab);
		double2 ph = bph;
		static_for<0, (int)NA>([&](auto kc) {
			constexpr unsigned k = (unsigned)decltype(kc)::value;
			constexpr unsigned rb = rev_bits(k, RB);
			if(k) ph = (!PART || (k & m2)) ? cmul(ph, us) : bph;
			v[rb] = rot<V>(v[rb], ph);
		});
#pragma unroll
		for(unsigned rb = 0; rb < NA; ++rb) nt_store(v[rb], base + ((u64)((ub << SB) | (rb << DL) | dB) << pb));
	}
}

struct WideTailParams {
	u64 ntiles;  // 2^(n - 10 - C)
	uint32_t n, f, xcd, pad;  // f = n - 10 - C: bits between the passengers and the targets
};

// Stage-A coordinates of lane t in the tail pass: its Lo bits yA and its h.
template <typename R>
__device__ __forceinline__ void wide_tail_lane(unsigned t, unsigned &yA, unsigned &hA) {
	using W = WideQft<R>;
	constexpr unsigned C = W::C, RB = W::RB, DL = W::DL, HM = (1u << (C - 1u)) - 1u;
	const unsigned h0 = (t >> (C - 1u)) & 1u;
	if constexpr(DL != 0) {
		constexpr unsigned NL = 10u - RB - (C - 1u) - 1u;
		const unsigned rest = t >> (C + 1u);
		yA = (t & HM) | ((rest & ((1u << NL) - 1u)) << (C - 1u)) | (((t >> C) & 1u) << (9u - RB));
		hA = h0 | ((rest >> NL) << 1);
	} else {
		constexpr unsigned NL = 10u - RB - (C - 1u);
		const unsigned rest = t >> C;
		yA = (t & HM) | ((rest & ((1u << NL) - 1u)) << (C - 1u));
		hA = h0 | ((rest >> NL) << 1);
	}
}

// QFT targets 9..0 followed by the full bit reversal, out of place.  A tile
// holds the ten target bits Lo and the top C bits h of the index (the
// passengers: after the reversal they are the bottom C bits, one 128 B run
// per output row); tile p has the middle bits F = rev_f(p), so its output
// index is (rev10(Lo), p, rev_C(h)) and consecutive tiles write consecutive
// runs.  Stage A lanes: h bit 0 on lane bit C-1, the low 10 - RB bits of Lo
// (Lo bit 9-RB on lane bit C in complex128) and the other h bits on top;
// each wavefront reads runs of 256 B or more (measured 7% faster than 1 KiB
// runs with all of h on top).  Stage B lanes: rev_C(h) on the bottom C lane
// bits, then Lo bit 0 (complex128) and Lo bits 9..5.  Persistent grid with
// the next tile loaded straight into LDS during stage B and the stores, as
// in k_qft_wide_block.  Prefetched tile layout: slot ra * 512 + t holds the
// amplitude lane t takes into register ra, so each load instruction fetches
// what one stage-A register load of the wavefront would (complex64: a
// lane's 16 bytes are the amplitudes of lanes t and t + 1, t even, which are
// adjacent in memory).  Measured at n = 30: 6.5 / 3.6 ms against 6.9 / 3.8
// ms with one tile per workgroup and register loads.
template <typename R>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
k_qft_wide_tail(const WV<R> *__restrict__ a, WV<R> *__restrict__ o, WideTailParams P, const double2 *__restrict__ ptab,
                const WP<R> *__restrict__ cst) {
	using V = WV<R>;
	using W = WideQft<R>;
	constexpr unsigned C = W::C, RB = W::RB, DL = W::DL, NA = 1u << RB, CM = (1u << C) - 1u;
	constexpr unsigned EPL = 16u / (unsigned)sizeof(V), NDMA = NA / EPL;
	static_assert(NDMA * 512u * EPL == (1u << (10u + C)), "tile");
	extern __shared__ char smem[];
	V *tile = (V *)smem;
	const unsigned t = threadIdx.x;
	unsigned yA, hA;
	wide_tail_lane<R>(t, yA, hA);
	const unsigned g = t & CM, u = t >> C;
	const unsigned dB = DL ? (u & 1u) : 0u, ub = u >> DL;
	const unsigned hB = rev_bits(g, C);
	const double2 stepA = phase_of_range((u64)yA << 54, ptab, 0, 1);
	double2 lfd = make_double2(1.0, 0.0);
	if(DL && (yA >> (9u - RB))) lfd = phase_of_range((u64)(yA & ((1u << (9u - RB)) - 1u)) << (54u + RB), ptab, 0, 1);
	const unsigned n = P.n, f = P.f;
	const unsigned wv = __builtin_amdgcn_readfirstlane(t >> 6), ln = t & 63u;
	auto request = [&](u64 vt) {
		const u64 tid = xcd_tile(vt, P.ntiles, P.xcd);
		const u64 F = f ? (u64)(__brev((unsigned)tid) >> (32u - f)) : 0ull;
#pragma unroll
		for(unsigned q = 0; q < NDMA; ++q) {
			const unsigned s0 = (q * 8u + wv) * 64u * EPL, s = s0 + ln * EPL;
			const unsigned ra = s >> 9, tt = s & 511u;
			unsigned y, h;
			wide_tail_lane<R>(tt, y, h);
			const V *src = a + ((((u64)h << (n - C)) | (F << 10) | y) + ((u64)ra << (10u - RB)));
			lds_direct_load16(src, tile + s0);
		}
	};
	V v[NA];
	u64 vt = blockIdx.x;
	if(vt < P.ntiles) request(vt);
	bool first = true;
	for(; vt < P.ntiles; vt += gridDim.x) {
		const u64 tid = xcd_tile(vt, P.ntiles, P.xcd);
		if(first) wait_vmcnt<0>();
		else wait_vmcnt<NA>();
		first = false;
		__syncthreads();  // every wavefront's part of the tile is in LDS
#pragma unroll
		for(unsigned ra = 0; ra < NA; ++ra) v[ra] = tile[ra * 512u + t];
		wide_stage_a<R, false>(v, 10u, stepA, lfd, cst);
		__syncthreads();  // every lane has read its stage-A amplitudes
#pragma unroll
		for(unsigned ra = 0; ra < NA; ++ra) {
			const unsigned L = (ra << (10u - RB)) | yA;
			tile[(L << C) | (hA ^ (L & CM))] = v[ra];
		}
		__syncthreads();
#pragma unroll
		for(unsigned rb = 0; rb < NA; ++rb) {
			const unsigned L = (ub << (RB + DL)) | (rb << DL) | dB;
			v[rb] = tile[(L << C) | (hB ^ (L & CM))];
		}
		__syncthreads();  // the tile is in registers: the LDS takes the next one
		if(vt + gridDim.x < P.ntiles) request(vt + gridDim.x);
		wide_stage_b<R, false>(v, 10u, dB, cst);
		V *dst = o + ((tid << C) | g);
#pragma unroll
		for(unsigned rb = 0; rb < NA; ++rb) {
			const unsigned L = (ub << (RB + DL)) | (rb << DL) | dB;
			const V x = v[rb];
			nt_store(mk<V>(x.x * 0.03125f, x.y * 0.03125f), dst + ((u64)rev_bits(L, 10u) << (n - 10u)));
		}
	}
}

unsigned qft_wide_min_qubits(int prec) { return 10u + (prec == 0 ? WideQft<double>::C : WideQft<float>::C); }

// Grids of the ten-target passes, measured at n = 30 (HIP events; rocprof
// agrees): blocks whose rows lie at least 2 MiB apart (pbot 20) are fastest
// with 512-1024 workgroups (complex128 8.3 ms, complex64 4.1 ms; 8.5-8.8 /
// 4.4-5.0 ms at 32768-65536); nearer rows (pbot 10) with two tiles per
// workgroup in complex128 (6.1 ms; 6.3-8.1 ms at 512-32768) and four in
// complex64 (3.05 ms; 4.0 ms at one tile each); the tail with four tiles per
// workgroup (6.5 / 3.6 ms).
static u64 wide_grid_floor(u64 ntiles) { return ntiles < 1024 ? ntiles : 1024; }

template <typename R, bool PART>
static hipError_t launch_qft_wide_block_t(const Launch &L, void *a, unsigned n, unsigned pbot, unsigned K, const double *dphase,
                                          const double *dcin, u64 grid_cap) {
	using V = WV<R>;
	WideBlockParams P;
	P.n = n;
	P.pbot = pbot;
	P.K = K;
	P.ntiles = 1ull << (n - 10 - WideQft<R>::C);
	P.xcd = P.ntiles >= 8 ? 1u : 0u;
	P.scale = std::ldexp(K & 1u ? 1.0 / std::sqrt(2.0) : 1.0, -(int)(K / 2));  // (1/sqrt(2))^K, one rounding at most
	u64 cap = grid_cap;
	if(!cap) {
		const bool far = ((u64)sizeof(V) << pbot) >= (2ull << 20);
		const u64 per = sizeof(R) == 8 ? 2 : 4;  // tiles per workgroup when the rows are near
		cap = far ? wide_grid_floor(P.ntiles) : std::max(wide_grid_floor(P.ntiles), P.ntiles / per);
	}
	const unsigned grid = (unsigned)(P.ntiles < cap ? P.ntiles : cap);
	const size_t shmem = ((size_t)1 << (10 + WideQft<R>::C)) * sizeof(V) + 5 * 256 * sizeof(double2);  // tile, phase tables
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_qft_wide_block<R, PART>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	const WP<R> *cinp = sizeof(R) == 8 ? (const WP<R> *)dcin : (const WP<R> *)(dcin + 126);
	hipLaunchKernelGGL((k_qft_wide_block<R, PART>), dim3(grid), dim3(512), shmem, L.stream, (V *)a, P, (const double2 *)dphase, cinp);
	return hipGetLastError();
}

hipError_t launch_qft_wide_block(const Launch &L, void *a, unsigned n, unsigned pbot, unsigned K, const double *dphase,
                                 const double *dcin, u64 grid_cap) {
	const unsigned C = L.prec == 0 ? WideQft<double>::C : WideQft<float>::C;
	if(n > 40 || pbot < C || pbot + 10 > n || K < 1 || K > 10) return hipErrorInvalidValue;
	if(L.prec == 0)
		return K == 10 ? launch_qft_wide_block_t<double, false>(L, a, n, pbot, K, dphase, dcin, grid_cap)
		               : launch_qft_wide_block_t<double, true>(L, a, n, pbot, K, dphase, dcin, grid_cap);
	return K == 10 ? launch_qft_wide_block_t<float, false>(L, a, n, pbot, K, dphase, dcin, grid_cap)
	               : launch_qft_wide_block_t<float, true>(L, a, n, pbot, K, dphase, dcin, grid_cap);
}

template <typename R>
static hipError_t launch_qft_wide_tail_t(const Launch &L, const void *src, void *dst, unsigned n, const double *dphase,
                                         const double *dcin, u64 grid_cap) {
	using V = WV<R>;
	WideTailParams P;
	P.n = n;
	P.f = n - 10 - WideQft<R>::C;
	P.pad = 0;
	P.ntiles = 1ull << P.f;
	P.xcd = P.ntiles >= 8 ? 1u : 0u;
	const u64 cap = grid_cap ? grid_cap : std::max(wide_grid_floor(P.ntiles), P.ntiles / 4);
	const unsigned grid = (unsigned)(P.ntiles < cap ? P.ntiles : cap);
	const size_t shmem = ((size_t)1 << (10 + WideQft<R>::C)) * sizeof(V);
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_qft_wide_tail<R>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	const WP<R> *cinp = sizeof(R) == 8 ? (const WP<R> *)dcin : (const WP<R> *)(dcin + 126);
	hipLaunchKernelGGL((k_qft_wide_tail<R>), dim3(grid), dim3(512), shmem, L.stream, (const V *)src, (V *)dst, P,
	                   (const double2 *)dphase, cinp);
	return hipGetLastError();
}

hipError_t launch_qft_wide_tail(const Launch &L, const void *src, void *dst, unsigned n, const double *dphase, const double *dcin,
                                u64 grid_cap) {
	if(n < qft_wide_min_qubits(L.prec) || n > 40 || src == dst) return hipErrorInvalidValue;
	return L.prec == 0 ? launch_qft_wide_tail_t<double>(L, src, dst, n, dphase, dcin, grid_cap)
	                   : launch_qft_wide_tail_t<float>(L, src, dst, n, dphase, dcin, grid_cap);
}

}  // namespace tmfqs_hip
