// gfx950 kernels of the dense state-vector engine.
//
// Storage is interleaved complex (double2 for complex128, float2 for
// complex64), basis index bit p holds qubit n-1-p.  Every kernel indexes with
// 64-bit integers.  The library is compiled with -ffp-contract=off so that the
// per-operation kernels reproduce the reference dense backend's arithmetic
// (src/tmfqs/storage/dense/dense_state_backend.cpp) bit for bit; kernels that
// are not bit-exact by design (phase ladders, twiddles) use explicit fma().
#include "engine.h"

#include <climits>
#include <cmath>

namespace tmfqs_hip {

template <typename R> struct V2T;
template <> struct V2T<double> { using t = double2; };
template <> struct V2T<float> { using t = float2; };

__device__ __forceinline__ u64 gtid() { return (u64)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ u64 gstride() { return (u64)gridDim.x * blockDim.x; }
__host__ __device__ __forceinline__ u64 ins0(u64 x, unsigned b) {
	return ((x >> b) << (b + 1)) | (x & ((1ull << b) - 1ull));
}

static unsigned grid_for(u64 work, unsigned threads = 256, u64 maxblocks = 8192) {
	u64 b = (work + threads - 1) / threads;
	if(b < 1) b = 1;
	if(b > maxblocks) b = maxblocks;
	return (unsigned)b;
}

template <typename V> __device__ __forceinline__ V mk(double x, double y);
template <> __device__ __forceinline__ double2 mk<double2>(double x, double y) { return make_double2(x, y); }
template <> __device__ __forceinline__ float2 mk<float2>(double x, double y) { return make_float2((float)x, (float)y); }

// Streaming (non-temporal) 16/8-byte loads and stores: a sweep touches each
// amplitude once, so it should not displace reusable lines in L2/MALL.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 nt_load(const double2 *p) {
	const nt_d2 v = __builtin_nontemporal_load((const nt_d2 *)p);
	return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 nt_load(const float2 *p) {
	const nt_f2 v = __builtin_nontemporal_load((const nt_f2 *)p);
	return make_float2(v.x, v.y);
}
__device__ __forceinline__ void nt_store(double2 v, double2 *p) {
	nt_d2 w;
	w.x = v.x;
	w.y = v.y;
	__builtin_nontemporal_store(w, (nt_d2 *)p);
}
__device__ __forceinline__ void nt_store(float2 v, float2 *p) {
	nt_f2 w;
	w.x = v.x;
	w.y = v.y;
	__builtin_nontemporal_store(w, (nt_f2 *)p);
}

// ------------------------------------------------------------------------
// Per-operation sweeps (one pass per reference backend call).

// dense_state_backend.cpp:255-273: a0' = (a0+a1)c, a1' = (a0-a1)c, all-zero pairs untouched.
template <typename R>
__global__ void __launch_bounds__(256) k_hadamard(typename V2T<R>::t *__restrict__ a, u64 npairs, unsigned bit, R c) {
	for(u64 t = gtid(); t < npairs; t += gstride()) {
		const u64 i0 = ins0(t, bit), i1 = i0 | (1ull << bit);
		const auto x = a[i0], y = a[i1];
		if(x.x == 0 && x.y == 0 && y.x == 0 && y.y == 0) continue;
		typename V2T<R>::t o0, o1;
		o0.x = (x.x + y.x) * c;
		o0.y = (x.y + y.y) * c;
		o1.x = (x.x - y.x) * c;
		o1.y = (x.y - y.y) * c;
		a[i0] = o0;
		a[i1] = o1;
	}
}

// dense_state_backend.cpp:276-286 / :306-318: swap the pair when state0 holds
// every control bit, unless the two amplitudes are equal.
template <typename R>
__global__ void __launch_bounds__(256) k_pauli_x(typename V2T<R>::t *__restrict__ a, u64 npairs, unsigned bit, u64 cmask) {
	for(u64 t = gtid(); t < npairs; t += gstride()) {
		const u64 i0 = ins0(t, bit), i1 = i0 | (1ull << bit);
		if((i0 & cmask) != cmask) continue;
		const auto x = a[i0], y = a[i1];
		if(x.x == y.x && x.y == y.y) continue;
		a[i0] = y;
		a[i1] = x;
	}
}

// dense_state_backend.cpp:289-303: states with the target and control bits set
// are multiplied by (cr, ci); zero amplitudes untouched.
template <typename R>
__global__ void __launch_bounds__(256) k_cphase(typename V2T<R>::t *__restrict__ a, u64 nquads, unsigned lo, unsigned hi,
                                                u64 setmask, R cr, R ci) {
	for(u64 t = gtid(); t < nquads; t += gstride()) {
		const u64 i = ins0(ins0(t, lo), hi) | setmask;
		const auto x = a[i];
		if(x.x == 0 && x.y == 0) continue;
		typename V2T<R>::t o;
		o.x = cr * x.x - ci * x.y;
		o.y = cr * x.y + ci * x.x;
		a[i] = o;
	}
}

// gate_block_apply.h:117-143: one-qubit matrix with the k==1 accumulation order.
template <typename R>
__global__ void __launch_bounds__(256) k_u1(typename V2T<R>::t *__restrict__ a, u64 npairs, unsigned bit, R g00r, R g00i,
                                            R g01r, R g01i, R g10r, R g10i, R g11r, R g11i) {
	for(u64 t = gtid(); t < npairs; t += gstride()) {
		const u64 i0 = ins0(t, bit), i1 = i0 | (1ull << bit);
		const auto x = a[i0], y = a[i1];
		if((x.x == 0 && x.y == 0) && (y.x == 0 && y.y == 0)) continue;
		typename V2T<R>::t o0, o1;
		o0.x = g00r * x.x - g00i * x.y + g01r * y.x - g01i * y.y;
		o0.y = g00r * x.y + g00i * x.x + g01r * y.y + g01i * y.x;
		o1.x = g10r * x.x - g10i * x.y + g11r * y.x - g11i * y.y;
		o1.y = g10r * x.y + g10i * x.x + g11r * y.y + g11i * y.x;
		a[i0] = o0;
		a[i1] = o1;
	}
}

// gate_block_apply.h:145-211: k-qubit matrix (k >= 2), blocks gathered in the
// order of the matrix index (bits[0] is its MSB), all-zero blocks skipped,
// rows accumulated column by column.
template <typename R, int K>
__global__ void __launch_bounds__(256) k_uk(typename V2T<R>::t *__restrict__ a, u64 nblocks, const double *__restrict__ dmat,
                                            unsigned b0, unsigned b1, unsigned b2, unsigned b3, unsigned b4) {
	constexpr int D = 1 << K;
	__shared__ R g[2 * D * D];
	for(int i = threadIdx.x; i < 2 * D * D; i += blockDim.x) g[i] = (R)dmat[i];
	__syncthreads();
	const unsigned bits[5] = {b0, b1, b2, b3, b4};
	unsigned sorted[5];
	for(int i = 0; i < K; ++i) sorted[i] = bits[i];
	for(int i = 0; i < K; ++i)
		for(int j = i + 1; j < K; ++j)
			if(sorted[j] < sorted[i]) {
				const unsigned t = sorted[i];
				sorted[i] = sorted[j];
				sorted[j] = t;
			}
	for(u64 blk = gtid(); blk < nblocks; blk += gstride()) {
		u64 base = blk;
		for(int i = 0; i < K; ++i) base = ins0(base, sorted[i]);
		R xr[D], xi[D];
		u64 idx[D];
		bool nz = false;
#pragma unroll
		for(int r = 0; r < D; ++r) {
			u64 off = 0;
#pragma unroll
			for(int i = 0; i < K; ++i)
				if((r >> (K - 1 - i)) & 1) off |= 1ull << bits[i];
			idx[r] = base | off;
			const auto v = a[idx[r]];
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
				const R gr = g[2 * (r * D + c)], gi = g[2 * (r * D + c) + 1];
				nr += gr * xr[c] - gi * xi[c];
				ni += gr * xi[c] + gi * xr[c];
			}
			typename V2T<R>::t o;
			o.x = nr;
			o.y = ni;
			a[idx[r]] = o;
		}
	}
}

// Swap bit positions p < q: exchange states with (p=1, q=0) and (p=0, q=1).
template <typename R>
__global__ void __launch_bounds__(256) k_bitswap(typename V2T<R>::t *__restrict__ a, u64 nquads, unsigned p, unsigned q) {
	for(u64 t = gtid(); t < nquads; t += gstride()) {
		const u64 i = ins0(ins0(t, p), q) | (1ull << p);
		const u64 j = i ^ (1ull << p) ^ (1ull << q);
		const auto x = a[i], y = a[j];
		if(x.x == y.x && x.y == y.y) continue;
		a[i] = y;
		a[j] = x;
	}
}

// In-place bit permutation by an involution perm (perm[perm[b]] = b).  The
// local set S = {0..C-1} U perm({0..C-1}) is closed under perm, so tile A
// (fixed bits outside S) and tile perm(A) swap contents; each workgroup owns
// one such pair and stages both tiles in LDS.
struct InvParams {
	u64 npairs_tiles;  // 2^(n - s)
	uint32_t n, C, s, nh;
	uint32_t perm[64];
	uint32_t hbits[16];  // S \ {0..C-1}, ascending
	uint32_t lperm[16];  // perm restricted to S in local-bit numbering
};

template <typename R>
__global__ void __launch_bounds__(256) k_involution(typename V2T<R>::t *__restrict__ a, InvParams P) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	const u64 nloc = 1ull << P.s;
	V *ta = (V *)smem;
	V *tb = ta + nloc;
	const u64 lowmask = (1ull << P.C) - 1ull;
	for(u64 tid = blockIdx.x; tid < P.npairs_tiles; tid += gridDim.x) {
		u64 base = tid << P.C;
		for(unsigned b = 0; b < P.nh; ++b) base = ins0(base, P.hbits[b]);
		u64 pb = 0;
		for(unsigned b = 0; b < P.n; ++b)
			if((base >> b) & 1ull) pb |= 1ull << P.perm[b];
		if(pb < base) continue;  // uniform: the partner tile's workgroup handles the pair
		const bool self = (pb == base);
		__syncthreads();
		for(u64 l = threadIdx.x; l < nloc; l += blockDim.x) {
			u64 hi = 0;
			for(unsigned b = 0; b < P.nh; ++b)
				if((l >> (P.C + b)) & 1ull) hi |= 1ull << P.hbits[b];
			const u64 off = (l & lowmask) | hi;
			ta[l] = a[base | off];
			if(!self) tb[l] = a[pb | off];
		}
		__syncthreads();
		for(u64 l = threadIdx.x; l < nloc; l += blockDim.x) {
			u64 hi = 0, src = 0;
			for(unsigned b = 0; b < P.nh; ++b)
				if((l >> (P.C + b)) & 1ull) hi |= 1ull << P.hbits[b];
			for(unsigned b = 0; b < P.s; ++b)
				if((l >> b) & 1ull) src |= 1ull << P.lperm[b];
			const u64 off = (l & lowmask) | hi;
			if(self) {
				a[base | off] = ta[src];
			} else {
				a[pb | off] = ta[src];
				a[base | off] = tb[src];
			}
		}
	}
}

// ------------------------------------------------------------------------
// Elementwise chains: the Grover diffusion recurrence and affine overlay
// flushes folded into one sweep.  Each step reproduces the reference's
// arithmetic (dense_state_backend.cpp:230-252); the chain runs in FP64
// registers for both precisions.
template <typename R>
__global__ void __launch_bounds__(256) k_ew_chain(typename V2T<R>::t *__restrict__ a, u64 N, const EwStep *__restrict__ steps,
                                                  unsigned nsteps) {
	using V = typename V2T<R>::t;
	for(u64 i = gtid(); i < N; i += gstride()) {
		const V v = nt_load(&a[i]);
		double re = v.x, im = v.y;
		for(unsigned s = 0; s < nsteps; ++s) {
			const EwStep st = steps[s];
			if(st.kind == EW_DIFFUSE) {
				// (ar, ai) already hold 2m: doubling is exact, so 2m - a rounds once as in the reference
				re = st.ar - re;
				im = st.ai - im;
			} else {
				const double nr = st.ar * re - st.ai * im + st.br;
				const double ni = st.ar * im + st.ai * re + st.bi;
				re = nr;
				im = ni;
			}
		}
		nt_store(mk<V>(re, im), &a[i]);
	}
}

// Plain streaming copy (bandwidth ceiling and counter calibration).
template <typename R>
__global__ void __launch_bounds__(256) k_copy(const typename V2T<R>::t *__restrict__ src, typename V2T<R>::t *__restrict__ dst, u64 N) {
	for(u64 i = gtid(); i < N; i += gstride()) nt_store(nt_load(&src[i]), &dst[i]);
}

// Streaming passes (copies, elementwise chains, per-operation sweeps) run on
// up to 2^18 workgroups: a 2^30-amplitude copy measured 5.69 TB/s at that
// grid against 5.10 TB/s at 2^14; per-operation Hadamard and X sweeps were
// 6-12% faster than at 2^13.
constexpr u64 kStreamGrid = 262144;

hipError_t launch_copy(const Launch &L, const void *src, void *dst, unsigned n) {
	const u64 N = 1ull << n;
	if(L.prec == 0)
		hipLaunchKernelGGL(k_copy<double>, dim3(grid_for(N, 256, kStreamGrid)), dim3(256), 0, L.stream, (const double2 *)src, (double2 *)dst, N);
	else
		hipLaunchKernelGGL(k_copy<float>, dim3(grid_for(N, 256, kStreamGrid)), dim3(256), 0, L.stream, (const float2 *)src, (float2 *)dst, N);
	return hipGetLastError();
}

// Pure diffusion chains (the Grover loop): steps hold 2m per diffusion.  Each
// thread streams 4 elements of one 1024-element block (4 x 16 B in flight per
// lane) and applies every step to all four in FP64 registers, so one sweep of
// HBM carries the whole chain.  a = 2m - a rounds once per step exactly as the
// reference's inversionAboutMean(mean) does.
//
// Real states (every imaginary part +-0, as in Grover from a basis state):
// when every step's Im(2m) is +-0 too (zero_ok), the imaginary chain only
// ever subtracts zeros, and IEEE subtraction makes its result depend on
// nothing but the sign of the starting zero.  The host runs the chain on +0
// and on -0 (zp, zm); a wavefront whose imaginary parts are all zero takes
// those and computes only the real chain (half the FP64 adds), with the same
// bits as the full chain.
template <typename R>
__global__ void __launch_bounds__(256) k_diffuse_chain(typename V2T<R>::t *__restrict__ a, u64 N, const double2 *__restrict__ m2,
                                                       unsigned nsteps, double zp, double zm, int zero_ok) {
	using V = typename V2T<R>::t;
	constexpr int U = 4;
	const u64 stride = (u64)gridDim.x * blockDim.x * U;
	for(u64 base = (u64)blockIdx.x * blockDim.x * U + threadIdx.x; base < N; base += stride) {
		double re[U], im[U];
#pragma unroll
		for(int u = 0; u < U; ++u) {
			const u64 i = base + (u64)u * blockDim.x;
			V v = i < N ? nt_load(&a[i]) : mk<V>(0.0, 0.0);
			re[u] = v.x;
			im[u] = v.y;
		}
		bool real = zero_ok != 0;
#pragma unroll
		for(int u = 0; u < U; ++u) real = real && im[u] == 0.0;
		if(__all(real)) {
#pragma unroll 8
			for(unsigned s = 0; s < nsteps; ++s) {
				const double mx = m2[s].x;
#pragma unroll
				for(int u = 0; u < U; ++u) re[u] = mx - re[u];
			}
#pragma unroll
			for(int u = 0; u < U; ++u) im[u] = signbit(im[u]) ? zm : zp;
		} else {
			for(unsigned s = 0; s < nsteps; ++s) {
				const double2 m = m2[s];
#pragma unroll
				for(int u = 0; u < U; ++u) {
					re[u] = m.x - re[u];
					im[u] = m.y - im[u];
				}
			}
		}
#pragma unroll
		for(int u = 0; u < U; ++u) {
			const u64 i = base + (u64)u * blockDim.x;
			if(i < N) nt_store(mk<V>(re[u], im[u]), &a[i]);
		}
	}
}

hipError_t launch_diffuse_chain(const Launch &L, void *a, unsigned n, const double *dm2, unsigned nsteps, double zp, double zm,
                                int zero_ok) {
	const u64 N = 1ull << n;
	u64 blocks = (N + 1023) / 1024;
	if(blocks > 262144) blocks = 262144;
	if(L.prec == 0)
		hipLaunchKernelGGL(k_diffuse_chain<double>, dim3((unsigned)blocks), dim3(256), 0, L.stream, (double2 *)a, N, (const double2 *)dm2, nsteps,
		                   zp, zm, zero_ok);
	else
		hipLaunchKernelGGL(k_diffuse_chain<float>, dim3((unsigned)blocks), dim3(256), 0, L.stream, (float2 *)a, N, (const double2 *)dm2, nsteps,
		                   zp, zm, zero_ok);
	return hipGetLastError();
}

template <typename R>
__global__ void k_scatter(typename V2T<R>::t *__restrict__ a, const Scatter *__restrict__ sc, unsigned count) {
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < count) a[sc[i].index] = mk<typename V2T<R>::t>(sc[i].re, sc[i].im);
}

// ------------------------------------------------------------------------
// Reductions (FP64 accumulation, deterministic fixed-shape trees).
__device__ __forceinline__ double wave_sum(double v) {
	for(int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
	return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
	for(int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
	return v;
}

// Block of 256 threads: returns the sum in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh) {
	v = wave_sum(v);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if(lane == 0) sh[w] = v;
	__syncthreads();
	double r = 0;
	if(threadIdx.x == 0)
		for(int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
	__syncthreads();
	return r;
}

__device__ __forceinline__ long long block_sum_ll(long long v, long long *sh) {
	v = wave_sum_ll(v);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if(lane == 0) sh[w] = v;
	__syncthreads();
	long long r = 0;
	if(threadIdx.x == 0)
		for(int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
	__syncthreads();
	return r;
}

template <typename R>
__global__ void __launch_bounds__(256) k_sum(const typename V2T<R>::t *__restrict__ a, u64 N, double *part) {
	__shared__ double sh[8];
	double sr = 0, si = 0;
	for(u64 i = gtid(); i < N; i += gstride()) {
		const auto v = a[i];
		sr += (double)v.x;
		si += (double)v.y;
	}
	sr = block_sum(sr, sh);
	si = block_sum(si, sh);
	if(threadIdx.x == 0) {
		part[2 * blockIdx.x] = sr;
		part[2 * blockIdx.x + 1] = si;
	}
}

// Logical amplitude (scale*a + bias with the reference's multiply-then-add,
// quantum_register.cpp:161-166) and its probability re^2 + im^2.
__device__ __forceinline__ double prob_of(double re, double im, const double *af) {
	if(af) {
		const double mr = af[0] * re - af[1] * im;
		const double mi = af[0] * im + af[1] * re;
		re = mr + af[2];
		im = mi + af[3];
	}
	return re * re + im * im;
}

template <typename R>
__global__ void __launch_bounds__(256) k_norm(const typename V2T<R>::t *__restrict__ a, u64 N, const double *af, double *part) {
	__shared__ double sh[8];
	double s = 0;
	for(u64 i = gtid(); i < N; i += gstride()) {
		const auto v = a[i];
		s += prob_of(v.x, v.y, af);
	}
	s = block_sum(s, sh);
	if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// Fixed-order sum of the partials (count <= kReduceBlocks), one block of 64
// lanes: lane t sums partials t, t + 64, ... in order, then lane 0 sums the 64
// lane results in order, so every run rounds the same way.
__global__ void k_finish(const double *part, unsigned count, unsigned width, double *out) {
	__shared__ double sh[64];
	for(unsigned w = 0; w < width; ++w) {
		double s = 0;
		for(unsigned i = threadIdx.x; i < count; i += 64) s += part[width * i + w];
		sh[threadIdx.x] = s;
		__syncthreads();
		if(threadIdx.x == 0) {
			double t = 0;
			for(unsigned l = 0; l < 64; ++l) t += sh[l];
			out[w] = t;
		}
		__syncthreads();
	}
}

// Value fed to a sequential chain: 0 = probability |logical a|^2,
// 1 = real part, 2 = imaginary part.
__device__ __forceinline__ double chain_value(double re, double im, int sel, const double *af) {
	if(sel == 0) return prob_of(re, im, af);
	return sel == 1 ? re : im;
}

// Per-chunk FP64 sums of the chain values (chunk = 2^kSampleChunkLog states;
// fewer when N is smaller).
template <typename R>
__global__ void __launch_bounds__(256) k_chunk_sum(const typename V2T<R>::t *__restrict__ a, u64 chunk, u64 nchunks,
                                                   const double *af, int sel, double *sums) {
	__shared__ double sh[8];
	for(u64 c = blockIdx.x; c < nchunks; c += gridDim.x) {
		double s = 0;
		for(u64 i = threadIdx.x; i < chunk; i += blockDim.x) {
			const auto v = a[c * chunk + i];
			s += chain_value(v.x, v.y, sel, af);
		}
		s = block_sum(s, sh);
		if(threadIdx.x == 0) sums[c] = s;
	}
}

// Per-chunk integer increments of the sequential chain S = fl(S + x_i) on the
// grid u = 2^ulog[c]: while every partial sum stays inside one binade of that
// ulp, each step adds rint(x_i / u) ulps exactly, unless x_i / u ends in
// exactly one half (a tie, flagged).  Each thread walks a run of consecutive
// states so that the chunk's prefix minimum and maximum (in ulps, relative to
// the chunk start) come out in state order.
template <typename R>
__global__ void __launch_bounds__(256) k_chunk_ticks(const typename V2T<R>::t *__restrict__ a, u64 chunk, u64 nchunks,
                                                     const double *af, int sel, const int32_t *ulog, ChunkTicks *out) {
	__shared__ long long sh[256];
	__shared__ long long smin[256];
	__shared__ long long smax[256];
	__shared__ int sflag[256];
	const unsigned T = blockDim.x;
	const u64 per = chunk >= T ? chunk / T : 1;
	const bool active = chunk >= T || threadIdx.x < chunk;
	for(u64 c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const int32_t ul = ulog[c];
		long long cum = 0, lmin = LLONG_MAX, lmax = LLONG_MIN;
		int tie = 0, nz = 0;
		if(active) {
			const u64 first = c * chunk + (u64)threadIdx.x * per;
			for(u64 j = 0; j < per; ++j) {
				const auto v = a[first + j];
				const double x = chain_value(v.x, v.y, sel, af);
				nz |= (x != 0.0);
				if(ul != INT32_MIN) {
					const double q = ldexp(x, -ul);
					if(q - floor(q) == 0.5) tie = 1;
					cum += (long long)rint(q);
					lmin = cum < lmin ? cum : lmin;
					lmax = cum > lmax ? cum : lmax;
				}
			}
		}
		sh[threadIdx.x] = cum;
		smin[threadIdx.x] = lmin;
		smax[threadIdx.x] = lmax;
		sflag[threadIdx.x] = tie | (nz << 1);
		__syncthreads();
		if(threadIdx.x == 0) {
			long long run = 0, gmin = LLONG_MAX, gmax = LLONG_MIN;
			int anyTie = 0, anyNz = 0;
			for(unsigned t = 0; t < T; ++t) {
				if(smin[t] != LLONG_MAX) {
					gmin = run + smin[t] < gmin ? run + smin[t] : gmin;
					gmax = run + smax[t] > gmax ? run + smax[t] : gmax;
				}
				run += sh[t];
				anyTie |= sflag[t] & 1;
				anyNz |= sflag[t] >> 1;
			}
			ChunkTicks o;
			o.ticks = run;
			o.pmin = gmin == LLONG_MAX ? 0 : gmin;
			o.pmax = gmax == LLONG_MIN ? 0 : gmax;
			o.tie = anyTie;
			o.nonzero = anyNz;
			out[c] = o;
		}
		__syncthreads();
	}
}

// ------------------------------------------------------------------------
// Two-state comparison (the reference's experiment::compareStates metrics,
// comparison.cpp:12-63, plus the overlap <r|c>): per-block partials of sums
// and maxima; the host combines the blocks.
__device__ __forceinline__ void block_max_idx(double &v, u64 &idx, double *shv, u64 *shi) {
	for(int off = 32; off > 0; off >>= 1) {
		const double ov = __shfl_down(v, off, 64);
		const u64 oi = __shfl_down(idx, off, 64);
		if(ov > v || (ov == v && oi < idx)) {
			v = ov;
			idx = oi;
		}
	}
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if(lane == 0) {
		shv[w] = v;
		shi[w] = idx;
	}
	__syncthreads();
	if(threadIdx.x == 0)
		for(int i = 1; i < (int)(blockDim.x >> 6); ++i)
			if(shv[i] > v || (shv[i] == v && shi[i] < idx)) {
				v = shv[i];
				idx = shi[i];
			}
	__syncthreads();
}

template <typename RA, typename RB>
__global__ void __launch_bounds__(256) k_compare(const typename V2T<RA>::t *__restrict__ r, const typename V2T<RB>::t *__restrict__ c, u64 N,
                                                 CompareBlock *out) {
	__shared__ double sh[8];
	__shared__ double shv[8];
	__shared__ u64 shi[8];
	double d2 = 0, r2 = 0, c2 = 0, ovr = 0, ovi = 0, neq = 0;
	double maxd = -1, maxcomp = 0, maxdp = 0;
	u64 worst = 0;
	for(u64 i = gtid(); i < N; i += gstride()) {
		const auto a = r[i];
		const auto b = c[i];
		const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
		if(ar != br || ai != bi) neq += 1;
		const double dr = br - ar, di = bi - ai;
		const double dd = dr * dr + di * di;
		const double ae = sqrt(dd);
		if(ae > maxd) {
			maxd = ae;
			worst = i;
		}
		maxcomp = fmax(maxcomp, fmax(fabs(dr), fabs(di)));
		d2 += dd;
		const double pa = ar * ar + ai * ai, pb = br * br + bi * bi;
		r2 += pa;
		c2 += pb;
		maxdp = fmax(maxdp, fabs(pb - pa));
		ovr += ar * br + ai * bi;
		ovi += ar * bi - ai * br;
	}
	d2 = block_sum(d2, sh);
	r2 = block_sum(r2, sh);
	c2 = block_sum(c2, sh);
	ovr = block_sum(ovr, sh);
	ovi = block_sum(ovi, sh);
	neq = block_sum(neq, sh);
	block_max_idx(maxd, worst, shv, shi);
	double tmp = maxcomp;
	u64 dummy = 0;
	block_max_idx(tmp, dummy, shv, shi);
	maxcomp = tmp;
	tmp = maxdp;
	block_max_idx(tmp, dummy, shv, shi);
	maxdp = tmp;
	if(threadIdx.x == 0) {
		CompareBlock &o = out[blockIdx.x];
		o.d2 = d2;
		o.r2 = r2;
		o.c2 = c2;
		o.ovr = ovr;
		o.ovi = ovi;
		o.neq = neq;
		o.maxd = maxd;
		o.worst = worst;
		o.maxcomp = maxcomp;
		o.maxdp = maxdp;
	}
}

hipError_t launch_compare(const Launch &L, int prec_ref, const void *ref, int prec_cand, const void *cand, unsigned n,
                          CompareBlock *dout, unsigned *blocks) {
	const u64 N = 1ull << n;
	const unsigned g = grid_for(N, 256, kReduceBlocks);
	*blocks = g;
	if(prec_ref == 0 && prec_cand == 0)
		hipLaunchKernelGGL((k_compare<double, double>), dim3(g), dim3(256), 0, L.stream, (const double2 *)ref, (const double2 *)cand, N, dout);
	else if(prec_ref == 0 && prec_cand == 1)
		hipLaunchKernelGGL((k_compare<double, float>), dim3(g), dim3(256), 0, L.stream, (const double2 *)ref, (const float2 *)cand, N, dout);
	else if(prec_ref == 1 && prec_cand == 0)
		hipLaunchKernelGGL((k_compare<float, double>), dim3(g), dim3(256), 0, L.stream, (const float2 *)ref, (const double2 *)cand, N, dout);
	else
		hipLaunchKernelGGL((k_compare<float, float>), dim3(g), dim3(256), 0, L.stream, (const float2 *)ref, (const float2 *)cand, N, dout);
	return hipGetLastError();
}

// ------------------------------------------------------------------------
// Tiled multi-operation pass.
__device__ __forceinline__ double2 cmul(double2 p, double2 q) {
	return make_double2(fma(p.x, q.x, -p.y * q.y), fma(p.x, q.y, p.y * q.x));
}

// e^{2 pi i F / 2^64} from five 256-entry tables of the top 40 bits of F.
__device__ __forceinline__ double2 phase_of(u64 F, const double2 *__restrict__ tab) {
	double2 p = tab[(F >> 56) & 255u];
	p = cmul(p, tab[256 + ((F >> 48) & 255u)]);
	p = cmul(p, tab[512 + ((F >> 40) & 255u)]);
	p = cmul(p, tab[768 + ((F >> 32) & 255u)]);
	p = cmul(p, tab[1024 + ((F >> 24) & 255u)]);
	return p;
}

// Same, over the table chunks [j0, j1] only (F has no bits outside them).
__device__ __forceinline__ double2 phase_of_range(u64 F, const double2 *__restrict__ tab, unsigned j0, unsigned j1) {
	double2 p = tab[j0 * 256u + ((F >> (56u - 8u * j0)) & 255u)];
	for(unsigned j = j0 + 1; j <= j1; ++j) p = cmul(p, tab[j * 256u + ((F >> (56u - 8u * j)) & 255u)]);
	return p;
}

// Rotation by a phase evaluated in FP64; complex64 amplitudes take it rounded
// to FP32 and rotate in FP32 (the bulk arithmetic of the complex64 form).
template <typename V> __device__ __forceinline__ V rot(V x, double2 p) {
	const double xr = x.x, xi = x.y;
	return mk<V>(fma(xr, p.x, -xi * p.y), fma(xr, p.y, xi * p.x));
}
template <> __device__ __forceinline__ float2 rot<float2>(float2 x, double2 p) {
	const float pr = (float)p.x, pi = (float)p.y;
	return make_float2(fmaf(x.x, pr, -x.y * pi), fmaf(x.x, pi, x.y * pr));
}

// A phase held in the amplitudes' own precision (rounded once where it is
// computed, for phases kept in registers across a loop), and the rotation by
// it: rotp(x, LanePhase<V>::of(p)) equals rot<V>(x, p) bit for bit.
template <typename V> struct LanePhase {
	using t = double2;
	static __device__ __forceinline__ t of(double2 p) { return p; }
};
template <> struct LanePhase<float2> {
	using t = float2;
	static __device__ __forceinline__ t of(double2 p) { return make_float2((float)p.x, (float)p.y); }
};
__device__ __forceinline__ double2 rotp(double2 x, double2 p) { return rot<double2>(x, p); }
__device__ __forceinline__ float2 rotp(float2 x, float2 p) {
	return make_float2(fmaf(x.x, p.x, -x.y * p.y), fmaf(x.x, p.y, x.y * p.x));
}

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

// Unscaled butterfly (x + y, x - y): transforms that apply several Hadamards
// in a row scale once at the end, by an exact power of two.
template <typename V> __device__ __forceinline__ void butterfly(V &x, V &y) {
	const V s = x;
	x = mk<V>(s.x + y.x, s.y + y.y);
	y = mk<V>(s.x - y.x, s.y - y.y);
}

__device__ __forceinline__ double2 csqr(double2 p) { return make_double2(fma(p.x, p.x, -p.y * p.y), 2.0 * p.x * p.y); }

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
template <typename R>
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
				case 5: tile_uk<R, 5>(tile, nloc, op, mats); break;
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

template <typename V, typename R, int W, int J>
__device__ __forceinline__ void reg_h(V *v, R hc) {
#pragma unroll
	for(int r = 0; r < (1 << W); ++r) {
		if(r & (1 << J)) continue;
		const V x = v[r], y = v[r | (1 << J)];
		if(is_zero(x) && is_zero(y)) continue;  // dense_state_backend.cpp:263 skips all-zero pairs
		V o0, o1;
		o0.x = (x.x + y.x) * hc;
		o0.y = (x.y + y.y) * hc;
		o1.x = (x.x - y.x) * hc;
		o1.y = (x.y - y.y) * hc;
		v[r] = o0;
		v[r | (1 << J)] = o1;
	}
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

template <typename V, typename R, int W, int J>
__device__ __forceinline__ void reg_u1(V *v, const double *g) {
	const R g00r = (R)g[0], g00i = (R)g[1], g01r = (R)g[2], g01i = (R)g[3];
	const R g10r = (R)g[4], g10i = (R)g[5], g11r = (R)g[6], g11i = (R)g[7];
#pragma unroll
	for(int r = 0; r < (1 << W); ++r) {
		if(r & (1 << J)) continue;
		const V x = v[r], y = v[r | (1 << J)];
		if(is_zero(x) && is_zero(y)) continue;
		V o0, o1;
		o0.x = g00r * x.x - g00i * x.y + g01r * y.x - g01i * y.y;
		o0.y = g00r * x.y + g00i * x.x + g01r * y.y + g01i * y.x;
		o1.x = g10r * x.x - g10i * x.y + g11r * y.x - g11i * y.y;
		o1.y = g10r * x.y + g10i * x.x + g11r * y.y + g11i * y.x;
		v[r] = o0;
		v[r | (1 << J)] = o1;
	}
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

// One tile of a staged pass: loads it into LDS, runs every stage, stores it.
// ta/tb/hitab/gacc are the workgroup's LDS buffers; hitab must hold P's table.
template <typename R, int W, bool DUAL, bool LDSOPS>
__device__ __forceinline__ void staged_tile(typename V2T<R>::t *__restrict__ a, typename V2T<R>::t *__restrict__ b,
                                            const TilePass &P, u64 tid, const TileOp *__restrict__ ops,
                                            const TileStage *__restrict__ stages, unsigned nstages,
                                            const double *__restrict__ mats, const TwiddleTerm *__restrict__ terms,
                                            const double2 *__restrict__ ptab, typename V2T<R>::t *ta,
                                            typename V2T<R>::t *tb, const u64 *hitab, double *gacc) {
	using V = typename V2T<R>::t;
	constexpr int NR = 1 << W;
	constexpr bool PAD = !LDSOPS;
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
		if(LDSOPS && st.lds) {  // k-qubit matrices: their registers would cap every other stage's occupancy
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
				}
				tile_apply<R>(ta, op, nloc, tbase, C, hitab, mats, terms, ptab);
				if(DUAL) tile_apply<R>(tb, op, nloc, tbase, C, hitab, mats, terms, ptab);
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
			for(unsigned o = 0; o < st.nops; ++o) {
				const TileOp op = ops[st.op0 + o];
				if((tbase & op.gcond) != op.gcond) continue;  // uniform across the block
				if(op.kind == TOP_GRAD) {
					if(DUAL) {
						double acc = 0.0;
						if(act) {
#pragma unroll
							for(int r = 0; r < NR; ++r)
								if((lidx[r] & op.lcond) == op.lcond)
									acc += (double)vb[r].x * (double)va[r].y - (double)vb[r].y * (double)va[r].x;
						}
						acc = wave_sum(acc);
						if(lane == 0) gacc[wave * kMaxGradSlots + op.mat] += acc;
					}
					continue;
				}
				if(!act) continue;
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

template <typename R, int W, bool DUAL, bool LDSOPS>
__global__ void __launch_bounds__(256) k_tile_staged(typename V2T<R>::t *__restrict__ a, typename V2T<R>::t *__restrict__ b,
                                                     TilePass P, const TileOp *__restrict__ ops,
                                                     const TileStage *__restrict__ stages, unsigned nstages,
                                                     const double *__restrict__ mats, const TwiddleTerm *__restrict__ terms,
                                                     const double2 *__restrict__ ptab, double *__restrict__ gpart,
                                                     unsigned nslots) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	const unsigned nlocp = staged_tile_elems(P.T, LDSOPS);
	V *ta = (V *)smem;
	V *tb = ta + nlocp;  // second tile (dual passes only)
	u64 *hitab = (u64 *)(ta + (DUAL ? 2u : 1u) * nlocp);
	double *gacc = (double *)(hitab + (1u << P.nh));  // dual: [4 waves][kMaxGradSlots]
	tile_hitab<R>(hitab, P);
	if(DUAL)
		for(unsigned i = threadIdx.x; i < 4 * kMaxGradSlots; i += blockDim.x) gacc[i] = 0.0;
	for(u64 tid = blockIdx.x; tid < P.ntiles; tid += gridDim.x)
		staged_tile<R, W, DUAL, LDSOPS>(a, b, P, tid, ops, stages, nstages, mats, terms, ptab, ta, tb, hitab, gacc);
	if(DUAL) {
		__syncthreads();
		for(unsigned s = threadIdx.x; s < nslots; s += blockDim.x)
			gpart[(u64)s * gridDim.x + blockIdx.x] =
			    ((gacc[s] + gacc[kMaxGradSlots + s]) + gacc[2 * kMaxGradSlots + s]) + gacc[3 * kMaxGradSlots + s];
	}
}

// ------------------------------------------------------------------------
// Batched circuits: one launch per round applies one pass of every circuit
// that has one, as a work list of (pass, tile range) items that persistent
// workgroups walk.
template <typename R, int W>
__global__ void __launch_bounds__(256) k_batch_tiles(const BatchPass *__restrict__ passes, const BatchItem *__restrict__ items,
                                                     unsigned nitems, const TileOp *__restrict__ ops,
                                                     const TileStage *__restrict__ stages, const double2 *__restrict__ ptab) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	V *ta = (V *)smem;
	u64 *hitab = (u64 *)(ta + staged_tile_elems(batch_max_tile_bits(sizeof(R) == 8 ? 0 : 1), false));
	for(unsigned it = blockIdx.x; it < nitems; it += gridDim.x) {
		const BatchItem item = items[it];
		const BatchPass &bp = passes[item.pass];  // read in place: its tables are indexed at run time
		__syncthreads();  // the previous item's tiles are done with hitab
		tile_hitab<R>(hitab, bp.tp);
		for(u64 tid = item.tile0; tid < item.tile0 + item.ntiles; ++tid)
			staged_tile<R, W, false, false>((V *)bp.state, nullptr, bp.tp, tid, ops + bp.op0, stages + bp.st0, bp.nst, nullptr,
			                                nullptr, ptab, ta, nullptr, hitab, nullptr);
	}
}

// Dual form of k_batch_tiles for the batched adjoint: each pass steps psi
// (state) and lambda (state2) back together; a work item's gradient partials
// (per wavefront, summed in a fixed order) go to its own gpart row.
template <typename R, int W>
__global__ void __launch_bounds__(256) k_batch_dual(const BatchPass *__restrict__ passes, const BatchItem *__restrict__ items,
                                                    unsigned nitems, const TileOp *__restrict__ ops,
                                                    const TileStage *__restrict__ stages, double *__restrict__ gpart) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	const unsigned nlocp = staged_tile_elems(batch_max_tile_bits(sizeof(R) == 8 ? 0 : 1) - 1, false);
	V *ta = (V *)smem;
	V *tb = ta + nlocp;
	u64 *hitab = (u64 *)(tb + nlocp);
	double *gacc = (double *)(hitab + (1u << kMaxHighBits));  // [4 waves][kMaxGradSlots]
	for(unsigned it = blockIdx.x; it < nitems; it += gridDim.x) {
		const BatchItem item = items[it];
		const BatchPass &bp = passes[item.pass];
		__syncthreads();  // the previous item is done with hitab and gacc
		tile_hitab<R>(hitab, bp.tp);
		for(unsigned i = threadIdx.x; i < 4 * kMaxGradSlots; i += blockDim.x) gacc[i] = 0.0;
		for(u64 tid = item.tile0; tid < item.tile0 + item.ntiles; ++tid)
			staged_tile<R, W, true, false>((V *)bp.state, (V *)bp.state2, bp.tp, tid, ops + bp.op0, stages + bp.st0, bp.nst, nullptr,
			                               nullptr, nullptr, ta, tb, hitab, gacc);
		__syncthreads();
		for(unsigned s = threadIdx.x; s < bp.nslots; s += blockDim.x)
			gpart[(u64)it * kMaxGradSlots + s] =
			    ((gacc[s] + gacc[kMaxGradSlots + s]) + gacc[2 * kMaxGradSlots + s]) + gacc[3 * kMaxGradSlots + s];
	}
}

__global__ void __launch_bounds__(256) k_batch_grad_finish(const BatchGrad *__restrict__ g, unsigned count,
                                                           const double *__restrict__ gpart, double scale,
                                                           double *__restrict__ grads) {
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= count) return;
	const BatchGrad e = g[i];
	double v = 0.0;
	for(unsigned k = 0; k < e.nitems; ++k) v += gpart[(u64)(e.item0 + k) * kMaxGradSlots + e.slot];
	grads[e.gindex] = scale * v;
}

// Basis-state initialisation of every circuit (items: chunks of one state).
template <typename R>
__global__ void __launch_bounds__(256) k_batch_init(const BatchChunk *__restrict__ chunks, unsigned nchunks) {
	using V = typename V2T<R>::t;
	for(unsigned c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const BatchChunk ch = chunks[c];
		V *s = (V *)ch.dst;
		for(u64 i = ch.first + threadIdx.x; i < ch.first + ch.count; i += blockDim.x)
			s[i] = i == ch.init ? mk<V>(1.0, 0.0) : mk<V>(0.0, 0.0);
	}
}

// Final bit layout of every circuit whose swaps were relabeled:
// dst[x] = src[y], bit map[b] of y = bit b of x.
template <typename R>
// The source index is assembled from per-byte deposit tables built in LDS
// for each chunk: y = tab[0][x & 255] | tab[1][(x >> 8) & 255] | ...
__global__ void __launch_bounds__(256) k_batch_permute(const BatchChunk *__restrict__ chunks, unsigned nchunks) {
	using V = typename V2T<R>::t;
	__shared__ u64 tab[5][256];
	__shared__ uint8_t map[64];
	for(unsigned c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const BatchChunk &ch = chunks[c];
		const unsigned n = ch.n;
		__syncthreads();  // the previous chunk is done with the tables
		if(threadIdx.x < 64) map[threadIdx.x] = threadIdx.x < n ? ch.map[threadIdx.x] : 0;
		__syncthreads();
		for(unsigned t = threadIdx.x; t < 5 * 256; t += blockDim.x) {
			const unsigned j = t >> 8, v = t & 255u;
			u64 d = 0;
			for(unsigned b = 0; b < 8; ++b)
				if(((v >> b) & 1u) && 8 * j + b < n) d |= 1ull << map[8 * j + b];
			tab[j][v] = d;
		}
		__syncthreads();
		const V *src = (const V *)ch.src;
		V *dst = (V *)ch.dst;
		const u64 end = ch.first + ch.count;
		auto gather = [&](u64 x) {
			return src[tab[0][x & 255u] | tab[1][(x >> 8) & 255u] | tab[2][(x >> 16) & 255u] | tab[3][(x >> 24) & 255u] |
			           tab[4][(x >> 32) & 255u]];
		};
		for(u64 x0 = ch.first + threadIdx.x; x0 < end; x0 += 8 * blockDim.x) {
			if(x0 + 7 * blockDim.x < end) {  // eight gathers in flight
				V v[8];
#pragma unroll
				for(unsigned u = 0; u < 8; ++u) v[u] = gather(x0 + u * blockDim.x);
#pragma unroll
				for(unsigned u = 0; u < 8; ++u) dst[x0 + u * blockDim.x] = v[u];
			} else {
				for(u64 x = x0; x < end; x += blockDim.x) dst[x] = gather(x);
			}
		}
	}
}

// grads[gidx[s]] = scale * sum over the blocks of gpart[s][.] (one block per slot).
__global__ void __launch_bounds__(256) k_grad_finish(const double *__restrict__ gpart, unsigned blocks,
                                                     const uint32_t *__restrict__ gidx, double scale, double *__restrict__ grads) {
	__shared__ double sh[8];
	const unsigned s = blockIdx.x;
	double v = 0.0;
	for(unsigned i = threadIdx.x; i < blocks; i += blockDim.x) v += gpart[(u64)s * blocks + i];
	v = block_sum(v, sh);
	if(threadIdx.x == 0) grads[gidx[s]] = scale * v;
}

// Diagonal observable O(x) = sum_z coef_z (-1)^popcount(x & mask_z)
// + sum_p coef_p [x == state_p]: per-block partials of sum O(x)|a_x|^2 and,
// when lam is given, lam = O a.
template <typename R>
__global__ void __launch_bounds__(256) k_observable(const typename V2T<R>::t *__restrict__ a, typename V2T<R>::t *__restrict__ lam,
                                                    u64 N, ObsParams O, double *__restrict__ part) {
	using V = typename V2T<R>::t;
	__shared__ double sh[8];
	double acc = 0.0;
	for(u64 i = gtid(); i < N; i += gstride()) {
		double o = 0.0;
		for(unsigned t = 0; t < O.nz; ++t) o += (__popcll(i & O.zmask[t]) & 1) ? -O.zcoef[t] : O.zcoef[t];
		for(unsigned t = 0; t < O.np; ++t)
			if(i == O.pstate[t]) o += O.pcoef[t];
		const V v = a[i];
		const double re = v.x, im = v.y;
		acc += o * (re * re + im * im);
		if(lam) lam[i] = mk<V>(o * re, o * im);
	}
	acc = block_sum(acc, sh);
	if(threadIdx.x == 0) part[blockIdx.x] = acc;
}

// ------------------------------------------------------------------------
// Host launch helpers.
#define TMFQS_LAUNCH(KERNEL, GRID, BLOCK, SHMEM, STREAM, ...) \
	hipLaunchKernelGGL(KERNEL, dim3(GRID), dim3(BLOCK), SHMEM, STREAM, __VA_ARGS__)

hipError_t launch_hadamard(const Launch &L, void *a, unsigned n, unsigned bit) {
	const u64 np = (1ull << n) >> 1;
	const double c = 1.0 / sqrt(2.0);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_hadamard<double>, grid_for(np, 256, kStreamGrid), 256, 0, L.stream, (double2 *)a, np, bit, c);
	else
		TMFQS_LAUNCH(k_hadamard<float>, grid_for(np, 256, kStreamGrid), 256, 0, L.stream, (float2 *)a, np, bit, (float)c);
	return hipGetLastError();
}

hipError_t launch_pauli_x(const Launch &L, void *a, unsigned n, unsigned bit, u64 cmask) {
	const u64 np = (1ull << n) >> 1;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_pauli_x<double>, grid_for(np, 256, kStreamGrid), 256, 0, L.stream, (double2 *)a, np, bit, cmask);
	else
		TMFQS_LAUNCH(k_pauli_x<float>, grid_for(np, 256, kStreamGrid), 256, 0, L.stream, (float2 *)a, np, bit, cmask);
	return hipGetLastError();
}

hipError_t launch_cphase(const Launch &L, void *a, unsigned n, unsigned tbit, u64 cmask, double cr, double ci) {
	unsigned cbit = 0;
	while(!((cmask >> cbit) & 1ull)) ++cbit;
	const unsigned lo = tbit < cbit ? tbit : cbit, hi = tbit < cbit ? cbit : tbit;
	const u64 nq = (1ull << n) >> 2;
	const u64 setm = (1ull << tbit) | cmask;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_cphase<double>, grid_for(nq, 256, kStreamGrid), 256, 0, L.stream, (double2 *)a, nq, lo, hi, setm, cr, ci);
	else
		TMFQS_LAUNCH(k_cphase<float>, grid_for(nq, 256, kStreamGrid), 256, 0, L.stream, (float2 *)a, nq, lo, hi, setm, (float)cr, (float)ci);
	return hipGetLastError();
}

hipError_t launch_u1(const Launch &L, void *a, unsigned n, unsigned bit, const double *m) {
	const u64 np = (1ull << n) >> 1;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_u1<double>, grid_for(np, 256, kStreamGrid), 256, 0, L.stream, (double2 *)a, np, bit, m[0], m[1], m[2], m[3], m[4], m[5],
		             m[6], m[7]);
	else
		TMFQS_LAUNCH(k_u1<float>, grid_for(np, 256, kStreamGrid), 256, 0, L.stream, (float2 *)a, np, bit, (float)m[0], (float)m[1],
		             (float)m[2], (float)m[3], (float)m[4], (float)m[5], (float)m[6], (float)m[7]);
	return hipGetLastError();
}

template <typename R>
static void launch_uk_t(const Launch &L, void *a, unsigned n, unsigned k, const unsigned *b, const double *dmat) {
	using V = typename V2T<R>::t;
	const u64 nb = (1ull << n) >> k;
	// 8192 workgroups: each stages the matrix in LDS first, and larger grids
	// measured up to 11% slower on 30 qubits.
	const unsigned g = grid_for(nb);
	switch(k) {
		case 2: TMFQS_LAUNCH((k_uk<R, 2>), g, 256, 0, L.stream, (V *)a, nb, dmat, b[0], b[1], 0u, 0u, 0u); break;
		case 3: TMFQS_LAUNCH((k_uk<R, 3>), g, 256, 0, L.stream, (V *)a, nb, dmat, b[0], b[1], b[2], 0u, 0u); break;
		case 4: TMFQS_LAUNCH((k_uk<R, 4>), g, 256, 0, L.stream, (V *)a, nb, dmat, b[0], b[1], b[2], b[3], 0u); break;
		case 5: TMFQS_LAUNCH((k_uk<R, 5>), g, 256, 0, L.stream, (V *)a, nb, dmat, b[0], b[1], b[2], b[3], b[4]); break;
		default: break;
	}
}

hipError_t launch_uk(const Launch &L, void *a, unsigned n, unsigned k, const unsigned *bits, const double *dmat) {
	if(L.prec == 0)
		launch_uk_t<double>(L, a, n, k, bits, dmat);
	else
		launch_uk_t<float>(L, a, n, k, bits, dmat);
	return hipGetLastError();
}

hipError_t launch_bitswap(const Launch &L, void *a, unsigned n, unsigned p, unsigned q) {
	if(p > q) {
		const unsigned t = p;
		p = q;
		q = t;
	}
	const u64 nq = (1ull << n) >> 2;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_bitswap<double>, grid_for(nq, 256, kStreamGrid), 256, 0, L.stream, (double2 *)a, nq, p, q);
	else
		TMFQS_LAUNCH(k_bitswap<float>, grid_for(nq, 256, kStreamGrid), 256, 0, L.stream, (float2 *)a, nq, p, q);
	return hipGetLastError();
}

hipError_t launch_involution(const Launch &L, void *a, unsigned n, const unsigned *perm, unsigned C) {
	InvParams P{};
	P.n = n;
	P.C = C;
	bool inS[64] = {false};
	for(unsigned b = 0; b < C; ++b) {
		inS[b] = true;
		inS[perm[b]] = true;
	}
	unsigned local[64];
	unsigned s = 0;
	for(unsigned b = 0; b < n; ++b)
		if(inS[b]) local[s++] = b;  // ascending; the first C are 0..C-1
	P.s = s;
	P.nh = s - C;
	for(unsigned i = 0; i < P.nh; ++i) P.hbits[i] = local[C + i];
	for(unsigned b = 0; b < n; ++b) P.perm[b] = perm[b];
	for(unsigned i = 0; i < s; ++i) {
		const unsigned img = perm[local[i]];
		unsigned j = 0;
		while(local[j] != img) ++j;
		P.lperm[i] = j;
	}
	P.npairs_tiles = 1ull << (n - s);
	const size_t esz = L.prec == 0 ? 16 : 8;
	const size_t shmem = 2 * esz * (1ull << s);
	const unsigned grid = (unsigned)(P.npairs_tiles < 16384 ? P.npairs_tiles : 16384);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_involution<double>, grid, 256, shmem, L.stream, (double2 *)a, P);
	else
		TMFQS_LAUNCH(k_involution<float>, grid, 256, shmem, L.stream, (float2 *)a, P);
	return hipGetLastError();
}

hipError_t launch_ew_chain(const Launch &L, void *a, unsigned n, const EwStep *dsteps, unsigned nsteps) {
	const u64 N = 1ull << n;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_ew_chain<double>, grid_for(N, 256, kStreamGrid), 256, 0, L.stream, (double2 *)a, N, dsteps, nsteps);
	else
		TMFQS_LAUNCH(k_ew_chain<float>, grid_for(N, 256, kStreamGrid), 256, 0, L.stream, (float2 *)a, N, dsteps, nsteps);
	return hipGetLastError();
}

hipError_t launch_scatter(const Launch &L, void *a, const Scatter *dsc, unsigned count) {
	if(count == 0) return hipSuccess;
	const unsigned g = (count + 63) / 64;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_scatter<double>, g, 64, 0, L.stream, (double2 *)a, dsc, count);
	else
		TMFQS_LAUNCH(k_scatter<float>, g, 64, 0, L.stream, (float2 *)a, dsc, count);
	return hipGetLastError();
}

hipError_t launch_sum(const Launch &L, const void *a, unsigned n, double *dpart, double *dout) {
	const u64 N = 1ull << n;
	const unsigned g = grid_for(N, 256, kReduceBlocks);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_sum<double>, g, 256, 0, L.stream, (const double2 *)a, N, dpart);
	else
		TMFQS_LAUNCH(k_sum<float>, g, 256, 0, L.stream, (const float2 *)a, N, dpart);
	TMFQS_LAUNCH(k_finish, 1, 64, 0, L.stream, dpart, g, 2u, dout);
	return hipGetLastError();
}

hipError_t launch_norm(const Launch &L, const void *a, unsigned n, const double *affine4, double *dpart, double *dout) {
	const u64 N = 1ull << n;
	const unsigned g = grid_for(N, 256, kReduceBlocks);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_norm<double>, g, 256, 0, L.stream, (const double2 *)a, N, affine4, dpart);
	else
		TMFQS_LAUNCH(k_norm<float>, g, 256, 0, L.stream, (const float2 *)a, N, affine4, dpart);
	TMFQS_LAUNCH(k_finish, 1, 64, 0, L.stream, dpart, g, 1u, dout);
	return hipGetLastError();
}

static void chunk_geometry(unsigned n, u64 &chunk, u64 &nchunks) {
	const unsigned cl = n < kSampleChunkLog ? n : kSampleChunkLog;
	chunk = 1ull << cl;
	nchunks = 1ull << (n - cl);
}

hipError_t launch_chunk_sum(const Launch &L, const void *a, unsigned n, const double *affine4, int sel, double *dsums) {
	u64 chunk, nch;
	chunk_geometry(n, chunk, nch);
	const unsigned g = (unsigned)(nch < 16384 ? nch : 16384);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_chunk_sum<double>, g, 256, 0, L.stream, (const double2 *)a, chunk, nch, affine4, sel, dsums);
	else
		TMFQS_LAUNCH(k_chunk_sum<float>, g, 256, 0, L.stream, (const float2 *)a, chunk, nch, affine4, sel, dsums);
	return hipGetLastError();
}

hipError_t launch_chunk_ticks(const Launch &L, const void *a, unsigned n, const double *affine4, int sel, const int32_t *dulog,
                              ChunkTicks *dout) {
	u64 chunk, nch;
	chunk_geometry(n, chunk, nch);
	const unsigned g = (unsigned)(nch < 16384 ? nch : 16384);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_chunk_ticks<double>, g, 256, 0, L.stream, (const double2 *)a, chunk, nch, affine4, sel, dulog, dout);
	else
		TMFQS_LAUNCH(k_chunk_ticks<float>, g, 256, 0, L.stream, (const float2 *)a, chunk, nch, affine4, sel, dulog, dout);
	return hipGetLastError();
}

unsigned tile_stage_width(int prec, unsigned T, bool dual) {
	// 16 amplitudes per thread (8 per state when dual).  Tiles one bit narrower
	// than the default leave LDS for four workgroups per CU, which only
	// three-bit stages (fewer registers) can use: there, 8 per thread.
	const unsigned def = prec == 0 ? 12u : 13u;
	const unsigned w = dual ? 3u : (T >= def ? 4u : 3u);
	return T < w ? T : w;
}

unsigned tile_grid(const TilePass &P, bool dual, int prec) {
	// dual: at most kMaxDualGrid gradient partials per slot.  Single-state passes
	// over 2^30 amplitudes (Hadamard layer, measured): complex128 fastest at
	// 65536 workgroups, complex64 at 32768 (6% faster than at 65536).
	const u64 cap = dual ? kMaxDualGrid : (prec == 0 ? 65536 : 32768);
	return (unsigned)(P.ntiles < cap ? P.ntiles : cap);
}

template <typename R, int W, bool DUAL, bool LDSOPS>
static hipError_t launch_staged_t(const Launch &L, void *a, void *b, const TilePass &P, const TileOp *dops, const TileStage *dstages,
                                  unsigned nstages, const double *dmats, const TwiddleTerm *dterms, const double *dphase,
                                  double *gpart, unsigned nslots) {
	using V = typename V2T<R>::t;
	const size_t nlocp = staged_tile_elems(P.T, LDSOPS);
	const size_t shmem = (DUAL ? 2 : 1) * sizeof(V) * nlocp + (8ull << P.nh) + (DUAL ? 4 * kMaxGradSlots * sizeof(double) : 0);
	if(shmem > 160 * 1024) return hipErrorInvalidValue;
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_tile_staged<R, W, DUAL, LDSOPS>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                          160 * 1024);
		attr = true;
	}
	TMFQS_LAUNCH((k_tile_staged<R, W, DUAL, LDSOPS>), tile_grid(P, DUAL, L.prec), 256, shmem, L.stream, (V *)a, (V *)b, P, dops, dstages, nstages,
	             dmats, dterms, (const double2 *)dphase, gpart, nslots);
	return hipGetLastError();
}

template <typename R, bool DUAL, bool LDSOPS>
static hipError_t launch_staged_w(const Launch &L, unsigned W, void *a, void *b, const TilePass &P, const TileOp *dops,
                                  const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                                  const double *dphase, double *gpart, unsigned nslots) {
#define TMFQS_STAGED(WW) \
	launch_staged_t<R, WW, DUAL, LDSOPS>(L, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots)
	switch(W) {
		case 1: return TMFQS_STAGED(1);
		case 2: return TMFQS_STAGED(2);
		case 3: return TMFQS_STAGED(3);
		case 4: return TMFQS_STAGED(4);
		default: return hipErrorInvalidValue;
	}
#undef TMFQS_STAGED
}

template <typename R>
static hipError_t launch_staged_r(const Launch &L, unsigned W, void *a, void *b, const TilePass &P, const TileOp *dops,
                                  const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                                  const double *dphase, double *gpart, unsigned nslots, bool ldsops) {
	if(b)
		return ldsops ? launch_staged_w<R, true, true>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots)
		              : launch_staged_w<R, true, false>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
	return ldsops ? launch_staged_w<R, false, true>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots)
	              : launch_staged_w<R, false, false>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

hipError_t launch_tile_staged(const Launch &L, void *a, void *b, const TilePass &P, unsigned W, const TileOp *dops,
                              const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                              const double *dphase, double *gpart, unsigned nslots, bool ldsops) {
	if(W > P.T || nslots > kMaxGradSlots) return hipErrorInvalidValue;
	if(L.prec == 0)
		return launch_staged_r<double>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots, ldsops);
	return launch_staged_r<float>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots, ldsops);
}

hipError_t launch_grad_finish(const Launch &L, const double *gpart, unsigned blocks, unsigned nslots, const uint32_t *gidx,
                              double scale, double *grads) {
	if(nslots == 0) return hipSuccess;
	TMFQS_LAUNCH(k_grad_finish, nslots, 256, 0, L.stream, gpart, blocks, gidx, scale, grads);
	return hipGetLastError();
}

hipError_t launch_observable(const Launch &L, const void *a, void *lam, unsigned n, const ObsParams &O, double *dpart,
                             double *dout) {
	const u64 N = 1ull << n;
	const unsigned g = grid_for(N, 256, kReduceBlocks);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_observable<double>, g, 256, 0, L.stream, (const double2 *)a, (double2 *)lam, N, O, dpart);
	else
		TMFQS_LAUNCH(k_observable<float>, g, 256, 0, L.stream, (const float2 *)a, (float2 *)lam, N, O, dpart);
	TMFQS_LAUNCH(k_finish, 1, 64, 0, L.stream, dpart, g, 1u, dout);
	return hipGetLastError();
}

hipError_t launch_batch_init(const Launch &L, const BatchChunk *dchunks, unsigned nchunks) {
	if(!nchunks) return hipSuccess;
	const unsigned g = nchunks < 8192 ? nchunks : 8192;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_batch_init<double>, g, 256, 0, L.stream, dchunks, nchunks);
	else
		TMFQS_LAUNCH(k_batch_init<float>, g, 256, 0, L.stream, dchunks, nchunks);
	return hipGetLastError();
}

hipError_t launch_batch_permute(const Launch &L, const BatchChunk *dchunks, unsigned nchunks) {
	if(!nchunks) return hipSuccess;
	const unsigned g = nchunks < 8192 ? nchunks : 8192;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_batch_permute<double>, g, 256, 0, L.stream, dchunks, nchunks);
	else
		TMFQS_LAUNCH(k_batch_permute<float>, g, 256, 0, L.stream, dchunks, nchunks);
	return hipGetLastError();
}

template <typename R, int W>
static hipError_t launch_batch_tiles_t(const Launch &L, const BatchPass *dpasses, const BatchItem *ditems, unsigned nitems,
                                       const TileOp *dops, const TileStage *dstages, const double *dphase, unsigned grid) {
	using V = typename V2T<R>::t;
	const size_t shmem = sizeof(V) * staged_tile_elems(batch_max_tile_bits(L.prec), false) + (8ull << kMaxHighBits);
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_batch_tiles<R, W>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	TMFQS_LAUNCH((k_batch_tiles<R, W>), grid, 256, shmem, L.stream, dpasses, ditems, nitems, dops, dstages,
	             (const double2 *)dphase);
	return hipGetLastError();
}

template <typename R, int W>
static hipError_t launch_batch_dual_t(const Launch &L, const BatchPass *dpasses, const BatchItem *ditems, unsigned nitems,
                                      const TileOp *dops, const TileStage *dstages, double *dgpart, unsigned grid) {
	using V = typename V2T<R>::t;
	const size_t shmem = 2 * sizeof(V) * staged_tile_elems(batch_max_tile_bits(L.prec) - 1, false) + (8ull << kMaxHighBits) +
	                     4 * kMaxGradSlots * sizeof(double);
	static bool attr = false;
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_batch_dual<R, W>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	TMFQS_LAUNCH((k_batch_dual<R, W>), grid, 256, shmem, L.stream, dpasses, ditems, nitems, dops, dstages, dgpart);
	return hipGetLastError();
}

hipError_t launch_batch_dual(const Launch &L, unsigned W, const BatchPass *dpasses, const BatchItem *ditems, unsigned nitems,
                             const TileOp *dops, const TileStage *dstages, double *dgpart, unsigned grid) {
	if(!nitems) return hipSuccess;
	if(grid > nitems) grid = nitems;
	if(L.prec == 0) {
		if(W == 3) return launch_batch_dual_t<double, 3>(L, dpasses, ditems, nitems, dops, dstages, dgpart, grid);
		if(W == 2) return launch_batch_dual_t<double, 2>(L, dpasses, ditems, nitems, dops, dstages, dgpart, grid);
	} else {
		if(W == 3) return launch_batch_dual_t<float, 3>(L, dpasses, ditems, nitems, dops, dstages, dgpart, grid);
		if(W == 2) return launch_batch_dual_t<float, 2>(L, dpasses, ditems, nitems, dops, dstages, dgpart, grid);
	}
	return hipErrorInvalidValue;
}

hipError_t launch_batch_grad_finish(const Launch &L, const BatchGrad *dg, unsigned count, const double *dgpart, double scale,
                                    double *dgrads) {
	if(!count) return hipSuccess;
	TMFQS_LAUNCH(k_batch_grad_finish, (count + 255) / 256, 256, 0, L.stream, dg, count, dgpart, scale, dgrads);
	return hipGetLastError();
}

hipError_t launch_batch_tiles(const Launch &L, unsigned W, const BatchPass *dpasses, const BatchItem *ditems, unsigned nitems,
                              const TileOp *dops, const TileStage *dstages, const double *dphase, unsigned grid) {
	if(!nitems) return hipSuccess;
	if(grid > nitems) grid = nitems;
	if(L.prec == 0) {
		if(W == 4) return launch_batch_tiles_t<double, 4>(L, dpasses, ditems, nitems, dops, dstages, dphase, grid);
		if(W == 3) return launch_batch_tiles_t<double, 3>(L, dpasses, ditems, nitems, dops, dstages, dphase, grid);
	} else {
		if(W == 4) return launch_batch_tiles_t<float, 4>(L, dpasses, ditems, nitems, dops, dstages, dphase, grid);
		if(W == 3) return launch_batch_tiles_t<float, 3>(L, dpasses, ditems, nitems, dops, dstages, dphase, grid);
	}
	return hipErrorInvalidValue;
}

}  // namespace tmfqs_hip
