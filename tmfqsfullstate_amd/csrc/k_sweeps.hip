// gfx950 kernels of the dense state-vector engine: per-operation sweeps,
// elementwise chains, reductions, sampling chains, comparison, observables.
//
// Storage is interleaved complex (double2 for complex128, float2 for
// complex64), basis index bit p holds qubit n-1-p.  Every kernel indexes with
// 64-bit integers.  The library is compiled with -ffp-contract=off so that the
// per-operation kernels reproduce the reference dense backend's arithmetic
// (src/tmfqs/storage/dense/dense_state_backend.cpp) bit for bit; kernels that
// are not bit-exact by design (phase ladders, twiddles) use explicit fma().
#include "device_common.h"

namespace tmfqs_hip {

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
// one such pair at a time and stages both tiles in LDS.  The low C bits give
// runs of 2^C contiguous amplitudes; the index bits of the other local bits
// and the permuted local index come from small LDS tables (S has at most C
// bits above the low C), and every thread issues all of its loads of a pair
// (up to U per tile) before it stores.
struct InvParams {
	u64 npairs_tiles;  // 2^(n - s)
	uint32_t n, C, s, nh;
	uint32_t perm[64];
	uint32_t hbits[16];  // S \ {0..C-1}, ascending
	uint32_t lperm[16];  // perm restricted to S in local-bit numbering
};

template <typename R, unsigned U>
__global__ void __launch_bounds__(256) k_involution(typename V2T<R>::t *__restrict__ a, InvParams P) {
	using V = typename V2T<R>::t;
	extern __shared__ char smem[];
	const unsigned nloc = 1u << P.s, nhi = 1u << P.nh, nlo = 1u << P.C, lowmask = nlo - 1u;
	V *ta = (V *)smem;
	V *tb = ta + nloc;
	u64 *hitab = (u64 *)(tb + nloc);            // high local bits -> their index bits
	unsigned *slo = (unsigned *)(hitab + nhi);  // low bits -> bits of the source's local index
	unsigned *shi = slo + nlo;                  // high local bits -> bits of the source's local index
	for(unsigned j = threadIdx.x; j < nhi; j += blockDim.x) {
		u64 h = 0;
		unsigned sh = 0;
		for(unsigned b = 0; b < P.nh; ++b)
			if((j >> b) & 1u) {
				h |= 1ull << P.hbits[b];
				sh |= 1u << P.lperm[P.C + b];
			}
		hitab[j] = h;
		shi[j] = sh;
	}
	for(unsigned j = threadIdx.x; j < nlo; j += blockDim.x) {
		unsigned sl = 0;
		for(unsigned b = 0; b < P.C; ++b)
			if((j >> b) & 1u) sl |= 1u << P.lperm[b];
		slo[j] = sl;
	}
	for(u64 tid = blockIdx.x; tid < P.npairs_tiles; tid += gridDim.x) {
		u64 base = tid << P.C;
		for(unsigned b = 0; b < P.nh; ++b) base = ins0(base, P.hbits[b]);
		u64 pb = 0;
		for(unsigned b = 0; b < P.n; ++b)
			if((base >> b) & 1ull) pb |= 1ull << P.perm[b];
		if(pb < base) continue;  // uniform: the partner tile's workgroup handles the pair
		const bool self = (pb == base);
		__syncthreads();  // the tables are built; the previous pair's LDS reads are done
		// complex128 values held as native vectors (with the HIP vector type the
		// compiler keeps these arrays in scratch)
		using NV = typename std::conditional<sizeof(R) == 8, R __attribute__((ext_vector_type(2))), V>::type;
		const NV *an = (const NV *)a;
		NV xa[U], xb[U];
#pragma unroll
		for(unsigned u = 0; u < U; ++u) {
			const unsigned l = threadIdx.x + u * blockDim.x;
			if(l < nloc) {
				const u64 off = (l & lowmask) | hitab[l >> P.C];
				xa[u] = an[base | off];
				if(!self) xb[u] = an[pb | off];
			}
		}
#pragma unroll
		for(unsigned u = 0; u < U; ++u) {
			const unsigned l = threadIdx.x + u * blockDim.x;
			if(l < nloc) {
				ta[l] = V{xa[u].x, xa[u].y};
				if(!self) tb[l] = V{xb[u].x, xb[u].y};
			}
		}
		__syncthreads();
#pragma unroll
		for(unsigned u = 0; u < U; ++u) {
			const unsigned l = threadIdx.x + u * blockDim.x;
			if(l < nloc) {
				const u64 off = (l & lowmask) | hitab[l >> P.C];
				const unsigned src = slo[l & lowmask] | shi[l >> P.C];
				if(self) {
					a[base | off] = ta[src];
				} else {
					a[pb | off] = ta[src];
					a[base | off] = tb[src];
				}
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
// exactly one half (a tie, flagged).  The chunk is read with coalesced loads
// (lane-consecutive states) into an LDS array of ticks; each thread then
// walks a run of consecutive ticks, and a workgroup scan of the runs gives
// the chunk's total and its prefix minimum and maximum (in ulps, relative to
// the chunk start) in state order.
constexpr unsigned kTickThreads = 256;
__device__ __forceinline__ unsigned tick_slot(unsigned i) { return i + (i >> 4); }  // skew: 2-way LDS banks at most

__device__ __forceinline__ long long shfl_up_ll(long long v, unsigned d) { return __shfl_up(v, d, 64); }

template <typename R>
__global__ void __launch_bounds__(256) k_chunk_ticks(const typename V2T<R>::t *__restrict__ a, u64 chunk, u64 nchunks,
                                                     const double *af, int sel, const int32_t *ulog, ChunkTicks *out) {
	__shared__ long long tk[(1u << kSampleChunkLog) + ((1u << kSampleChunkLog) >> 4)];
	__shared__ long long wsum[kTickThreads / 64], wmin[kTickThreads / 64], wmax[kTickThreads / 64];
	__shared__ int wflag[kTickThreads / 64];
	const unsigned t = threadIdx.x, lane = t & 63u, wave = t >> 6;
	const unsigned per = chunk >= kTickThreads ? (unsigned)(chunk / kTickThreads) : 1u;
	const bool active = chunk >= kTickThreads || t < chunk;
	for(u64 c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const int32_t ul = ulog[c];
		const bool grid = ul != INT32_MIN;
		int tie = 0, nz = 0;
		// 1. coalesced: lane-consecutive states, their ticks into LDS
		if(active)
			for(unsigned j = 0; j < per; ++j) {
				const unsigned i = j * kTickThreads + t;
				const auto v = a[c * chunk + i];
				const double x = chain_value(v.x, v.y, sel, af);
				nz |= (x != 0.0);
				if(grid) {
					const double q = ldexp(x, -ul);
					if(q - floor(q) == 0.5) tie = 1;
					tk[tick_slot(i)] = (long long)rint(q);
				}
			}
		__syncthreads();
		// 2. each thread's run of consecutive states, in state order
		long long cum = 0, lmin = LLONG_MAX, lmax = LLONG_MIN;
		if(active && grid)
			for(unsigned j = 0; j < per; ++j) {
				cum += tk[tick_slot(t * per + j)];
				lmin = cum < lmin ? cum : lmin;
				lmax = cum > lmax ? cum : lmax;
			}
		// 3. exclusive prefix of the runs over the workgroup, then the extremes
		long long inc = cum;
		for(unsigned d = 1; d < 64; d <<= 1) {
			const long long o = shfl_up_ll(inc, d);
			if(lane >= d) inc += o;
		}
		if(lane == 63) wsum[wave] = inc;
		__syncthreads();
		long long before = inc - cum;
		for(unsigned w = 0; w < wave; ++w) before += wsum[w];
		long long mn = lmin == LLONG_MAX ? LLONG_MAX : before + lmin;
		long long mx = lmax == LLONG_MIN ? LLONG_MIN : before + lmax;
		for(unsigned d = 32; d > 0; d >>= 1) {
			const long long omn = __shfl_down(mn, d, 64), omx = __shfl_down(mx, d, 64);
			mn = omn < mn ? omn : mn;
			mx = omx > mx ? omx : mx;
		}
		const int anyTie = __any(tie) ? 1 : 0, anyNz = __any(nz) ? 1 : 0;
		if(lane == 0) {
			wmin[wave] = mn;
			wmax[wave] = mx;
			wflag[wave] = anyTie | (anyNz << 1);
		}
		__syncthreads();
		if(t == 0) {
			long long total = 0, gmin = LLONG_MAX, gmax = LLONG_MIN;
			int flags = 0;
			for(unsigned w = 0; w < kTickThreads / 64; ++w) {
				total += wsum[w];
				gmin = wmin[w] < gmin ? wmin[w] : gmin;
				gmax = wmax[w] > gmax ? wmax[w] : gmax;
				flags |= wflag[w];
			}
			ChunkTicks o;
			o.ticks = total;
			o.pmin = gmin == LLONG_MAX ? 0 : gmin;
			o.pmax = gmax == LLONG_MIN ? 0 : gmax;
			o.tie = flags & 1;
			o.nonzero = flags >> 1;
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
	const unsigned umax = L.prec == 0 ? 4u : 16u;  // loads per thread and tile the kernel holds
	if(P.nh > C || (1ull << s) > 256ull * umax) return hipErrorInvalidValue;
	const size_t shmem = 2 * esz * (1ull << s) + (8 + 4) * (1ull << P.nh) + 4 * (1ull << C);
	const unsigned grid = (unsigned)(P.npairs_tiles < 16384 ? P.npairs_tiles : 16384);
	static bool attr = false;  // complex64 tiles of 2^12 amplitudes take 64 KiB plus the tables
	if(!attr) {
		(void)hipFuncSetAttribute((const void *)k_involution<double, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		(void)hipFuncSetAttribute((const void *)k_involution<float, 16>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
		attr = true;
	}
	if(L.prec == 0)
		TMFQS_LAUNCH((k_involution<double, 4>), grid, 256, shmem, L.stream, (double2 *)a, P);
	else
		TMFQS_LAUNCH((k_involution<float, 16>), grid, 256, shmem, L.stream, (float2 *)a, P);
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

}  // namespace tmfqs_hip
