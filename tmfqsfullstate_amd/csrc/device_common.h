// Device helpers shared by the gfx950 kernel translation units: storage
// types, 64-bit index arithmetic, streaming loads and stores, wavefront and
// workgroup reductions, and phase arithmetic.  Internal to the library.
#pragma once

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

static inline unsigned grid_for(u64 work, unsigned threads = 256, u64 maxblocks = 8192) {
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

// s_waitcnt vmcnt(N): at most N of the wave's vector-memory accesses stay
// outstanding (gfx9 encoding: vmcnt bits 3:0 and 15:14; expcnt 6:4 and
// lgkmcnt 11:8 at their maxima, not waited for).
template <unsigned N> __device__ __forceinline__ void wait_vmcnt() {
	static_assert(N == 0 || N == 16 || N == 32, "wait_vmcnt");
	__builtin_amdgcn_s_waitcnt((int)((N & 15u) | ((N >> 4) << 14) | (7u << 4) | (15u << 8)));
}

typedef __attribute__((address_space(3))) void lds_void_t;

// A 16-byte LDS-direct load per lane (gfx950 global_load_lds_dwordx4,
// streaming): lane l writes its 16 bytes at the LDS base (M0) + 16 l, and no
// registers receive data.  Written as inline asm with no output registers:
// the compiler does not track it, so it adds no vmcnt(0) of its own before
// later LDS reads or barriers (for the builtin it must, as its wait model
// lets loads and stores on the vector-memory counter complete out of order,
// which would make every tile wait for the previous tile's stores).  The
// hardware counts it with every other vector-memory access of the wave in
// issue order; the caller waits for it with wait_vmcnt before its data is
// read, and its address registers are read when it issues.  M0 is saved and
// restored around it (the compiler reserves M0), with a wait state after each
// write of M0.
__device__ __forceinline__ void lds_direct_load16(const void *g, void *lds) {
	const unsigned base = (unsigned)(size_t)(lds_void_t *)lds;
	unsigned saved;
	asm volatile(
	    "s_mov_b32 %0, m0\n\t"
	    "s_mov_b32 m0, %2\n\t"
	    "s_nop 0\n\t"
	    "global_load_lds_dwordx4 %1, off nt\n\t"
	    "s_nop 0\n\t"
	    "s_mov_b32 m0, %0"
	    : "=&s"(saved)
	    : "v"(g), "s"(base)
	    : "memory");
}


// Streaming passes (copies, elementwise chains, per-operation sweeps) run on
// up to 2^18 workgroups: a 2^30-amplitude copy measured 5.69 TB/s at that
// grid against 5.10 TB/s at 2^14; per-operation Hadamard and X sweeps were
// 6-12% faster than at 2^13.
constexpr u64 kStreamGrid = 262144;

// ------------------------------------------------------------------------
// Reductions (FP64 accumulation, deterministic fixed-shape trees).
__device__ __forceinline__ double wave_sum(double v) {
	for(int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
	return v;
}

// The same sum through DPP lane moves (no LDS traffic): prefix sums within
// each row of 16 lanes (row_shr 1, 2, 4, 8), then the row totals carried up
// by row_bcast 15 and 31; lane 63 holds the total, which every lane gets.  A
// different fixed order than wave_sum (deterministic, not bit-identical).
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_f64(double v) {
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, true);
	const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, true);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
	v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
	v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
	v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
	v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
	v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
	v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
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

// Unscaled butterfly (x + y, x - y): transforms that apply several Hadamards
// in a row scale once at the end, by an exact power of two.
template <typename V> __device__ __forceinline__ void butterfly(V &x, V &y) {
	const V s = x;
	x = mk<V>(s.x + y.x, s.y + y.y);
	y = mk<V>(s.x - y.x, s.y - y.y);
}

__device__ __forceinline__ double2 csqr(double2 p) { return make_double2(fma(p.x, p.x, -p.y * p.y), 2.0 * p.x * p.y); }

// ------------------------------------------------------------------------
// Host launch helpers.
#define TMFQS_LAUNCH(KERNEL, GRID, BLOCK, SHMEM, STREAM, ...) \
	hipLaunchKernelGGL(KERNEL, dim3(GRID), dim3(BLOCK), SHMEM, STREAM, __VA_ARGS__)

}  // namespace tmfqs_hip
