// randomPhase on the device: mt19937 jump-ahead across workgroups, the
// word stream, generate_canonical, and cos/sin in double-double with the
// near-midpoint flags of phase_math.h.
#include "device_common.h"
#include "engine.h"
#include "mt_jump.h"
#include "phase_math.h"

namespace tmfqs_hip {

using namespace tmfqs_phase;

namespace {

constexpr int kJumpThreads = 256;
constexpr int kGenThreads = 320;  // one thread per state of a 624-word block (312 states)

// states[dst0 + b] = the state J words after states[src0 + b], J given by
// its jump polynomial g (deg g < 19937): x_{J+m} = XOR_{j in g} x_{j+m}.
__global__ __launch_bounds__(kJumpThreads) void k_mt_jump(uint32_t *states, unsigned src0, unsigned dst0, const uint64_t *g) {
	__shared__ uint32_t seq[tmfqs_mt::kJumpSeqWords];
	const int tid = threadIdx.x;
	const uint32_t *src = states + (size_t)(src0 + blockIdx.x) * kMtN;
	for(int i = tid; i < kMtN; i += kJumpThreads) seq[i] = src[i];
	__syncthreads();
	// x_{k+624} from x_k, x_{k+1}, x_{k+397}: 227 words at a time never read
	// a word of the same round.
	for(int c = 0; c + kMtN < tmfqs_mt::kJumpSeqWords; c += kMtN - kMtM) {
		const int k = c + tid;
		if(tid < kMtN - kMtM && k + kMtN < tmfqs_mt::kJumpSeqWords) seq[k + kMtN] = mt_next(seq[k], seq[k + 1], seq[k + kMtM]);
		__syncthreads();
	}
	const int m0 = tid, m1 = tid + 256, m2 = min(tid + 512, kMtN - 1);
	uint32_t a0 = 0, a1 = 0, a2 = 0;
	for(int w = 0; w < kPolyWords; ++w) {
		for(uint64_t b = g[w]; b; b &= b - 1) {
			const int j = 64 * w + __builtin_ctzll(b);
			a0 ^= seq[j + m0];
			a1 ^= seq[j + m1];
			a2 ^= seq[j + m2];
		}
	}
	uint32_t *dst = states + (size_t)(dst0 + blockIdx.x) * kMtN;
	dst[m0] = a0;
	dst[m1] = a1;
	if(tid + 512 < kMtN) dst[tid + 512] = a2;
}

template <typename V>
__device__ __forceinline__ void store_amp(V *amp, u64 s, double re, double im);
template <>
__device__ __forceinline__ void store_amp<double2>(double2 *amp, u64 s, double re, double im) {
	amp[s] = make_double2(re, im);
}
template <>
__device__ __forceinline__ void store_amp<float2>(float2 *amp, u64 s, double re, double im) {
	amp[s] = make_float2((float)re, (float)im);
}

// Appends the lanes' flagged values (key = state << 1 | sin) with one atomic
// per wavefront.
__device__ __forceinline__ void append_flags(bool fc, bool fs, u64 state, double x, PhaseFlag *flags,
                                             unsigned long long *count, u64 cap) {
	const unsigned long long mc = __ballot(fc), ms = __ballot(fs);
	const int nc = __popcll(mc), ns = __popcll(ms);
	if(nc + ns == 0) return;
	const int lane = __lane_id();
	unsigned long long base = 0;
	if(lane == 0) base = atomicAdd(count, (unsigned long long)(nc + ns));
	base = __shfl(base, 0);
	const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	if(fc) {
		const u64 i = base + __popcll(mc & below);
		if(i < cap) flags[i] = PhaseFlag{state << 1, x};
	}
	if(fs) {
		const u64 i = base + nc + __popcll(ms & below);
		if(i < cap) flags[i] = PhaseFlag{(state << 1) | 1u, x};
	}
}

// One stream per workgroup: words [p L, (p + 1) L) of the seed's output,
// i.e. states [p L / 2, (p + 1) L / 2).
template <typename V>
__global__ __launch_bounds__(kGenThreads) void k_random_phase(const uint32_t *states, unsigned stream0, u64 words, double norm,
                                                              double range, V *amp, PhaseFlag *flags,
                                                              unsigned long long *count, u64 cap) {
	__shared__ uint32_t mt[kMtN];
	const int tid = threadIdx.x;
	const uint32_t *src = states + (size_t)blockIdx.x * kMtN;
	for(int i = tid; i < kMtN; i += kGenThreads) mt[i] = src[i];
	__syncthreads();
	const u64 first = (u64)(stream0 + blockIdx.x) * (words / 2);
	constexpr int kA = kMtN - kMtM;  // 227
	for(u64 done = 0; done < words; done += kMtN) {
		// The twist in three rounds, each reading only words that are final.
		uint32_t v = 0;
		if(tid < kA) v = mt_next(mt[tid], mt[tid + 1], mt[tid + kMtM]);
		__syncthreads();
		if(tid < kA) mt[tid] = v;
		__syncthreads();
		if(tid < kA) v = mt_next(mt[kA + tid], mt[kA + tid + 1], mt[tid]);
		__syncthreads();
		if(tid < kA) mt[kA + tid] = v;
		__syncthreads();
		constexpr int kC = kMtN - 2 * kA;  // 170
		if(tid < kC) v = mt_next(mt[2 * kA + tid], mt[(2 * kA + tid + 1) % kMtN], mt[kA + tid]);
		__syncthreads();
		if(tid < kC) mt[2 * kA + tid] = v;
		__syncthreads();
		const u64 left = (words - done) / 2;
		const bool active = tid < kMtN / 2 && (u64)tid < left;
		bool fc = false, fs = false;
		double x = 0.0;
		const u64 state = first + done / 2 + tid;
		if(active) {
			x = canonical53(mt_temper(mt[2 * tid]), mt_temper(mt[2 * tid + 1])) * range;
			DD S, C;
			dd_sincos(x, S, C);
			const double c = round_flagged(C, &fc), s = round_flagged(S, &fs);
			store_amp(amp, state, norm * c, norm * s);
		}
		append_flags(fc, fs, state, x, flags, count, cap);
		__syncthreads();  // the next twist rewrites mt
	}
}

// The host's glibc values for the flagged entries.
template <typename T>
__global__ void k_phase_patch(const PhaseFlag *flags, u64 count, double norm, T *amp) {
	for(u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (u64)gridDim.x * blockDim.x) {
		const PhaseFlag f = flags[i];
		amp[2 * (f.key >> 1) + (f.key & 1u)] = (T)(norm * f.x);
	}
}

}  // namespace

hipError_t launch_mt_jump(hipStream_t stream, uint32_t *dstates, unsigned src0, unsigned dst0, unsigned count,
                          const uint64_t *dpoly) {
	if(count == 0) return hipSuccess;
	TMFQS_LAUNCH(k_mt_jump, count, kJumpThreads, 0, stream, dstates, src0, dst0, dpoly);
	return hipGetLastError();
}

hipError_t launch_random_phase(const Launch &L, const uint32_t *dstates, unsigned stream0, unsigned streams, u64 words,
                               double norm, double range, void *amp, PhaseFlag *dflags, unsigned long long *dcount, u64 cap) {
	if(streams == 0) return hipSuccess;
	if(words % 2 != 0) return hipErrorInvalidValue;
	if(L.prec == 0)
		TMFQS_LAUNCH(k_random_phase<double2>, streams, kGenThreads, 0, L.stream, dstates, stream0, words, norm, range,
		             (double2 *)amp, dflags, dcount, cap);
	else
		TMFQS_LAUNCH(k_random_phase<float2>, streams, kGenThreads, 0, L.stream, dstates, stream0, words, norm, range,
		             (float2 *)amp, dflags, dcount, cap);
	return hipGetLastError();
}

hipError_t launch_phase_patch(const Launch &L, const PhaseFlag *dflags, u64 count, double norm, void *amp) {
	if(count == 0) return hipSuccess;
	const unsigned g = grid_for(count, 256, 4096);
	if(L.prec == 0)
		TMFQS_LAUNCH(k_phase_patch<double>, g, 256, 0, L.stream, dflags, count, norm, (double *)amp);
	else
		TMFQS_LAUNCH(k_phase_patch<float>, g, 256, 0, L.stream, dflags, count, norm, (float *)amp);
	return hipGetLastError();
}

}  // namespace tmfqs_hip
