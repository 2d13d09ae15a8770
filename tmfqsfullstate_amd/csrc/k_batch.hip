// gfx950 batched-circuit kernels: rounds of persistent work-list passes,
// initialisation and final relabeling gathers.
#include "tile_device.h"

namespace tmfqs_hip {

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
			staged_tile<R, W, false, 0>((V *)bp.state, nullptr, bp.tp, tid, ops + bp.op0, stages + bp.st0, bp.nst, nullptr,
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
			staged_tile<R, W, true, 0>((V *)bp.state, (V *)bp.state2, bp.tp, tid, ops + bp.op0, stages + bp.st0, bp.nst, nullptr,
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
