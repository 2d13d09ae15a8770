// Tiled-pass geometry and the launch dispatch over precision and form.
#include "tile_device.h"

namespace tmfqs_hip {

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

hipError_t launch_tile_staged(const Launch &L, void *a, void *b, const TilePass &P, unsigned W, const TileOp *dops,
                              const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                              const double *dphase, double *gpart, unsigned nslots, int ldsops) {
	if(W > P.T || nslots > kMaxGradSlots) return hipErrorInvalidValue;
	if(ldsops < 0 || ldsops > 2) return hipErrorInvalidValue;
	// ldsops: 0 register stages only, 1 with k-qubit matrix stages, 2 with
	// five-qubit blocks on FP64 MFMA (complex128; complex64 blocks always use
	// FP32 MFMA in build 1)
	static const decltype(&launch_tile_c128) entry[2][2][3] = {
	    {{launch_tile_c128, launch_tile_c128_lds, launch_tile_c128_mfma},
	     {launch_tile_c128_dual, launch_tile_c128_dual_lds, launch_tile_c128_dual_mfma}},
	    {{launch_tile_c64, launch_tile_c64_lds, launch_tile_c64_lds}, {launch_tile_c64_dual, launch_tile_c64_dual_lds, launch_tile_c64_dual_lds}}};
	return entry[L.prec == 0 ? 0 : 1][b ? 1 : 0][ldsops](L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

}  // namespace tmfqs_hip
