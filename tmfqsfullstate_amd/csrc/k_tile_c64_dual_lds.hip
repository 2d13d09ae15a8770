// Tiled-pass kernels: complex64, dual (adjoint) form, with k-qubit matrix stages on the LDS tile.
#include "tile_device.h"

namespace tmfqs_hip {

TMFQS_TILE_ENTRY(launch_tile_c64_dual_lds) {
	return launch_staged_w<float, true, 1>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

}  // namespace tmfqs_hip
