// Tiled-pass kernels: complex64, with k-qubit matrix stages on the LDS tile.
#include "tile_device.h"

namespace tmfqs_hip {

TMFQS_TILE_ENTRY(launch_tile_c64_lds) {
	return launch_staged_w<float, false, 1>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

}  // namespace tmfqs_hip
