// Tiled-pass kernels: complex64.
#include "tile_device.h"

namespace tmfqs_hip {

TMFQS_TILE_ENTRY(launch_tile_c64) {
	return launch_staged_w<float, false, 0>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

}  // namespace tmfqs_hip
