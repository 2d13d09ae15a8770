// Tiled-pass kernels: complex128, with k-qubit matrix stages on the LDS
// tile and five-qubit blocks on FP64 MFMA.
#include "tile_device.h"

namespace tmfqs_hip {

TMFQS_TILE_ENTRY(launch_tile_c128_mfma) {
	return launch_staged_w<double, false, 2>(L, W, a, b, P, dops, dstages, nstages, dmats, dterms, dphase, gpart, nslots);
}

}  // namespace tmfqs_hip
