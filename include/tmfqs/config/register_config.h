// Register / backend configuration (reference: include/tmfqs/config/register_config.h:9-109).
// The device backend adds StorageStrategyKind::Hip (appended after Auto so the
// existing ordinals keep their values) and HipConfig.
#ifndef TMFQS_CONFIG_REGISTER_CONFIG_H
#define TMFQS_CONFIG_REGISTER_CONFIG_H

#include <cstddef>

namespace tmfqs {

enum class StorageStrategyKind {
	Dense,
	Blosc,
	Zfp,
	Auto,
	/** MI355X (gfx950) device state vector: the default strategy of this
	 *  library and what Auto resolves to (see RegisterConfig). */
	Hip,
	/** Chunked register in pinned host memory run through device windows
	 *  (HipConfig::chunkStates, HipConfig::windowChunks): the chunked
	 *  storage of the Blosc and Zfp strategies without their codecs. */
	HipChunked
};

enum class StorageWorkloadHint { Generic, Grover, Qft };

enum class ZfpCompressionMode { FixedRate, FixedPrecision, FixedAccuracy };

struct BloscConfig {
	size_t chunkStates = 32768;
	int clevel = 1;
	int nthreads = 4;
	int compcode = 1;
	bool useShuffle = true;
	size_t gateCacheSlots = 8;
};

struct BloscTuningOverrides {
	bool chunkStates = false;
	bool clevel = false;
	bool nthreads = false;
	bool gateCacheSlots = false;
};

struct ZfpConfig {
	ZfpCompressionMode mode = ZfpCompressionMode::FixedPrecision;
	double rate = 32.0;
	unsigned int precision = 40;
	double accuracy = 1e-8;
	size_t chunkStates = 32768;
	int nthreads = 4;
	size_t gateCacheSlots = 8;
};

struct ZfpTuningOverrides {
	bool mode = false;
	bool rate = false;
	bool precision = false;
	bool accuracy = false;
	bool chunkStates = false;
	bool nthreads = false;
	bool gateCacheSlots = false;
};

/** Amplitude storage precision of the device backend. */
enum class HipPrecision {
	/** FP64 amplitudes: matches the dense backend (bit-exact except fused QFT phase ladders). */
	Complex128,
	/** FP32 amplitudes; FP64 reductions, Grover means and phase factors. */
	Complex64
};

/** Device backend settings (0 means "choose automatically"). */
struct HipConfig {
	HipPrecision precision = HipPrecision::Complex128;
	int device = 0;
	/** Queue operations and fuse them into tiled passes. */
	bool fusion = true;
	/** Qubits per LDS tile (0: 12 for complex128, 13 for complex64). */
	unsigned int tileQubits = 0;
	/** Low qubits always inside a tile (0: automatic). */
	unsigned int coalesceQubits = 0;
	/** Most Grover/affine steps folded into one sweep (0: automatic). */
	unsigned int groverBlock = 0;
	/** The dense backend's arithmetic bit for bit everywhere: every controlled
	 *  phase as its own product and five-qubit blocks in its accumulation
	 *  order on VALU (otherwise fused ladders and MFMA blocks, within 1e-12). */
	bool exactPhases = false;
	/** Gate-run composition width: runs of consecutive gates on at most this
	 *  many qubits (2..5) become one dense block (0: off). */
	unsigned int fuseQubits = 0;
	/** Fewest gates a run needs to be composed (0: 24). */
	unsigned int fuseMinGates = 0;
	/** Complete QFTs: 0 runs ten-target passes and an out-of-place tail fused
	 *  with the bit reversal (three passes at 30 qubits) when device memory
	 *  has room for a second state buffer; 1 keeps the in-place eight-target
	 *  plan (four passes at 30 qubits, no second buffer). */
	unsigned int qftPlan = 0;
	/** HipChunked: amplitudes per host chunk, a power of two (0: 2^24, or the
	 *  whole register when it is smaller), the counterpart of
	 *  BloscConfig::chunkStates; and chunks per device window, a power of two
	 *  >= 2 (0: 16, or all chunks when there are fewer), the counterpart of
	 *  BloscConfig::gateCacheSlots. */
	size_t chunkStates = 0;
	size_t windowChunks = 0;
};

struct HipTuningOverrides {
	bool precision = false;
	bool tileQubits = false;
	bool coalesceQubits = false;
	bool groverBlock = false;
};

struct RegisterConfig {
	/** Differs from the reference, whose default is Dense.  This library ships
	 *  only the device backend, so registers and simulators built without a
	 *  strategy run on the GPU, Auto resolves to Hip whatever the state size,
	 *  and Dense, Blosc and Zfp report unavailable (constructing them throws
	 *  std::runtime_error, as an unavailable codec backend does in the
	 *  reference).  Inside the reference itself, where Hip is added to its
	 *  own registry, the default stays Dense and Hip is opt-in. */
	StorageStrategyKind strategy = StorageStrategyKind::Hip;
	size_t autoThresholdBytes = 8u * 1024u * 1024u;
	StorageWorkloadHint workloadHint = StorageWorkloadHint::Generic;
	BloscConfig blosc;
	BloscTuningOverrides bloscOverrides;
	ZfpConfig zfp;
	ZfpTuningOverrides zfpOverrides;
	HipConfig hip;
	HipTuningOverrides hipOverrides;
};

}  // namespace tmfqs

#endif
