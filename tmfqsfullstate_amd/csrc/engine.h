// Internal interfaces shared by the gfx950 kernels (k_*.hip) and the host
// engine (engine.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "tmfqs_hip.h"

#include <cstddef>
#include <cstdint>

namespace tmfqs_hip {

using u64 = uint64_t;

// ---------------------------------------------------------------- tile ops
// A tiled pass loads 2^T amplitudes whose indices vary over T "local" bit
// positions (the low C bits 0..C-1 plus up to kMaxHighBits higher positions)
// into LDS, applies a list of TileOps, and stores them back.
constexpr int kMaxHighBits = 10;
constexpr int kMaxTwiddleTerms = 16;
constexpr int kMaxTileOps = 512;

enum TileOpKind : uint32_t {
	TOP_H = 0,        // Hadamard on local bit j0 (zero pairs untouched)
	TOP_X = 1,        // Pauli-X / controlled-NOT on local bit j0 (equal pairs untouched)
	TOP_U1 = 2,       // 2x2 matrix on local bit j0 (reference k==1 arithmetic)
	TOP_UK = 3,       // 2^k x 2^k matrix on local bits jl[0..k) (reference k>=2 arithmetic)
	TOP_CP = 4,       // multiply by (cr, ci) where all bits of the condition are set
	TOP_LADDER = 5,   // multiply by e^{2 pi i F / 2^64}, F = (g & amask) << shift, where the condition holds
	TOP_TWIDDLE = 6,  // multiply by e^{2 pi i F / 2^64}, F = sum over terms whose bit is set of (g & mask) << shift
	TOP_GRAD = 7,     // dual passes only: accumulate Im sum conj(b) a where the condition holds into slot `mat`
	TOP_PHASES = 8,   // multiply by e^{i theta}, theta = sum of the terms [twid, twid + nterms) whose global
	                  // bits `cond` are all set, each term's angle the double stored in `mask`
};

// Most angles one TOP_PHASES op sums.
constexpr unsigned kMaxPhaseTerms = 64;

// Gradient slots one dual-pass launch can accumulate.
constexpr unsigned kMaxGradSlots = 64;
// Most workgroups of a dual (adjoint) tiled pass: its gradient partials
// buffer holds kMaxGradSlots * kMaxDualGrid doubles.
constexpr unsigned kMaxDualGrid = 8192;

struct TwiddleTerm {
	u64 cond;   // global bit that enables the term
	u64 mask;   // global bits feeding F
	uint32_t shift;
	uint32_t pad;
};

struct TileOp {
	uint32_t kind;
	uint32_t j0;          // local bit (H, X, U1)
	uint32_t k;           // UK arity
	uint32_t mat;         // offset (doubles) into the matrix pool
	uint32_t jl[5];       // UK local bits, jl[0] = MSB of the matrix index
	uint32_t nterms;      // TWIDDLE terms
	u64 lcond;            // condition on local index bits (all set)
	u64 gcond;            // condition on global index bits (all set)
	u64 amask;            // LADDER mask
	uint32_t shift;       // LADDER shift
	uint32_t twid;        // TWIDDLE: index of the first term in the term pool
	uint32_t cj0, cj1;    // LADDER/TWIDDLE: phase-table chunks [cj0, cj1] F can occupy
	uint32_t sj;          // register stages: position of j0 among the stage's bits
	uint32_t pad;
	double cr, ci;        // CP phase
};

// TOP_UK with k = 5: TileOp::twid is the offset (doubles) in the matrix pool
// of the matrix's MFMA operands, laid out lane by lane, or this value when the
// block runs on VALU in the reference's accumulation order.
constexpr uint32_t kNoMfmaOperands = 0xffffffffu;

// A run of tile ops applied between two LDS round trips.  In a register
// stage every thread holds the 2^W amplitudes whose local indices differ in
// the stage's W bit positions (ascending) and applies the stage's ops in
// registers; ops that need more bits than that (k-qubit matrices) run one by
// one on the LDS tile (lds = 1).
constexpr int kMaxStageBits = 5;
struct TileStage {
	uint32_t op0, nops;
	uint32_t lds;
	uint32_t bits[kMaxStageBits];
};

struct TilePass {
	u64 ntiles;           // 2^(n - T)
	uint32_t n;           // qubits
	uint32_t T;           // local bits
	uint32_t C;           // contiguous low local bits
	uint32_t nh;          // high local bits (T - C)
	uint32_t hbits[kMaxHighBits];
	uint32_t nops;
	uint32_t op0;         // first op in the op pool
};

// ------------------------------------------------------ elementwise chains
enum EwKind : uint32_t {
	EW_DIFFUSE = 0,  // a = 2 m - a
	EW_AFFINE = 1,   // a = s a + b (reference applyAffineTransform arithmetic)
};

struct EwStep {
	uint32_t kind;
	uint32_t pad;
	double ar, ai;   // mean (DIFFUSE) or scale (AFFINE)
	double br, bi;   // bias (AFFINE)
};

struct Scatter {
	u64 index;
	double re, im;
};

// ------------------------------------------------------------ launches
// All launch helpers enqueue on `stream` and return hipError_t of the launch.
// `prec` selects complex128 (0) or complex64 (1) storage.
struct Launch {
	hipStream_t stream;
	int prec;
};

hipError_t launch_hadamard(const Launch &L, void *a, unsigned n, unsigned bit);
hipError_t launch_pauli_x(const Launch &L, void *a, unsigned n, unsigned bit, u64 cmask);
hipError_t launch_cphase(const Launch &L, void *a, unsigned n, unsigned tbit, u64 cmask, double cr, double ci);
hipError_t launch_u1(const Launch &L, void *a, unsigned n, unsigned bit, const double *mat8);
hipError_t launch_uk(const Launch &L, void *a, unsigned n, unsigned k, const unsigned *bits, const double *dmat);
hipError_t launch_bitswap(const Launch &L, void *a, unsigned n, unsigned p, unsigned q);
hipError_t launch_involution(const Launch &L, void *a, unsigned n, const unsigned *perm, unsigned C);
hipError_t launch_ew_chain(const Launch &L, void *a, unsigned n, const EwStep *dsteps, unsigned nsteps);
hipError_t launch_scatter(const Launch &L, void *a, const Scatter *dsc, unsigned count);
// QFT block pass over the block bits pbot+7 .. pbot (see k_qft.hip): the
// top K of them are targets (K = 8: a full block; fewer: a partial leading
// pass), each with its whole ladder.  dcin holds the constants
// e^{i pi v / 2^j} at 2^j - 1 + v (FP64, then the same in FP32).
// C contiguous low bits per tile: 4 or 5 (complex128), 5 or 6 (complex64).
hipError_t launch_qft_block(const Launch &L, void *a, unsigned n, unsigned pbot, const double *dphase, const double *dcin,
                            unsigned C, unsigned K);
// Low contiguous bits a QFT block tile keeps: `requested` when supported,
// else 5 for both precisions.
unsigned qft_block_coalesce_bits(int prec, int requested);
// QFT targets 5..0 (H + complete ladders, no lower bits) followed by the full
// bit reversal, in place, for 12 <= n <= 40; dcq holds the 63 constants
// e^{i pi v / 2^t}, t = 0..5, v < 2^t (at 2^t - 1 + v).
hipError_t launch_qft_tail_reverse(const Launch &L, void *a, unsigned n, const double *dcq);
// Ten-target QFT passes (see k_qft.hip).  Block pass, in place, over the
// block bits pbot+9 .. pbot whose top K (1..10) bits are targets, each with
// its whole ladder; needs C <= pbot and pbot + 10 <= n (C = 3 complex128,
// 4 complex64).  Tail: targets 9..0 and the full bit reversal, from src into
// dst (distinct buffers of 2^n amplitudes), for qft_wide_min_qubits(prec) <=
// n <= 40.  Persistent workgroups that load their next tile into LDS during
// the current one; grid_cap sets the workgroup count (0: the measured
// default for the pass).
hipError_t launch_qft_wide_block(const Launch &L, void *a, unsigned n, unsigned pbot, unsigned K, const double *dphase,
                                 const double *dcin, u64 grid_cap = 0);
hipError_t launch_qft_wide_tail(const Launch &L, const void *src, void *dst, unsigned n, const double *dphase, const double *dcin,
                                u64 grid_cap = 0);
unsigned qft_wide_min_qubits(int prec);
// Chain of diffusions only: dm2 holds nsteps pairs (2 Re m, 2 Im m).  When
// zero_ok (every 2 Im m is +-0), zp and zm are the chain applied to an
// imaginary part +0 and -0; wavefronts of real amplitudes use them.
hipError_t launch_diffuse_chain(const Launch &L, void *a, unsigned n, const double *dm2, unsigned nsteps, double zp, double zm,
                                int zero_ok);
hipError_t launch_copy(const Launch &L, const void *src, void *dst, unsigned n);

struct CompareBlock {
	double d2, r2, c2, ovr, ovi, neq, maxd, maxcomp, maxdp;
	u64 worst;
};
hipError_t launch_compare(const Launch &L, int prec_ref, const void *ref, int prec_cand, const void *cand, unsigned n,
                          CompareBlock *dout, unsigned *blocks);
// Register-stage width for a tile of T bits.
unsigned tile_stage_width(int prec, unsigned T, bool dual);
// Workgroups of a tiled pass over P; a dual pass's gpart holds nslots * grid doubles.
unsigned tile_grid(const TilePass &P, bool dual, int prec);
// Tiled pass over one state (b == nullptr) or two states stepped together
// (b != nullptr: TOP_GRAD partials into gpart); ldsops: 0 register stages
// only, 1 some stage runs k-qubit matrices on the LDS tile (a separate kernel
// build, so the common case keeps its register budget), 2 as 1 with every
// five-qubit block on FP64 MFMA (complex128 tiles of >= 11 bits whose
// blocks carry MFMA operands; no VALU five-qubit code, fewer registers).
hipError_t launch_tile_staged(const Launch &L, void *a, void *b, const TilePass &P, unsigned W, const TileOp *dops,
                              const TileStage *dstages, unsigned nstages, const double *dmats, const TwiddleTerm *dterms,
                              const double *dphase, double *gpart, unsigned nslots, int ldsops);
// grads[gidx[s]] = scale * sum_i gpart[s * blocks + i], s < nslots.
hipError_t launch_grad_finish(const Launch &L, const double *gpart, unsigned blocks, unsigned nslots, const uint32_t *gidx,
                              double scale, double *grads);

// Diagonal observable: Z-string terms (basis-index bit masks) and basis-state projectors.
constexpr unsigned kMaxObsTerms = 16;
struct ObsParams {
	u64 zmask[kMaxObsTerms];
	double zcoef[kMaxObsTerms];
	u64 pstate[kMaxObsTerms];
	double pcoef[kMaxObsTerms];
	uint32_t nz, np;
};
// *dout = sum_x O(x) |a_x|^2 (FP64); when lam is non-null also lam = O a.
// dpart holds kReduceBlocks doubles.
hipError_t launch_observable(const Launch &L, const void *a, void *lam, unsigned n, const ObsParams &O, double *dpart,
                             double *dout);
// The observable of a C-ABI term list on n qubits (throws std::invalid_argument
// / std::out_of_range as tmfqs_hip_expectation reports them).
ObsParams observable_params(unsigned n, const tmfqs_hip_diag_term *terms, size_t nterms);
// Throws the library's exception for a failed HIP call (std::bad_alloc when
// out of memory).
void check_device(hipError_t e, const char *what);
// Inside a catch block: records the message for tmfqs_hip_last_error and
// returns the matching TMFQS_HIP_ERR_* status.
int status_of_current_exception();
// Device windows of chunked states (chunked.cpp) are ordinary handles whose
// buffer the chunked layer fills and drains with copies on the handle's
// stream.  window_reset issues any queued work and drops the handle's
// host-side knowledge of its amplitudes (values cached for phase flips), as
// before new contents arrive; window_flush issues every queued operation;
// both return the device buffer (2^n amplitudes) and its stream.
void *window_reset(tmfqs_hip_state *h, hipStream_t *stream);
void *window_flush(tmfqs_hip_state *h, hipStream_t *stream);
// Reductions write `blocks` partials into dpart, then reduce into dout
// (sum: 2 doubles; norm: 1 double).
constexpr unsigned kReduceBlocks = 1024;
hipError_t launch_sum(const Launch &L, const void *a, unsigned n, double *dpart, double *dout);
hipError_t launch_norm(const Launch &L, const void *a, unsigned n, const double *affine4, double *dpart,
                       double *dout);
// Sampling support: per-chunk FP64 probability sums (chunk = kSampleChunk states).
constexpr unsigned kSampleChunkLog = 12;
// Chain values: sel 0 = |scale*a + bias|^2 (affine4 may be null), 1 = Re a, 2 = Im a.
hipError_t launch_chunk_sum(const Launch &L, const void *a, unsigned n, const double *affine4, int sel, double *dsums);
// Per-chunk integer increments of the sequential chain on the grid
// u = 2^ulog[c] (int64) with their prefix extremes, a tie flag and a non-zero
// flag; ulog[c] == INT32_MIN marks a chunk that is resolved on the host.
struct ChunkTicks {
	long long ticks, pmin, pmax;
	int32_t tie, nonzero;
};
hipError_t launch_chunk_ticks(const Launch &L, const void *a, unsigned n, const double *affine4, int sel,
                              const int32_t *dulog, ChunkTicks *dout);

// ------------------------------------------------------------ randomPhase
// A value whose double-double cos/sin lay too close to a rounding midpoint:
// key = state << 1 | (0 cos, 1 sin); x is the phase, replaced on the host by
// glibc's value before the patch kernel stores norm * x.
struct PhaseFlag {
	u64 key;
	double x;
};
// Jump states[src0 + b] -> states[dst0 + b], b < count, by the polynomial
// dpoly (tmfqs_phase::kPolyWords words); states are 624-word buffers.
hipError_t launch_mt_jump(hipStream_t stream, uint32_t *dstates, unsigned src0, unsigned dst0, unsigned count,
                          const uint64_t *dpoly);
// Streams [stream0, stream0 + streams) of `words` words each (dstates points
// at stream0's state) write their words / 2 amplitudes norm * (cos, sin) of
// canonical * range and append near-midpoint values to dflags (*dcount
// counts every one; entries at or past cap are dropped).
hipError_t launch_random_phase(const Launch &L, const uint32_t *dstates, unsigned stream0, unsigned streams, u64 words,
                               double norm, double range, void *amp, PhaseFlag *dflags, unsigned long long *dcount, u64 cap);
hipError_t launch_phase_patch(const Launch &L, const PhaseFlag *dflags, u64 count, double norm, void *amp);

// ------------------------------------------------------ batched circuits
// Batched tile passes hold at most 2^batch_max_tile_bits(prec) amplitudes per
// tile (64 KiB of amplitudes: two workgroups per CU).
__host__ __device__ constexpr unsigned batch_max_tile_bits(int prec) { return prec == 0 ? 12u : 13u; }
struct BatchPass {
	void *state;      // the circuit's state buffer
	void *state2;     // dual passes: the second state (lambda)
	TilePass tp;      // geometry of the pass
	uint32_t op0;     // its ops in the batch op pool
	uint32_t st0;     // its stages in the batch stage pool
	uint32_t nst;
	uint32_t nslots;  // dual passes: gradient slots (TOP_GRAD ops)
};
// One gradient of a batched dual round: the sum over the pass's work items
// [item0, item0 + nitems) of their partials for `slot`, scaled, goes to
// grads[gindex].
struct BatchGrad {
	uint32_t item0, nitems, slot, pad;
	u64 gindex;
};
struct BatchItem {
	uint32_t pass;    // index into the round's pass table
	uint32_t ntiles;  // tiles [tile0, tile0 + ntiles) of that pass
	u64 tile0;
};
// A chunk of one circuit's state: initialisation (dst, init) or final
// permutation (src -> dst through map).
struct BatchChunk {
	const void *src;
	void *dst;
	u64 first, count;
	u64 init;
	uint32_t n, pad;
	uint8_t map[64];
};
hipError_t launch_batch_init(const Launch &L, const BatchChunk *dchunks, unsigned nchunks);
hipError_t launch_batch_permute(const Launch &L, const BatchChunk *dchunks, unsigned nchunks);
// One round: W is the register-stage width every pass of the round uses.
hipError_t launch_batch_tiles(const Launch &L, unsigned W, const BatchPass *dpasses, const BatchItem *ditems, unsigned nitems,
                              const TileOp *dops, const TileStage *dstages, const double *dphase, unsigned grid);
// One dual round (adjoint backward): every pass applies its ops to state and
// state2 and writes per-item partials of its TOP_GRAD slots to
// gpart[item * kMaxGradSlots + slot]; W is the round's register-stage width.
hipError_t launch_batch_dual(const Launch &L, unsigned W, const BatchPass *dpasses, const BatchItem *ditems, unsigned nitems,
                             const TileOp *dops, const TileStage *dstages, double *dgpart, unsigned grid);
hipError_t launch_batch_grad_finish(const Launch &L, const BatchGrad *dg, unsigned count, const double *dgpart, double scale,
                                    double *dgrads);

}  // namespace tmfqs_hip
