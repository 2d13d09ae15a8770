/*
 * tmfqs_hip.h -- C-ABI of the MI355X (gfx950) dense state-vector engine.
 *
 * One opaque handle owns one state vector in HBM (complex128 or complex64,
 * interleaved re/im, basis index bit n-1-q holds qubit q), one HIP stream,
 * pinned staging buffers and device scalars.  The entry points below are the
 * backend operations of the reference's storage plugin interface
 * (include/tmfqs/storage/i_state_backend.h in the reference) plus the few
 * register-level operations the device backend accelerates; each declaration
 * cites the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns a tmfqs_hip_status; 0 is success.  Negative codes
 *     map 1:1 onto the exceptions the reference throws (std::out_of_range,
 *     std::invalid_argument, std::logic_error, std::runtime_error,
 *     std::bad_alloc).  tmfqs_hip_last_error() returns the message of the
 *     calling thread's last failure.
 *   - Qubit arguments use the reference's qubit numbering (qubit 0 is the most
 *     significant bit of the basis index).  State indices are 64-bit.
 *   - Mutating calls are queued on the handle's stream and may be fused with
 *     neighbouring calls; calls that return values (amplitude, probability,
 *     sample, export, sums) flush the queue and synchronise first.
 *   - A handle must not be used by two threads at once; distinct handles may.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - No exception crosses this boundary.
 */
#ifndef TMFQS_HIP_H
#define TMFQS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMFQS_HIP_ABI_VERSION 3

typedef struct tmfqs_hip_state tmfqs_hip_state;

typedef enum tmfqs_hip_status {
	TMFQS_HIP_OK = 0,
	TMFQS_HIP_ERR_OUT_OF_RANGE = -1,      /* std::out_of_range */
	TMFQS_HIP_ERR_INVALID_ARGUMENT = -2,  /* std::invalid_argument */
	TMFQS_HIP_ERR_LOGIC = -3,             /* std::logic_error ("not initialized") */
	TMFQS_HIP_ERR_RUNTIME = -4,           /* std::runtime_error */
	TMFQS_HIP_ERR_OUT_OF_MEMORY = -5,     /* std::bad_alloc */
	TMFQS_HIP_ERR_DEVICE = -6             /* HIP runtime failure (std::runtime_error) */
} tmfqs_hip_status;

typedef enum tmfqs_hip_precision {
	TMFQS_HIP_COMPLEX128 = 0, /* FP64 amplitudes; matches the reference dense backend */
	TMFQS_HIP_COMPLEX64 = 1   /* FP32 amplitudes; FP64 reductions, means and phase factors */
} tmfqs_hip_precision;

/* Backend configuration: the device analogue of the reference's per-backend
 * config structs (include/tmfqs/config/register_config.h:41-91). */
typedef struct tmfqs_hip_config {
	int precision;       /* tmfqs_hip_precision */
	int device;          /* HIP device ordinal */
	int fusion;          /* 1: queue and fuse operations into tiled passes; 0: one sweep per call */
	int tile_qubits;     /* tiled-pass tile size in qubits (0: automatic) */
	int coalesce_qubits; /* low qubits always inside a tile (0: automatic) */
	int grover_block;    /* most elementwise steps folded into one sweep (0: 1024 for complex128, which replays
	                        every step's rounding; complex64 composes up to 65536 steps into one affine map) */
	int exact_phases;    /* 1: the reference's arithmetic bit for bit everywhere: every controlled
	                        phase as its own complex product, five-qubit blocks accumulated in
	                        the dense backend's order on VALU; 0: fuse QFT phase ladders and run
	                        five-qubit blocks on MFMA (both within 1e-12 of the reference) */
	int fuse_qubits;     /* gate-run composition: a run of consecutive gates acting on at most this many
	                        qubits (2..5) is multiplied into one dense matrix on the host in FP64 and
	                        applied as one k-qubit block (five-qubit blocks on MFMA); 0: off.  Not in
	                        the exact mode.  tmfqs_hip_tune chooses it per circuit. */
	int fuse_min_gates;  /* fewest gates a run needs to be composed (0: 24) */
	int qft_plan;        /* complete QFTs (makeQft's op order, fused ladders): 0 automatic: passes of
	                        ten targets and the last ten targets fused with the bit reversal out of
	                        place (three full-state passes at 30 qubits; the handle keeps a second
	                        state buffer, used when device memory has room for it, else 1);
	                        1: in place, passes of eight targets and the last six targets exchanged
	                        with the reversal in place (four passes at 30 qubits, no second buffer) */
} tmfqs_hip_config;

/* One circuit operation (the reference's circuit::Operation variant,
 * include/tmfqs/circuit/circuit.h:16-62, as a flat record). */
typedef enum tmfqs_hip_op_kind {
	TMFQS_HIP_OP_H = 0,
	TMFQS_HIP_OP_X = 1,
	TMFQS_HIP_OP_CX = 2,
	TMFQS_HIP_OP_CPHASE = 3,
	TMFQS_HIP_OP_SWAP = 4,
	TMFQS_HIP_OP_PHASE_FLIP = 5,
	TMFQS_HIP_OP_DIFFUSE_LAZY = 6,
	TMFQS_HIP_OP_DIFFUSE_MATERIALIZED = 7,
	TMFQS_HIP_OP_MATRIX = 8,
	TMFQS_HIP_OP_PHASE = 9  /* diag(1, e^{i theta}) on qubit q0; no circuit::Operation counterpart
	                           (the reference applies it as QuantumGate::Phase through applyGate) */
} tmfqs_hip_op_kind;

typedef struct tmfqs_hip_op {
	int32_t kind;     /* tmfqs_hip_op_kind */
	uint32_t q0;      /* target (H, X); control (CX, CPHASE); first (SWAP); MATRIX: offset into qubit pool */
	uint32_t q1;      /* target (CX, CPHASE); second (SWAP) */
	uint32_t k;       /* MATRIX: number of target qubits */
	double theta;     /* CPHASE angle (radians) */
	uint64_t state;   /* PHASE_FLIP basis state; MATRIX: offset (in doubles) into matrix pool */
} tmfqs_hip_op;

/* ---- library ---------------------------------------------------------- */
const char *tmfqs_hip_last_error(void);
int tmfqs_hip_abi_version(void);
int tmfqs_hip_device_count(int *count);
void tmfqs_hip_config_default(tmfqs_hip_config *cfg);

/* ---- lifecycle -------------------------------------------------------- */
/* createDenseStateBackend / createBloscStateBackend factory pattern
 * (src/tmfqs/storage/blosc/blosc_state_backend.cpp:14-21).  The handle holds
 * no state until one of the init/load calls below. */
int tmfqs_hip_create(const tmfqs_hip_config *cfg, tmfqs_hip_state **out);
int tmfqs_hip_destroy(tmfqs_hip_state *h);
/* IStateBackend::clone (i_state_backend.h:31). */
int tmfqs_hip_clone(tmfqs_hip_state *h, tmfqs_hip_state **out);
int tmfqs_hip_qubit_count(tmfqs_hip_state *h, unsigned *out);
int tmfqs_hip_get_config(tmfqs_hip_state *h, tmfqs_hip_config *out);

/* ---- initialisation and I/O ------------------------------------------- */
/* IStateBackend::initZero (i_state_backend.h:37). */
int tmfqs_hip_init_zero(tmfqs_hip_state *h, unsigned num_qubits);
/* IStateBackend::initBasis (i_state_backend.h:44). */
int tmfqs_hip_init_basis(tmfqs_hip_state *h, unsigned num_qubits, uint64_t state, double re, double im);
/* IStateBackend::initUniformSuperposition (i_state_backend.h:50). */
int tmfqs_hip_init_uniform(tmfqs_hip_state *h, unsigned num_qubits, const uint64_t *states, size_t count);
/* IStateBackend::loadAmplitudes (i_state_backend.h:56): elem_count = 2 * 2^n doubles. */
int tmfqs_hip_load(tmfqs_hip_state *h, unsigned num_qubits, const double *interleaved, uint64_t elem_count);
/* QuantumState::randomPhase(n, seed) materialized on this backend
 * (src/tmfqs/state/quantum_state.cpp:264-277 and materialize, :290-307),
 * generated in device memory: std::mt19937(seed) split into jump-ahead
 * streams, uniform_real_distribution(0, 2 acos(-1)) per state, and
 * norm * (cos, sin) of the phase, bit for bit the values glibc's sincos gives
 * (near-midpoint values are recomputed on the host with glibc).  Validates
 * the probability mass as fromAmplitudes does (TMFQS_HIP_ERR_INVALID beyond
 * 1e-9).  Replaces the host construction plus loadAmplitudes. */
int tmfqs_hip_init_random_phase(tmfqs_hip_state *h, unsigned num_qubits, unsigned seed);
/* IStateBackend::exportAmplitudes / forEachAmplitudeChunk (i_state_backend.h:61,66):
 * copies states [first, first + count) as interleaved FP64 into out. */
int tmfqs_hip_export(tmfqs_hip_state *h, uint64_t first, uint64_t count, double *out);
/* IStateBackend::amplitude (i_state_backend.h:73). */
int tmfqs_hip_amplitude(tmfqs_hip_state *h, uint64_t state, double *re, double *im);
/* IStateBackend::setAmplitude (i_state_backend.h:79). */
int tmfqs_hip_set_amplitude(tmfqs_hip_state *h, uint64_t state, double re, double im);
/* IStateBackend::probability (i_state_backend.h:85). */
int tmfqs_hip_probability(tmfqs_hip_state *h, uint64_t state, double *out);
/* IStateBackend::totalProbability (i_state_backend.h:90). */
int tmfqs_hip_total_probability(tmfqs_hip_state *h, double *out);
/* IStateBackend::sampleMeasurement (i_state_backend.h:96): first state whose
 * sequential FP64 cumulative probability reaches rnd (bit-exact index). */
int tmfqs_hip_sample(tmfqs_hip_state *h, double rnd, uint64_t *out);

/* ---- register-level reads the device accelerates ----------------------- */
/* QuantumRegister::logicalAmplitudeSum without an overlay
 * (src/tmfqs/register/quantum_register.cpp:176-189). */
int tmfqs_hip_amplitude_sum(tmfqs_hip_state *h, double *re, double *im);
/* QuantumRegister::totalProbability with the lazy affine overlay active
 * (quantum_register.cpp:269-281): sum |scale*a + bias|^2. */
int tmfqs_hip_total_probability_affine(tmfqs_hip_state *h, double s_re, double s_im, double b_re, double b_im,
                                       double *out);
/* QuantumRegister::measure with the lazy affine overlay active
 * (quantum_register.cpp:284-310): no range check on rnd, as there. */
int tmfqs_hip_sample_affine(tmfqs_hip_state *h, double rnd, double s_re, double s_im, double b_re, double b_im,
                            uint64_t *out);

/* ---- operations ------------------------------------------------------- */
/* IStateBackend::phaseFlipBasisState (i_state_backend.h:110). */
int tmfqs_hip_phase_flip(tmfqs_hip_state *h, uint64_t state);
/* IStateBackend::inversionAboutMean() (i_state_backend.h:112). */
int tmfqs_hip_inversion_about_mean(tmfqs_hip_state *h);
/* IStateBackend::inversionAboutMean(mean) (i_state_backend.h:117). */
int tmfqs_hip_inversion_about_mean_value(tmfqs_hip_state *h, double mean_re, double mean_im);
/* IStateBackend::applyAffineTransform (i_state_backend.h:123). */
int tmfqs_hip_affine(tmfqs_hip_state *h, double s_re, double s_im, double b_re, double b_im);
/* IStateBackend::applyHadamard (i_state_backend.h:128). */
int tmfqs_hip_hadamard(tmfqs_hip_state *h, unsigned qubit);
/* IStateBackend::applyPauliX (i_state_backend.h:133). */
int tmfqs_hip_pauli_x(tmfqs_hip_state *h, unsigned qubit);
/* IStateBackend::applyControlledPhaseShift (i_state_backend.h:140). */
int tmfqs_hip_controlled_phase(tmfqs_hip_state *h, unsigned control, unsigned target, double theta);
/* IStateBackend::applyControlledNot (i_state_backend.h:146). */
int tmfqs_hip_controlled_not(tmfqs_hip_state *h, unsigned control, unsigned target);
/* QuantumRegister::applySwap (quantum_register.cpp:518-527): bit-identical to
 * its three controlled-NOTs, executed as one bit permutation. */
int tmfqs_hip_swap(tmfqs_hip_state *h, unsigned q1, unsigned q2);
/* IStateBackend::applyGate (i_state_backend.h:153): mat is a row-major
 * interleaved 2^k x 2^k complex matrix, qubits[0] the MSB of its index. */
int tmfqs_hip_apply_gate(tmfqs_hip_state *h, const double *mat, unsigned k, const unsigned *qubits);
/* IStateBackend::beginOperationBatch / endOperationBatch (i_state_backend.h:159-163):
 * end flushes every queued operation. */
int tmfqs_hip_begin_batch(tmfqs_hip_state *h);
int tmfqs_hip_end_batch(tmfqs_hip_state *h);
/* Flushes the queue and waits for the device. */
int tmfqs_hip_synchronize(tmfqs_hip_state *h);

/* ---- comparison ------------------------------------------------------- */
/* experiment::StateComparison (src/tmfqs/experiment/comparison.cpp:12-63 in
 * the reference) computed on the device for two states of equal qubit count
 * on one device (either precision), plus the overlap <ref|cand> and the
 * fidelity |<ref|cand>|^2 / (<ref|ref> <cand|cand>). */
typedef struct tmfqs_hip_comparison {
	int bitwise_equal;
	double max_abs_amplitude_error;
	double max_abs_component_error;
	double rmse_amplitude;
	double rel_l2;
	double max_abs_probability_error;
	double total_probability_diff;
	uint64_t worst_state;
	double overlap_re, overlap_im;
	double fidelity;
} tmfqs_hip_comparison;
int tmfqs_hip_compare(tmfqs_hip_state *ref, tmfqs_hip_state *cand, tmfqs_hip_comparison *out);

/* ---- observability ---------------------------------------------------- */
/* Per-kernel device timing with HIP events on the handle's stream (the device
 * analogue of ObservabilityConfig.traceOperations, simulator.h:16-19). */
int tmfqs_hip_profile_enable(tmfqs_hip_state *h, int on);
int tmfqs_hip_profile_reset(tmfqs_hip_state *h);
/* Kernel-name table: for i < *count writes names[i] (NUL-terminated, up to
 * 63 chars), launches[i], milliseconds[i] and algorithmic bytes[i].  On entry
 * *count is the capacity; on return the number of entries. */
int tmfqs_hip_profile_read(tmfqs_hip_state *h, char (*names)[64], uint64_t *launches, double *milliseconds,
                           double *bytes, size_t *count);
/* Streaming copy of the state into a scratch buffer of the same size (reads
 * and writes 2^n amplitudes once each), repeated reps times; *ms is the mean
 * time of one copy.  A plain copy kernel of known size: the practical
 * bandwidth ceiling and the calibration for hardware byte counters.  No
 * counterpart in the reference. */
int tmfqs_hip_copy_probe(tmfqs_hip_state *h, int reps, double *ms);
/* Raw stream handle (hipStream_t) for callers that time with their own events. */
int tmfqs_hip_stream(tmfqs_hip_state *h, void **stream);

/* ---- operation lists -------------------------------------------------- */
/* Enqueues a list of operation records, as sim::Simulator::executeOperation
 * (src/tmfqs/sim/simulator.cpp:17-51) dispatches them to the backend.  MATRIX
 * records take their qubits from qubit_pool[q0 ..] and their row-major
 * interleaved matrix from mat_pool[state ..].  PHASE_FLIP calls
 * phaseFlipBasisState; both DIFFUSE kinds call the backend's
 * inversionAboutMean() (register-level laziness lives above this layer). */
int tmfqs_hip_run(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *qubit_pool);

/* ---- diagonal observables and adjoint differentiation ----------------- */
/* No counterpart in the reference (its experiments only compare states).
 * A diagonal observable is a sum of terms: Z-strings coef * prod_{q in bits}
 * Z_q (bits = mask of qubits in the reference numbering; Z_q is +1 when qubit
 * q is 0) and basis projectors coef * |bits><bits|.  At most 16 terms of
 * each kind. */
typedef enum tmfqs_hip_diag_kind {
	TMFQS_HIP_DIAG_Z = 0,
	TMFQS_HIP_DIAG_PROJECTOR = 1
} tmfqs_hip_diag_kind;

typedef struct tmfqs_hip_diag_term {
	int32_t kind;       /* tmfqs_hip_diag_kind */
	uint32_t reserved;
	double coef;
	uint64_t bits;      /* Z: qubit mask; PROJECTOR: basis state */
} tmfqs_hip_diag_term;

/* No counterpart in the reference (its tuning hook,
 * StorageStrategyRegistry::tuneConfig in
 * src/tmfqs/storage/factory/storage_strategy_registry.cpp:77-85, returns its
 * input unchanged).  Autotuner: times the gate list (H, X, CX, CPHASE, PHASE,
 * SWAP, MATRIX records) on a scratch copy of the handle's state for every
 * candidate of the tile width (automatic, default - 1, default) and the
 * gate-run composition width fuse_qubits (0, 3, 4, 5), two runs each, and
 * keeps the fastest in the handle's configuration (fuse_qubits stays as it
 * was in the exact mode).  The state is unchanged; one extra state buffer is
 * allocated.  best_ms (may be null) receives the winner's device time. */
int tmfqs_hip_tune(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *qubit_pool,
                   int *tile_qubits, int *fuse_qubits, double *best_ms);
/* Sets the two tuned parameters directly (a cached tuning result). */
int tmfqs_hip_set_tuning(tmfqs_hip_state *h, int tile_qubits, int fuse_qubits);

/* Per-operation device timing (the traced RunReport of sim::Simulator, whose
 * reference counterpart times host calls, src/tmfqs/sim/simulator.cpp:71-82):
 * tmfqs_hip_flush_timed executes everything queued, bracketed by a pair of
 * stream events, without waiting; tmfqs_hip_timed_read waits for the stream
 * and returns the device milliseconds of every bracket since the last read,
 * in order (the first cap of them; count receives how many there were). */
int tmfqs_hip_flush_timed(tmfqs_hip_state *h);
int tmfqs_hip_timed_read(tmfqs_hip_state *h, double *ms, size_t cap, size_t *count);

/* <psi|O|psi> of the current state (FP64 accumulation in both precisions). */
int tmfqs_hip_expectation(tmfqs_hip_state *h, const tmfqs_hip_diag_term *terms, size_t nterms, double *out);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * tmfqs_hip_expectation over a state in the caller's device memory (2^n
 * interleaved complex128 or complex64 amplitudes at amps), enqueued on the
 * caller's stream (a hipStream_t; null = the default stream) of device
 * `device`; *out (device memory, one double) receives <psi|O|psi>, the same
 * bits as tmfqs_hip_expectation.  Returns without waiting; allocates
 * nothing: its workgroup partials live in a fixed 8 KiB device array per GPU
 * (any n <= 63 fits).  Calls that overlap on one GPU share that array and
 * race; order them on one stream. */
int tmfqs_hip_expectation_device(int device, void *stream, int precision, const void *amps, unsigned n,
                                 const tmfqs_hip_diag_term *terms, size_t nterms, double *out);

/* Adjoint gradient: runs the gate list (H, X, CX, CPHASE, PHASE, SWAP,
 * MATRIX) forward from the handle's current state, leaving the final state in
 * the handle, and returns E = <psi|O|psi> and dE/dtheta for every CPHASE and
 * PHASE record, in record order (grad_count must equal their number).  The
 * forward sweep stores the state at a pass boundary roughly every
 * checkpoint_gates gates (0: ceil(sqrt(count))); the backward sweep restarts
 * psi from each checkpoint and steps psi and lambda = O psi back through the
 * inverse gates on two device buffers, reducing -2 Im <lambda|P|psi> per
 * parameter per wavefront.  Extra device memory: one state per checkpoint plus
 * one (fewer checkpoints are taken when HBM is short). */
int tmfqs_hip_adjoint(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *qubit_pool,
                      const tmfqs_hip_diag_term *terms, size_t nterms, unsigned checkpoint_gates, double *expval, double *grads,
                      size_t grad_count);

/* ---- batched circuits -------------------------------------------------- */
/* No counterpart in the reference (its campaigns run circuits one
 * Simulator::run at a time).  A batch owns the states of many independent
 * circuits of uneven qubit counts (4..34), each starting from a basis state.
 * Every circuit is planned like a register's gate run; round r of a batch run
 * applies pass r of every circuit in one persistent work-list kernel, and the
 * initialisation, the rounds and the final bit permutations are captured in
 * one HIP graph, so tmfqs_hip_batch_run is a single graph launch.  Gate kinds:
 * H, X, CX, CPHASE, PHASE, SWAP.  complex128 results are bit-identical to
 * running each circuit through the reference's dense backend. */
typedef struct tmfqs_hip_batch tmfqs_hip_batch;

typedef struct tmfqs_hip_circuit {
	unsigned qubits;
	uint32_t reserved;
	uint64_t init_state;     /* basis state the circuit starts from */
	const tmfqs_hip_op *ops; /* borrowed for the duration of tmfqs_hip_batch_create */
	size_t count;
} tmfqs_hip_circuit;

typedef struct tmfqs_hip_batch_stats {
	size_t circuits, rounds, passes, items;
	int tile_qubits; /* tile width of the current plan (config tile_qubits, or tuned) */
	int graph;       /* 1 when runs are one graph launch */
	double state_bytes;
} tmfqs_hip_batch_stats;

int tmfqs_hip_batch_create(const tmfqs_hip_config *cfg, const tmfqs_hip_circuit *circuits, size_t count, tmfqs_hip_batch **out);
/* Autotuner: times the batch at each candidate tile width (fusion width of
 * the tiled passes) and keeps the fastest; a config tile_qubits > 0 pins it. */
int tmfqs_hip_batch_tune(tmfqs_hip_batch *b, int *tile_qubits);
/* Initialises every state and runs every circuit; *seconds is the device time. */
int tmfqs_hip_batch_run(tmfqs_hip_batch *b, double *seconds);
/* Interleaved FP64 amplitudes (2 * 2^qubits doubles) of circuit `index`. */
int tmfqs_hip_batch_export(tmfqs_hip_batch *b, size_t index, double *out);
int tmfqs_hip_batch_info(tmfqs_hip_batch *b, tmfqs_hip_batch_stats *out);
/* Adjoint gradients of one diagonal observable (see tmfqs_hip_adjoint) for
 * every circuit of the batch, each from its basis state: expvals[c] per
 * circuit; grads holds the CPHASE/PHASE gradients of circuit 0 in record
 * order, then those of circuit 1, and so on (grad_count in total).  One batch
 * run gives the final states, then round r steps pass r of every circuit's
 * backward sweep (psi and lambda = O psi in the circuit's two buffers,
 * without checkpoints) in one work-list kernel, with per-item gradient
 * partials summed in a fixed order.  Every final state is in place afterwards,
 * as after tmfqs_hip_batch_run: the sweep steps copies of them (one more
 * buffer per circuit while it runs) when the device has room, else the
 * states themselves followed by a second batch run. */
int tmfqs_hip_batch_adjoint(tmfqs_hip_batch *b, const tmfqs_hip_diag_term *terms, size_t nterms, double *expvals, double *grads,
                            size_t grad_count);
int tmfqs_hip_batch_destroy(tmfqs_hip_batch *b);

/* ---- chunked host-resident states --------------------------------------
 * A register of 2^n amplitudes kept as 2^(n-c) chunks of 2^c amplitudes in
 * pinned host memory, uncompressed, for states beyond one GPU's HBM: the
 * reference's chunked storage (CompressedStateBackend,
 * include/tmfqs/storage/common/compressed_state_backend.h:103-747, with its
 * ChunkLayout, chunk_layout.h:22-90, and decompressed-chunk cache,
 * chunk_cache.h:83-122) without the codecs.  Operation lists stream through
 * device windows of 2^g chunks, each an engine handle of c + g qubits: the
 * list is cut into segments whose gates mix at most g chunk bits (the index
 * bits above the low c), and a segment visits every window holding those
 * bits, two windows in flight on two streams.  A gate inside a chunk is the
 * reference's IntraChunk pair mode, one across chunks its InterChunk mode
 * (pair_kernel_executor.h:26-103).  Controls and phases on chunk bits
 * outside a window are constant over it and become a condition or a phase
 * per window.  Both diffusion kinds reduce the amplitude sum over every
 * window, then update each; a run of phase flips alone (Grover's oracle
 * between two diffusions) negates the amplitudes where they are stored.  Kinds: H, X, CX, CPHASE, PHASE, SWAP, MATRIX,
 * PHASE_FLIP, DIFFUSE_*.  Device memory: two windows of 2^(c+g) amplitudes
 * (and their handles' working buffers).  Results equal the device register's
 * up to the rounding of phases and sums that a window regroups. */
typedef struct tmfqs_hip_chunked tmfqs_hip_chunked;

typedef struct tmfqs_hip_chunked_stats {
	uint64_t segments;  /* window sweeps run (one per segment, two per diffusion) */
	uint64_t windows;   /* window visits */
	double bytes_h2d, bytes_d2h;
} tmfqs_hip_chunked_stats;

/* 1 <= chunk_qubits <= num_qubits <= 40; window_bits >= 1 unless
 * chunk_qubits == num_qubits; chunk_qubits + window_bits <= num_qubits.
 * The state starts as |0...0>. */
int tmfqs_hip_chunked_create(const tmfqs_hip_config *cfg, unsigned num_qubits, unsigned chunk_qubits, unsigned window_bits,
                             tmfqs_hip_chunked **out);
int tmfqs_hip_chunked_init_basis(tmfqs_hip_chunked *c, uint64_t state, double re, double im);
/* elem_count = 2 * 2^n interleaved doubles. */
int tmfqs_hip_chunked_load(tmfqs_hip_chunked *c, const double *interleaved, uint64_t elem_count);
/* States [first, first + count) as interleaved FP64. */
int tmfqs_hip_chunked_export(tmfqs_hip_chunked *c, uint64_t first, uint64_t count, double *out);
/* As tmfqs_hip_run; a record that mixes more than g chunk bits is refused
 * (TMFQS_HIP_ERR_INVALID_ARGUMENT) before anything runs. */
int tmfqs_hip_chunked_run(tmfqs_hip_chunked *c, const tmfqs_hip_op *ops, size_t count, const double *mat_pool,
                          const unsigned *qubit_pool);
/* Sum of |a|^2, window by window. */
int tmfqs_hip_chunked_total_probability(tmfqs_hip_chunked *c, double *out);
/* The register-level operations of IStateBackend on a chunked register
 * (i_state_backend.h:37-120; the reference's CompressedStateBackend
 * implements them over its chunk cache, compressed_state_backend.h:330-747):
 * a deep copy; setAmplitude; applyAffineTransform (a -> s a + b);
 * inversionAboutMean(mean) (a -> 2 mean - a); the amplitude sum (window
 * sums added in window order); sampleMeasurement (the cumulative walk
 * window by window in state order, each window's walk exact). */
int tmfqs_hip_chunked_clone(tmfqs_hip_chunked *c, tmfqs_hip_chunked **out);
int tmfqs_hip_chunked_set_amplitude(tmfqs_hip_chunked *c, uint64_t state, double re, double im);
int tmfqs_hip_chunked_affine(tmfqs_hip_chunked *c, double s_re, double s_im, double b_re, double b_im);
int tmfqs_hip_chunked_inversion_about_mean_value(tmfqs_hip_chunked *c, double mean_re, double mean_im);
int tmfqs_hip_chunked_amplitude_sum(tmfqs_hip_chunked *c, double *re, double *im);
int tmfqs_hip_chunked_sample(tmfqs_hip_chunked *c, double rnd, uint64_t *out);
int tmfqs_hip_chunked_info(tmfqs_hip_chunked *c, tmfqs_hip_chunked_stats *out);
int tmfqs_hip_chunked_destroy(tmfqs_hip_chunked *c);

/* ---- host-side checks of the randomPhase arithmetic -------------------
 * No counterpart in the reference; for tests of the device path, no device
 * needed.  tmfqs_hip_mt_jump_outputs: the count outputs of std::mt19937(seed)
 * that follow its first jump_words outputs, reached with the jump-ahead
 * polynomials the device streams use.  tmfqs_hip_phase_eval: the device's
 * double-double cos/sin of x (out = cos hi, cos lo, sin hi, sin lo) and its
 * near-midpoint flags (flags[0] cos, flags[1] sin). */
int tmfqs_hip_mt_jump_outputs(uint32_t seed, uint64_t jump_words, size_t count, uint32_t *out);
int tmfqs_hip_phase_eval(double x, double out[4], int flags[2]);

#ifdef __cplusplus
}
#endif

#endif /* TMFQS_HIP_H */
