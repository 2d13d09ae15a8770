/*
 * tmfqs_sim.h -- C entry points into the C++ API (libtmfqs.so), for callers
 * that bind C (Python ctypes, the benchmark).  Each function drives the same
 * objects a C++ caller would: QuantumState -> QuantumRegister with
 * StorageStrategyKind::Hip -> sim::Simulator -> circuit::Circuit, so results
 * carry the reference's register semantics (cached amplitude sum, lazy
 * affine overlay, swap, measurement).  The reference exposes no FFI; these
 * mirror sim::Simulator::run (src/tmfqs/sim/simulator.cpp:54-96) and the
 * QuantumState / QuantumRegister queries a binding would need.
 *
 * Status codes are those of tmfqs_hip.h; tmfqs_sim_last_error() describes the
 * calling thread's last failure.  No exception crosses this boundary.
 */
#ifndef TMFQS_SIM_H
#define TMFQS_SIM_H

#include <stddef.h>
#include <stdint.h>

#include "tmfqs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tmfqs_reg tmfqs_reg;

const char *tmfqs_sim_last_error(void);

/* QuantumRegister(n, state, cfg{strategy = Hip}) (quantum_register.cpp:85-89). */
int tmfqs_reg_create_basis(unsigned n, uint64_t state, const tmfqs_hip_config *cfg, tmfqs_reg **out);
/* QuantumRegister(n, amplitudes, cfg) (quantum_register.cpp:99-103); 2 * 2^n doubles. */
int tmfqs_reg_create_amplitudes(unsigned n, const double *amps, const tmfqs_hip_config *cfg, tmfqs_reg **out);
/* QuantumState::randomPhase(n, seed) materialised on the device (quantum_state.cpp:264-277). */
int tmfqs_reg_create_random_phase(unsigned n, unsigned seed, const tmfqs_hip_config *cfg, tmfqs_reg **out);
/* QuantumRegister(n, state, cfg{strategy = HipChunked, hip.chunkStates =
 * chunk_states, hip.windowChunks = window_chunks}): a register held as chunks
 * in host memory (0: the defaults of HipConfig); the reference's chunked
 * strategies are selected the same way (register_config.h, BloscConfig). */
int tmfqs_reg_create_chunked(unsigned n, uint64_t state, const tmfqs_hip_config *cfg, uint64_t chunk_states,
                             uint64_t window_chunks, tmfqs_reg **out);
/* QuantumRegister copy constructor (deep copy). */
int tmfqs_reg_clone(tmfqs_reg *r, tmfqs_reg **out);
int tmfqs_reg_free(tmfqs_reg *r);

/* The timed region of Simulator::run (simulator.cpp:66-93) on this register:
 * ops (see tmfqs_hip_op; MATRIX ops take their matrix from mat_pool at
 * op.state doubles and their k targets from mat_qubits at op.q0) inside one
 * operation batch.  seconds receives RunReport.executionSeconds; trace runs
 * every operation on its own (ObservabilityConfig.traceOperations). */
int tmfqs_reg_run(tmfqs_reg *r, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *mat_qubits,
                  int trace, double *seconds);

/* As tmfqs_reg_run with the per-operation trace on: op_seconds receives
 * RunReport.operations[i].seconds for each of the count operations (their
 * device time on a Hip register). */
int tmfqs_reg_run_traced(tmfqs_reg *r, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *mat_qubits,
                         double *op_seconds, double *seconds);

/* Full Simulator::run from QuantumState::basis(n, basis) (init_amps == NULL)
 * or QuantumState::fromAmplitudes(n, init_amps, AllowUnnormalized); the result
 * register is returned in *result. */
int tmfqs_sim_run(unsigned n, uint64_t basis, const double *init_amps, const tmfqs_hip_config *cfg, const tmfqs_hip_op *ops,
                  size_t count, const double *mat_pool, const unsigned *mat_qubits, double *seconds, tmfqs_reg **result);

/* QuantumRegister queries. */
int tmfqs_reg_qubit_count(tmfqs_reg *r, unsigned *out);
int tmfqs_reg_export(tmfqs_reg *r, double *out); /* amplitudes(): 2 * 2^n doubles */
int tmfqs_reg_amplitude(tmfqs_reg *r, uint64_t state, double *re, double *im);
int tmfqs_reg_total_probability(tmfqs_reg *r, double *out);
/* measure(Mt19937RandomSource(seed)). */
int tmfqs_reg_measure_seed(tmfqs_reg *r, uint32_t seed, uint64_t *out);
/* measure() with a source that always returns rnd (no range check of its own). */
int tmfqs_reg_measure_fixed(tmfqs_reg *r, double rnd, uint64_t *out);
/* The device handle behind the register (profiling, streams). */
int tmfqs_reg_hip_handle(tmfqs_reg *r, tmfqs_hip_state **out);

/* QuantumState::randomPhase(n, seed).amplitudes() into out (2 * 2^n doubles);
 * host only (quantum_state.cpp:264-277). */
int tmfqs_state_random_phase(unsigned n, unsigned seed, double *out);

/* circuit::makeQft / makeGrover as op records (circuit.cpp:87-147).  With out
 * == NULL only *count is written. */
int tmfqs_circuit_qft(unsigned n, tmfqs_hip_op *out, size_t cap, size_t *count);
int tmfqs_circuit_grover(unsigned n, const uint64_t *marked, size_t m, int materialized, tmfqs_hip_op *out, size_t cap,
                         size_t *count);

/* circuit::makeRingLayers(n, layers, seed): the adjoint benchmark circuit. */
int tmfqs_circuit_ring(unsigned n, unsigned layers, unsigned seed, tmfqs_hip_op *out, size_t cap, size_t *count);

/* sim::adjointGradient on this register (ops as for tmfqs_reg_run, gate
 * kinds only); grads receives one value per CPHASE op (grad_count of them);
 * seconds receives AdjointResult.executionSeconds. */
int tmfqs_reg_adjoint(tmfqs_reg *r, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *mat_qubits,
                      const tmfqs_hip_diag_term *terms, size_t nterms, unsigned checkpoint_gates, double *expval,
                      double *grads, size_t grad_count, double *seconds);
/* sim::expectation of the register's state. */
int tmfqs_reg_expectation(tmfqs_reg *r, const tmfqs_hip_diag_term *terms, size_t nterms, double *out);

/* experiment::compareStates (src/tmfqs/experiment/comparison.cpp:12-63) of
 * two simulations of one circuit: ops run through sim::Simulator from
 * basis(n, basis) (or init_amps when given) once with ref_cfg and once with
 * cand_cfg, and the two result states compared (on the device when both are
 * Hip registers).  out's overlap and fidelity fields are left zero. */
int tmfqs_sim_compare(unsigned n, uint64_t basis, const double *init_amps, const tmfqs_hip_config *ref_cfg,
                      const tmfqs_hip_config *cand_cfg, const tmfqs_hip_op *ops, size_t count, const double *mat_pool,
                      const unsigned *mat_qubits, tmfqs_hip_comparison *out);
/* experiment::compareStates of two host amplitude vectors (2 * 2^n doubles
 * each), as QuantumState::fromAmplitudes values. */
int tmfqs_compare_amplitudes(unsigned n, const double *ref_amps, const double *cand_amps, tmfqs_hip_comparison *out);

/* QuantumGate factories (src/tmfqs/gates/quantum_gate.cpp:96-210 in the
 * reference), by index: 0 Identity(dim), 1 Hadamard, 2 PauliX, 3 PauliY,
 * 4 PauliZ, 5 PhaseShift(theta), 6 PiOverEight, 7 ControlledNot, 8 Toffoli,
 * 9 ControlledPhaseShift(theta), 10 Swap, 11 Ising(theta).  out receives
 * the row-major interleaved matrix (2 * dim^2 doubles, at most cap), *dim
 * its dimension. */
int tmfqs_gate_matrix(int which, double theta, unsigned identity_dim, double *out, size_t cap, unsigned *dim);
/* QuantumGate products: out = A * B (matrix) or A * (s_re + i s_im) (scalar
 * when b == NULL); A, B, out row-major interleaved dim x dim. */
int tmfqs_gate_product(unsigned dim, const double *a, const double *b, double s_re, double s_im, double *out);

#ifdef __cplusplus
}
#endif

#endif /* TMFQS_SIM_H */
