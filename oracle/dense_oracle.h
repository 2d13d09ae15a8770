/*
 * CPU oracle for the dense state-vector path: a clean-room restatement in C of
 * the reference's dense backend, register semantics, circuit builders and RNG.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product (libtmfqs_hip.so and the C++ API) never links or calls it.
 *
 * Conventions (reference file:line):
 *   - amplitudes are interleaved FP64 [re0, im0, re1, im1, ...]
 *     (include/tmfqs/core/types.h:20-28);
 *   - qubit q is bit position n-1-q (MSB first)
 *     (include/tmfqs/storage/common/backend_validation.h:37-39);
 *   - in a k-qubit matrix gate, qubits[0] is the MSB of the local row/column
 *     index (include/tmfqs/storage/gate_block_apply.h:76-96).
 *
 * Return codes mirror the exception classes the reference throws:
 *   0 ok, -1 std::out_of_range, -2 std::invalid_argument, -3 std::logic_error,
 *   -4 std::runtime_error.
 */
#ifndef TMFQS_ORACLE_DENSE_ORACLE_H
#define TMFQS_ORACLE_DENSE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	ORC_OK = 0,
	ORC_OUT_OF_RANGE = -1,
	ORC_INVALID_ARGUMENT = -2,
	ORC_LOGIC_ERROR = -3,
	ORC_RUNTIME_ERROR = -4
};

/* ---- std::mt19937 + uniform_real_distribution<double> (libstdc++) ---- */
typedef struct {
	uint32_t mt[624];
	int idx;
} orc_mt19937;

void orc_mt_seed(orc_mt19937 *g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937 *g);
/* uniform_real_distribution<double>(a, b) draw via generate_canonical<double,53>. */
double orc_mt_uniform(orc_mt19937 *g, double a, double b);
/* Mt19937RandomSource(seed).nextUnitDouble() x count. */
void orc_mt_unit_draws(uint32_t seed, size_t count, double *out);
/* QuantumState::randomPhase(n, seed) amplitudes (2 * 2^n doubles). */
void orc_random_phase(unsigned n, uint32_t seed, double *out);

/* ---- dense backend kernels on a raw interleaved buffer of 2^n states ---- */
int orc_hadamard(double *a, unsigned n, unsigned qubit);
int orc_pauli_x(double *a, unsigned n, unsigned qubit);
int orc_controlled_not(double *a, unsigned n, unsigned control, unsigned target);
int orc_controlled_phase(double *a, unsigned n, unsigned control, unsigned target, double theta);
/* Register-level swap: three controlled-NOTs. */
int orc_swap(double *a, unsigned n, unsigned q1, unsigned q2);
/* mat: row-major interleaved 2^k x 2^k complex matrix. */
int orc_apply_gate(double *a, unsigned n, const double *mat, unsigned k, const unsigned *qubits);
int orc_phase_flip(double *a, unsigned n, uint64_t state);
int orc_inversion_about_mean_value(double *a, unsigned n, double mean_re, double mean_im);
int orc_inversion_about_mean(double *a, unsigned n);
int orc_affine(double *a, unsigned n, double s_re, double s_im, double b_re, double b_im);
double orc_total_probability(const double *a, unsigned n);
int orc_sample(const double *a, unsigned n, double rnd, uint64_t *out);
/* Sequential sum of all amplitudes (the register's logicalAmplitudeSum sweep). */
void orc_amplitude_sum(const double *a, unsigned n, double *re, double *im);
int orc_init_uniform(double *a, unsigned n, const uint64_t *states, size_t count);

/* ---- circuit operations (shared layout with the C-ABI's op record) ---- */
enum {
	ORC_OP_H = 0,
	ORC_OP_X = 1,
	ORC_OP_CX = 2,
	ORC_OP_CPHASE = 3,
	ORC_OP_SWAP = 4,
	ORC_OP_PHASE_FLIP = 5,
	ORC_OP_DIFFUSE_LAZY = 6,
	ORC_OP_DIFFUSE_MATERIALIZED = 7,
	ORC_OP_MATRIX = 8,
	ORC_OP_PHASE = 9  /* QuantumGate::PhaseShift(theta) on qubit q0 through applyGate */
};

typedef struct {
	int32_t kind;
	uint32_t q0;
	uint32_t q1;
	uint32_t k;       /* ORC_OP_MATRIX: number of target qubits (targets in q0, q1, ... see mat_qubits) */
	double theta;     /* ORC_OP_CPHASE angle */
	uint64_t state;   /* ORC_OP_PHASE_FLIP basis state; ORC_OP_MATRIX: offset into the matrix pool */
} orc_op;

/* ---- QuantumRegister semantics over a dense buffer ---- */
typedef struct {
	unsigned n;
	uint64_t N;
	double *amp;
	int has_sum;
	double sum_re, sum_im;
	int overlay;
	double sc_re, sc_im, bi_re, bi_im;
} orc_register;

orc_register *orc_reg_create(unsigned n);
void orc_reg_free(orc_register *r);
int orc_reg_load(orc_register *r, const double *amps);
int orc_reg_init_basis(orc_register *r, uint64_t state, double re, double im);
/* Logical amplitudes (overlay applied) into out. */
int orc_reg_export(const orc_register *r, double *out);
int orc_reg_amplitude(const orc_register *r, uint64_t state, double *re, double *im);
int orc_reg_total_probability(const orc_register *r, double *out);
int orc_reg_measure(const orc_register *r, double rnd, uint64_t *out);
int orc_reg_hadamard(orc_register *r, unsigned q);
int orc_reg_pauli_x(orc_register *r, unsigned q);
int orc_reg_controlled_not(orc_register *r, unsigned c, unsigned t);
int orc_reg_controlled_phase(orc_register *r, unsigned c, unsigned t, double theta);
int orc_reg_swap(orc_register *r, unsigned q1, unsigned q2);
int orc_reg_apply_gate(orc_register *r, const double *mat, unsigned k, const unsigned *qubits);
int orc_reg_phase_flip(orc_register *r, uint64_t state);
int orc_reg_inversion_lazy(orc_register *r);
int orc_reg_inversion_lazy_value(orc_register *r, double m_re, double m_im);
int orc_reg_inversion_materialized(orc_register *r);
int orc_reg_inversion_materialized_value(orc_register *r, double m_re, double m_im);
/* Runs ops in order (Simulator::executeOperation dispatch).  mat_pool holds
 * ORC_OP_MATRIX matrices (op.state = double offset) and mat_qubits holds their
 * targets (op.q0 = offset into mat_qubits, op.k = count). */
int orc_reg_run(orc_register *r, const orc_op *ops, size_t count,
                const double *mat_pool, const unsigned *mat_qubits);

/* ---- circuit builders (circuit.cpp) ---- */
/* Number of ops of makeQft(n). */
size_t orc_qft_op_count(unsigned n);
void orc_qft_ops(unsigned n, orc_op *out);
unsigned orc_grover_iterations(uint64_t state_count, size_t marked_count);
/* Ops of makeGrover: returns count, or a negative error; out may be NULL to query. */
long long orc_grover_ops(unsigned n, const uint64_t *marked, size_t m, int materialized, orc_op *out, size_t cap);
/* Wall seconds of ops on a uniform-amplitude n-qubit register (CPU-baseline timing). */
int orc_time_ops(unsigned n, const orc_op *ops, size_t count, double *seconds);
int orc_time_ops_each(unsigned n, const orc_op *ops, size_t count, double *seconds, double *per_op);

/* ---- diagonal observables and the adjoint benchmark circuit ---- */
/* O = sum of Z-strings coef * prod_{q in bits} Z_q (Z_q = +1 when qubit q is 0)
 * and projectors coef * |bits><bits|; *out = sum_x O(x) |a_x|^2 in index order. */
typedef struct {
	int32_t kind; /* 0: Z-string (bits = qubit mask), 1: projector (bits = basis state) */
	uint32_t reserved;
	double coef;
	uint64_t bits;
} orc_diag_term;
int orc_expectation(const double *a, unsigned n, const orc_diag_term *terms, size_t nterms, double *out);
/* `layers` x (H on every qubit, then controlledPhase(i, (i+1) mod n, theta) for
 * i = 0..n-1), theta from mt19937(seed) uniform [0, 2 acos(-1)) in order
 * n * l + i.  Writes 2 n layers ops; n >= 2. */
size_t orc_ring_ops(unsigned n, unsigned layers, uint32_t seed, orc_op *out);

#ifdef __cplusplus
}
#endif

#endif
