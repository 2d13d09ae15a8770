/*
 * CPU oracle: clean-room C restatement of the reference's dense path.
 *
 * TEST INFRASTRUCTURE ONLY (see dense_oracle.h).  Each function names the
 * reference file:line whose behaviour it restates.  Arithmetic is written in
 * the same operation order as the reference and this file is compiled with
 * -ffp-contract=off so that no multiply-add is fused (the reference is built
 * for baseline x86-64, which has no FMA).
 */
#define _POSIX_C_SOURCE 199309L
#include "dense_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846 /* include/tmfqs/core/constants.h:6 */

/* ------------------------------------------------------------------------- */
/* RNG: std::mt19937 (Matsumoto & Nishimura 1998, parameters fixed by the C++
 * standard [rand.predef]) and libstdc++'s generate_canonical<double, 53>.
 * Reference use: include/tmfqs/core/random.h:32-58, src/tmfqs/core/random.cpp:6-19,
 * src/tmfqs/state/quantum_state.cpp:264-277. */

void orc_mt_seed(orc_mt19937 *g, uint32_t seed) {
	g->mt[0] = seed;
	for(int i = 1; i < 624; ++i) {
		const uint32_t prev = g->mt[i - 1];
		g->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
	}
	g->idx = 624;
}

static void orc_mt_twist(orc_mt19937 *g) {
	for(int i = 0; i < 624; ++i) {
		const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
		uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
		if(y & 1u) v ^= 0x9908b0dfu;
		g->mt[i] = v;
	}
	g->idx = 0;
}

uint32_t orc_mt_next(orc_mt19937 *g) {
	if(g->idx >= 624) orc_mt_twist(g);
	uint32_t y = g->mt[g->idx++];
	y ^= (y >> 11);
	y ^= (y << 7) & 0x9d2c5680u;
	y ^= (y << 15) & 0xefc60000u;
	y ^= (y >> 18);
	return y;
}

/* generate_canonical<double,53>: two 32-bit draws, sum = d1 + d2*2^32 in
 * double, divided by 2^64, clamped below 1; uniform_real_distribution maps it
 * as c*(b-a)+a. */
double orc_mt_uniform(orc_mt19937 *g, double a, double b) {
	double sum = 0.0;
	double tmp = 1.0;
	for(int k = 0; k < 2; ++k) {
		sum += (double)orc_mt_next(g) * tmp;
		tmp *= 4294967296.0;
	}
	double c = sum / tmp;
	if(c >= 1.0) c = nextafter(1.0, 0.0);
	return c * (b - a) + a;
}

void orc_mt_unit_draws(uint32_t seed, size_t count, double *out) {
	orc_mt19937 g;
	orc_mt_seed(&g, seed);
	for(size_t i = 0; i < count; ++i) out[i] = orc_mt_uniform(&g, 0.0, 1.0);
}

/* QuantumState::randomPhase: src/tmfqs/state/quantum_state.cpp:264-277. */
void orc_random_phase(unsigned n, uint32_t seed, double *out) {
	const uint64_t states = (uint64_t)1 << n;
	orc_mt19937 g;
	orc_mt_seed(&g, seed);
	const double hi = 2.0 * acos(-1.0);
	const double norm = 1.0 / sqrt((double)states);
	for(uint64_t s = 0; s < states; ++s) {
		const double phase = orc_mt_uniform(&g, 0.0, hi);
		out[2 * s] = norm * cos(phase);
		out[2 * s + 1] = norm * sin(phase);
	}
}

/* ------------------------------------------------------------------------- */
/* Dense backend kernels: src/tmfqs/storage/dense/dense_state_backend.cpp. */

/* backend_validation.h:37-39: qubit q -> mask 1 << (n-q-1). */
static uint64_t orc_mask(unsigned q, unsigned n) { return (uint64_t)1 << (n - q - 1u); }

/* Pair traversal, pair_kernel_executor.h:42-50: the pairs (s, s + m) with
 * bit m of s clear.  Pairs are disjoint and each is computed on its own, so
 * visiting them in any order (here: split over OpenMP threads for states of
 * 2^16 amplitudes and more) gives the same bits as the reference's
 * base-then-offset order.  The one-trip inner loop lets a body `continue`. */
#define ORC_PAR_MIN ((uint64_t)1 << 16)
#define ORC_PAIR_LOOP(N, M, S0, S1)                                                                      \
	_Pragma("omp parallel for schedule(static) if((N) >= ORC_PAR_MIN)")                                  \
	for(int64_t t_ = 0; t_ < (int64_t)((N) >> 1); ++t_)                                                  \
		for(uint64_t S0 = (((uint64_t)t_ & ~((M) - 1u)) << 1) | ((uint64_t)t_ & ((M) - 1u)), S1 = S0 + (M), \
		             once_ = 1;                                                                          \
		    once_; once_ = 0)

/* dense_state_backend.cpp:255-273 (zero pairs skipped). */
int orc_hadamard(double *a, unsigned n, unsigned qubit) {
	if(qubit >= n) return ORC_OUT_OF_RANGE;
	const uint64_t N = (uint64_t)1 << n;
	const uint64_t m = orc_mask(qubit, n);
	const double c = 1.0 / sqrt(2.0);
	ORC_PAIR_LOOP(N, m, s0, s1) {
		double *a0 = a + 2 * s0;
		double *a1 = a + 2 * s1;
		const double ar = a0[0], ai = a0[1], br = a1[0], bi = a1[1];
		if(ar == 0.0 && ai == 0.0 && br == 0.0 && bi == 0.0) continue;
		a0[0] = (ar + br) * c;
		a0[1] = (ai + bi) * c;
		a1[0] = (ar - br) * c;
		a1[1] = (ai - bi) * c;
	}
	return ORC_OK;
}

/* dense_state_backend.cpp:276-286 (equal pairs skipped). */
int orc_pauli_x(double *a, unsigned n, unsigned qubit) {
	if(qubit >= n) return ORC_OUT_OF_RANGE;
	const uint64_t N = (uint64_t)1 << n;
	const uint64_t m = orc_mask(qubit, n);
	ORC_PAIR_LOOP(N, m, s0, s1) {
		double *a0 = a + 2 * s0;
		double *a1 = a + 2 * s1;
		if(a0[0] == a1[0] && a0[1] == a1[1]) continue;
		double t = a0[0]; a0[0] = a1[0]; a1[0] = t;
		t = a0[1]; a0[1] = a1[1]; a1[1] = t;
	}
	return ORC_OK;
}

static int orc_two_qubit_check(unsigned n, unsigned q0, unsigned q1) {
	if(q0 >= n || q1 >= n) return ORC_OUT_OF_RANGE;
	if(q0 == q1) return ORC_INVALID_ARGUMENT;
	return ORC_OK;
}

/* dense_state_backend.cpp:306-318: swap pairs whose state0 has the control bit. */
int orc_controlled_not(double *a, unsigned n, unsigned control, unsigned target) {
	const int rc = orc_two_qubit_check(n, control, target);
	if(rc) return rc;
	const uint64_t N = (uint64_t)1 << n;
	const uint64_t cm = orc_mask(control, n);
	const uint64_t m = orc_mask(target, n);
	ORC_PAIR_LOOP(N, m, s0, s1) {
		if((s0 & cm) == 0) continue;
		double *a0 = a + 2 * s0;
		double *a1 = a + 2 * s1;
		if(a0[0] == a1[0] && a0[1] == a1[1]) continue;
		double t = a0[0]; a0[0] = a1[0]; a1[0] = t;
		t = a0[1]; a0[1] = a1[1]; a1[1] = t;
	}
	return ORC_OK;
}

/* dense_state_backend.cpp:289-303: multiply state1 of each pair by e^{i theta}
 * when it has the control bit (zero amplitudes skipped). */
int orc_controlled_phase(double *a, unsigned n, unsigned control, unsigned target, double theta) {
	const int rc = orc_two_qubit_check(n, control, target);
	if(rc) return rc;
	const uint64_t N = (uint64_t)1 << n;
	const uint64_t cm = orc_mask(control, n);
	const uint64_t m = orc_mask(target, n);
	const double pr = cos(theta);
	const double pi = sin(theta);
	ORC_PAIR_LOOP(N, m, s0, s1) {
		(void)s0;
		if((s1 & cm) == 0) continue;
		double *a1 = a + 2 * s1;
		const double re = a1[0], im = a1[1];
		if(re == 0.0 && im == 0.0) continue;
		a1[0] = pr * re - pi * im;
		a1[1] = pr * im + pi * re;
	}
	return ORC_OK;
}

/* quantum_register.cpp:518-527: SWAP = CX(q1,q2) CX(q2,q1) CX(q1,q2). */
int orc_swap(double *a, unsigned n, unsigned q1, unsigned q2) {
	const int rc = orc_two_qubit_check(n, q1, q2);
	if(rc) return rc;
	orc_controlled_not(a, n, q1, q2);
	orc_controlled_not(a, n, q2, q1);
	orc_controlled_not(a, n, q1, q2);
	return ORC_OK;
}

/* validation.h:56-78: size/width, dimension, then per-target range and duplicates. */
static int orc_validate_targets(unsigned n, unsigned k, const unsigned *qubits) {
	if(k > 63u) return ORC_INVALID_ARGUMENT;
	if(k > n) return ORC_INVALID_ARGUMENT;
	unsigned char seen[64];
	memset(seen, 0, sizeof(seen));
	for(unsigned i = 0; i < k; ++i) {
		if(qubits[i] >= n) return ORC_OUT_OF_RANGE;
		if(seen[qubits[i]]) return ORC_INVALID_ARGUMENT;
		seen[qubits[i]] = 1;
	}
	return ORC_OK;
}

/* Block apply: gate_block_apply.h:55-99 (layout) and :103-212 (k==1, k==2,
 * general paths with their own accumulation orders; all-zero blocks skipped). */
int orc_apply_gate(double *a, unsigned n, const double *mat, unsigned k, const unsigned *qubits) {
	const int rc = orc_validate_targets(n, k, qubits);
	if(rc) return rc;
	const uint64_t numBlocks = (uint64_t)1 << (n - k);
	const uint64_t dim = (uint64_t)1 << k;
	uint64_t tmask[64];
	int isTarget[64];
	memset(isTarget, 0, sizeof(isTarget));
	for(unsigned i = 0; i < k; ++i) {
		const unsigned bit = n - qubits[i] - 1u;
		tmask[i] = (uint64_t)1 << bit;
		isTarget[bit] = 1;
	}
	unsigned nonTarget[64];
	unsigned nNon = 0;
	for(unsigned bit = 0; bit < n; ++bit)
		if(!isTarget[bit]) nonTarget[nNon++] = bit;
	uint64_t *offsets = (uint64_t *)malloc(sizeof(uint64_t) * dim);
	if(!offsets) return ORC_RUNTIME_ERROR;
	for(uint64_t r = 0; r < dim; ++r) {
		uint64_t off = 0;
		for(unsigned bit = 0; bit < k; ++bit)
			if((r >> (k - bit - 1u)) & 1u) off |= tmask[bit];
		offsets[r] = off;
	}
	int failed = 0;
	/* Blocks are disjoint and computed on their own: any visiting order (here
	 * split over threads) gives the reference's bits. */
#pragma omp parallel if((uint64_t)1 << n >= ORC_PAR_MIN) reduction(| : failed)
	{
	double *loc = (double *)malloc(sizeof(double) * 2 * dim);
	uint64_t *idx = (uint64_t *)malloc(sizeof(uint64_t) * dim);
	if(!loc || !idx) failed = 1;
#pragma omp for schedule(static)
	for(int64_t block_ = 0; block_ < (int64_t)numBlocks; ++block_) {
		const uint64_t block = (uint64_t)block_;
		if(!loc || !idx) continue;
		uint64_t baseState = 0;
		for(unsigned j = 0; j < nNon; ++j) baseState |= ((block >> j) & 1u) << nonTarget[j];
		if(k == 1) {
			const uint64_t s0 = baseState, s1 = baseState | tmask[0];
			const double a0r = a[2 * s0], a0i = a[2 * s0 + 1];
			const double a1r = a[2 * s1], a1i = a[2 * s1 + 1];
			if((a0r == 0.0 && a0i == 0.0) && (a1r == 0.0 && a1i == 0.0)) continue;
			const double g00r = mat[0], g00i = mat[1], g01r = mat[2], g01i = mat[3];
			const double g10r = mat[4], g10i = mat[5], g11r = mat[6], g11i = mat[7];
			a[2 * s0] = g00r * a0r - g00i * a0i + g01r * a1r - g01i * a1i;
			a[2 * s0 + 1] = g00r * a0i + g00i * a0r + g01r * a1i + g01i * a1r;
			a[2 * s1] = g10r * a0r - g10i * a0i + g11r * a1r - g11i * a1i;
			a[2 * s1 + 1] = g10r * a0i + g10i * a0r + g11r * a1i + g11i * a1r;
			continue;
		}
		int nonZero = 0;
		for(uint64_t r = 0; r < dim; ++r) {
			/* k==2 uses the order {00, 01, 10, 11} with 01 = qubits[1]'s mask,
			 * which equals offsets[] order. */
			idx[r] = baseState | offsets[r];
			loc[2 * r] = a[2 * idx[r]];
			loc[2 * r + 1] = a[2 * idx[r] + 1];
			if(loc[2 * r] != 0.0 || loc[2 * r + 1] != 0.0) nonZero = 1;
		}
		if(!nonZero) continue;
		for(uint64_t r = 0; r < dim; ++r) {
			double nr = 0.0, ni = 0.0;
			const double *row = mat + 2 * dim * r;
			for(uint64_t c = 0; c < dim; ++c) {
				const double gr = row[2 * c], gi = row[2 * c + 1];
				const double xr = loc[2 * c], xi = loc[2 * c + 1];
				nr += gr * xr - gi * xi;
				ni += gr * xi + gi * xr;
			}
			a[2 * idx[r]] = nr;
			a[2 * idx[r] + 1] = ni;
		}
	}
	free(loc);
	free(idx);
	}
	free(offsets);
	return failed ? ORC_RUNTIME_ERROR : ORC_OK;
}

/* dense_state_backend.cpp:205-210. */
int orc_phase_flip(double *a, unsigned n, uint64_t state) {
	if(state >= ((uint64_t)1 << n)) return ORC_OUT_OF_RANGE;
	a[2 * state] = -a[2 * state];
	a[2 * state + 1] = -a[2 * state + 1];
	return ORC_OK;
}

/* dense_state_backend.cpp:230-240: a = 2*mean - a. */
int orc_inversion_about_mean_value(double *a, unsigned n, double mean_re, double mean_im) {
	const uint64_t N = (uint64_t)1 << n;
#pragma omp parallel for schedule(static) if(N >= ORC_PAR_MIN)
	for(uint64_t s = 0; s < N; ++s) {
		const double re = a[2 * s], im = a[2 * s + 1];
		a[2 * s] = 2.0 * mean_re - re;
		a[2 * s + 1] = 2.0 * mean_im - im;
	}
	return ORC_OK;
}

/* dense_state_backend.cpp:213-227: sequential sums, mean = sum / N. */
int orc_inversion_about_mean(double *a, unsigned n) {
	const uint64_t N = (uint64_t)1 << n;
	double sr = 0.0, si = 0.0;
	for(uint64_t s = 0; s < N; ++s) {
		sr += a[2 * s];
		si += a[2 * s + 1];
	}
	return orc_inversion_about_mean_value(a, n, sr / (double)N, si / (double)N);
}

/* dense_state_backend.cpp:242-252: a = scale*a + bias. */
int orc_affine(double *a, unsigned n, double s_re, double s_im, double b_re, double b_im) {
	const uint64_t N = (uint64_t)1 << n;
#pragma omp parallel for schedule(static) if(N >= ORC_PAR_MIN)
	for(uint64_t s = 0; s < N; ++s) {
		const double re = a[2 * s], im = a[2 * s + 1];
		a[2 * s] = s_re * re - s_im * im + b_re;
		a[2 * s + 1] = s_re * im + s_im * re + b_im;
	}
	return ORC_OK;
}

/* dense_state_backend.cpp:161-170: sequential sum of re^2 + im^2. */
double orc_total_probability(const double *a, unsigned n) {
	const uint64_t N = (uint64_t)1 << n;
	double sum = 0.0;
	for(uint64_t s = 0; s < N; ++s) sum += a[2 * s] * a[2 * s] + a[2 * s + 1] * a[2 * s + 1];
	return sum;
}

/* dense_state_backend.cpp:173-187: first state with rnd <= sequential cumulative. */
int orc_sample(const double *a, unsigned n, double rnd, uint64_t *out) {
	if(rnd < 0.0 || rnd >= 1.0) return ORC_INVALID_ARGUMENT;
	const uint64_t N = (uint64_t)1 << n;
	double cum = 0.0;
	for(uint64_t s = 0; s < N; ++s) {
		cum += a[2 * s] * a[2 * s] + a[2 * s + 1] * a[2 * s + 1];
		if(rnd <= cum) {
			*out = s;
			return ORC_OK;
		}
	}
	return ORC_RUNTIME_ERROR;
}

/* quantum_register.cpp:176-189 (no overlay): sequential sums. */
void orc_amplitude_sum(const double *a, unsigned n, double *re, double *im) {
	const uint64_t N = (uint64_t)1 << n;
	double sr = 0.0, si = 0.0;
	for(uint64_t s = 0; s < N; ++s) {
		sr += a[2 * s];
		si += a[2 * s + 1];
	}
	*re = sr;
	*im = si;
}

static int cmp_u64(const void *x, const void *y) {
	const uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
	return a < b ? -1 : (a > b ? 1 : 0);
}

/* dense_state_backend.cpp:92-115: zero, validate, sort+unique, 1/sqrt(count). */
int orc_init_uniform(double *a, unsigned n, const uint64_t *states, size_t count) {
	const uint64_t N = (uint64_t)1 << n;
	memset(a, 0, sizeof(double) * 2 * N);
	for(size_t i = 0; i < count; ++i)
		if(states[i] >= N) return ORC_OUT_OF_RANGE;
	if(count == 0) return ORC_INVALID_ARGUMENT;
	uint64_t *sel = (uint64_t *)malloc(sizeof(uint64_t) * count);
	memcpy(sel, states, sizeof(uint64_t) * count);
	qsort(sel, count, sizeof(uint64_t), cmp_u64);
	size_t u = 0;
	for(size_t i = 0; i < count; ++i)
		if(i == 0 || sel[i] != sel[u - 1]) sel[u++] = sel[i];
	const double amp = 1.0 / sqrt((double)u);
	for(size_t i = 0; i < u; ++i) {
		a[2 * sel[i]] = amp;
		a[2 * sel[i] + 1] = 0.0;
	}
	free(sel);
	return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* QuantumRegister semantics: src/tmfqs/register/quantum_register.cpp. */

orc_register *orc_reg_create(unsigned n) {
	orc_register *r = (orc_register *)calloc(1, sizeof(orc_register));
	if(!r) return NULL;
	r->n = n;
	r->N = (uint64_t)1 << n;
	r->amp = (double *)calloc(2 * r->N, sizeof(double));
	if(!r->amp) {
		free(r);
		return NULL;
	}
	r->amp[0] = 1.0;
	r->sc_re = 1.0;
	return r;
}

void orc_reg_free(orc_register *r) {
	if(!r) return;
	free(r->amp);
	free(r);
}

/* :152-158 resetAffineOverlay */
static void orc_reg_reset_overlay(orc_register *r) {
	r->sc_re = 1.0; r->sc_im = 0.0;
	r->bi_re = 0.0; r->bi_im = 0.0;
	r->overlay = 0;
	r->has_sum = 0;
}

int orc_reg_load(orc_register *r, const double *amps) {
	orc_reg_reset_overlay(r);
	memcpy(r->amp, amps, sizeof(double) * 2 * r->N);
	return ORC_OK;
}

int orc_reg_init_basis(orc_register *r, uint64_t state, double re, double im) {
	if(state >= r->N) return ORC_OUT_OF_RANGE;
	orc_reg_reset_overlay(r);
	memset(r->amp, 0, sizeof(double) * 2 * r->N);
	r->amp[2 * state] = re;
	r->amp[2 * state + 1] = im;
	return ORC_OK;
}

/* :161-166 applyAffineOverlay: scale*a + bias (multiplyAmplitude then add). */
static void orc_apply_overlay(const orc_register *r, double br, double bi, double *lr, double *li) {
	if(!r->overlay) {
		*lr = br;
		*li = bi;
		return;
	}
	const double mr = r->sc_re * br - r->sc_im * bi;
	const double mi = r->sc_re * bi + r->sc_im * br;
	*lr = mr + r->bi_re;
	*li = mi + r->bi_im;
}

/* :168-174 removeAffineOverlay: (l - bias) / scale (divideAmplitude, :38-47). */
static int orc_remove_overlay(const orc_register *r, double lr, double li, double *br, double *bi) {
	if(!r->overlay) {
		*br = lr;
		*bi = li;
		return ORC_OK;
	}
	const double nr = lr - r->bi_re, ni = li - r->bi_im;
	const double d = r->sc_re * r->sc_re + r->sc_im * r->sc_im;
	if(d == 0.0) return ORC_LOGIC_ERROR;
	*br = (nr * r->sc_re + ni * r->sc_im) / d;
	*bi = (ni * r->sc_re - nr * r->sc_im) / d;
	return ORC_OK;
}

/* :224-231 flushAffineOverlay (keeps the cached sum). */
static void orc_reg_flush(orc_register *r) {
	if(!r->overlay) return;
	const int has = r->has_sum;
	const double sr = r->sum_re, si = r->sum_im;
	orc_affine(r->amp, r->n, r->sc_re, r->sc_im, r->bi_re, r->bi_im);
	orc_reg_reset_overlay(r);
	r->has_sum = has;
	r->sum_re = sr;
	r->sum_im = si;
}

int orc_reg_export(const orc_register *r, double *out) {
	for(uint64_t s = 0; s < r->N; ++s) orc_apply_overlay(r, r->amp[2 * s], r->amp[2 * s + 1], &out[2 * s], &out[2 * s + 1]);
	return ORC_OK;
}

/* :249-253 */
int orc_reg_amplitude(const orc_register *r, uint64_t state, double *re, double *im) {
	if(state >= r->N) return ORC_OUT_OF_RANGE;
	orc_apply_overlay(r, r->amp[2 * state], r->amp[2 * state + 1], re, im);
	return ORC_OK;
}

/* :176-189 logicalAmplitudeSum (caches). */
static void orc_reg_logical_sum(orc_register *r, double *re, double *im) {
	if(r->has_sum) {
		*re = r->sum_re;
		*im = r->sum_im;
		return;
	}
	double sr = 0.0, si = 0.0;
	for(uint64_t s = 0; s < r->N; ++s) {
		double lr, li;
		orc_apply_overlay(r, r->amp[2 * s], r->amp[2 * s + 1], &lr, &li);
		sr += lr;
		si += li;
	}
	r->has_sum = 1;
	r->sum_re = sr;
	r->sum_im = si;
	*re = sr;
	*im = si;
}

/* :269-281 */
int orc_reg_total_probability(const orc_register *r, double *out) {
	if(!r->overlay) {
		*out = orc_total_probability(r->amp, r->n);
		return ORC_OK;
	}
	double t = 0.0;
	for(uint64_t s = 0; s < r->N; ++s) {
		double lr, li;
		orc_apply_overlay(r, r->amp[2 * s], r->amp[2 * s + 1], &lr, &li);
		t += lr * lr + li * li;
	}
	*out = t;
	return ORC_OK;
}

/* :284-310 (the overlay branch does not range-check rnd). */
int orc_reg_measure(const orc_register *r, double rnd, uint64_t *out) {
	if(!r->overlay) return orc_sample(r->amp, r->n, rnd, out);
	double cum = 0.0;
	for(uint64_t s = 0; s < r->N; ++s) {
		double lr, li;
		orc_apply_overlay(r, r->amp[2 * s], r->amp[2 * s + 1], &lr, &li);
		cum += lr * lr + li * li;
		if(rnd <= cum) {
			*out = s;
			return ORC_OK;
		}
	}
	return ORC_RUNTIME_ERROR;
}

/* :470-515 gate entry points: validate, flush overlay, apply, drop cached sum. */
int orc_reg_hadamard(orc_register *r, unsigned q) {
	if(q >= r->n) return ORC_OUT_OF_RANGE;
	orc_reg_flush(r);
	orc_hadamard(r->amp, r->n, q);
	r->has_sum = 0;
	return ORC_OK;
}

int orc_reg_pauli_x(orc_register *r, unsigned q) {
	if(q >= r->n) return ORC_OUT_OF_RANGE;
	orc_reg_flush(r);
	orc_pauli_x(r->amp, r->n, q);
	r->has_sum = 0;
	return ORC_OK;
}

int orc_reg_controlled_not(orc_register *r, unsigned c, unsigned t) {
	const int rc = orc_two_qubit_check(r->n, c, t);
	if(rc) return rc;
	orc_reg_flush(r);
	orc_controlled_not(r->amp, r->n, c, t);
	r->has_sum = 0;
	return ORC_OK;
}

int orc_reg_controlled_phase(orc_register *r, unsigned c, unsigned t, double theta) {
	const int rc = orc_two_qubit_check(r->n, c, t);
	if(rc) return rc;
	orc_reg_flush(r);
	orc_controlled_phase(r->amp, r->n, c, t, theta);
	r->has_sum = 0;
	return ORC_OK;
}

/* :518-527 */
int orc_reg_swap(orc_register *r, unsigned q1, unsigned q2) {
	const int rc = orc_two_qubit_check(r->n, q1, q2);
	if(rc) return rc;
	orc_reg_controlled_not(r, q1, q2);
	orc_reg_controlled_not(r, q2, q1);
	orc_reg_controlled_not(r, q1, q2);
	return ORC_OK;
}

/* :394-402 */
int orc_reg_apply_gate(orc_register *r, const double *mat, unsigned k, const unsigned *qubits) {
	const int rc = orc_validate_targets(r->n, k, qubits);
	if(rc) return rc;
	orc_reg_flush(r);
	orc_apply_gate(r->amp, r->n, mat, k, qubits);
	r->has_sum = 0;
	return ORC_OK;
}

/* :405-420 */
int orc_reg_phase_flip(orc_register *r, uint64_t state) {
	if(state >= r->N) return ORC_OUT_OF_RANGE;
	if(r->has_sum || r->overlay) {
		double lr, li;
		orc_apply_overlay(r, r->amp[2 * state], r->amp[2 * state + 1], &lr, &li);
		if(r->has_sum) {
			r->sum_re -= 2.0 * lr;
			r->sum_im -= 2.0 * li;
		}
		if(r->overlay) {
			double br, bi;
			const int rc = orc_remove_overlay(r, -lr, -li, &br, &bi);
			if(rc) return rc;
			r->amp[2 * state] = br;
			r->amp[2 * state + 1] = bi;
			return ORC_OK;
		}
	}
	return orc_phase_flip(r->amp, r->n, state);
}

/* :433-444 */
int orc_reg_inversion_lazy_value(orc_register *r, double m_re, double m_im) {
	if(r->has_sum) {
		r->sum_re = 2.0 * m_re * (double)r->N - r->sum_re;
		r->sum_im = 2.0 * m_im * (double)r->N - r->sum_im;
	}
	const double scr = r->overlay ? r->sc_re : 1.0, sci = r->overlay ? r->sc_im : 0.0;
	const double bir = r->overlay ? r->bi_re : 0.0, bii = r->overlay ? r->bi_im : 0.0;
	r->sc_re = -scr;
	r->sc_im = -sci;
	r->bi_re = m_re * 2.0 - bir;
	r->bi_im = m_im * 2.0 - bii;
	r->overlay = 1;
	return ORC_OK;
}

/* :423-430 */
int orc_reg_inversion_lazy(orc_register *r) {
	double sr, si;
	orc_reg_logical_sum(r, &sr, &si);
	return orc_reg_inversion_lazy_value(r, sr / (double)r->N, si / (double)r->N);
}

/* :457-467 */
int orc_reg_inversion_materialized_value(orc_register *r, double m_re, double m_im) {
	if(r->has_sum) {
		r->sum_re = 2.0 * m_re * (double)r->N - r->sum_re;
		r->sum_im = 2.0 * m_im * (double)r->N - r->sum_im;
	}
	orc_reg_flush(r);
	return orc_inversion_about_mean_value(r->amp, r->n, m_re, m_im);
}

/* :447-454 */
int orc_reg_inversion_materialized(orc_register *r) {
	double sr, si;
	orc_reg_logical_sum(r, &sr, &si);
	return orc_reg_inversion_materialized_value(r, sr / (double)r->N, si / (double)r->N);
}

/* simulator.cpp:17-44 executeOperation dispatch. */
int orc_reg_run(orc_register *r, const orc_op *ops, size_t count, const double *mat_pool, const unsigned *mat_qubits) {
	for(size_t i = 0; i < count; ++i) {
		const orc_op *op = &ops[i];
		int rc = ORC_OK;
		switch(op->kind) {
			case ORC_OP_H: rc = orc_reg_hadamard(r, op->q0); break;
			case ORC_OP_X: rc = orc_reg_pauli_x(r, op->q0); break;
			case ORC_OP_CX: rc = orc_reg_controlled_not(r, op->q0, op->q1); break;
			case ORC_OP_CPHASE: rc = orc_reg_controlled_phase(r, op->q0, op->q1, op->theta); break;
			case ORC_OP_SWAP: rc = orc_reg_swap(r, op->q0, op->q1); break;
			case ORC_OP_PHASE_FLIP: rc = orc_reg_phase_flip(r, op->state); break;
			case ORC_OP_DIFFUSE_LAZY: rc = orc_reg_inversion_lazy(r); break;
			case ORC_OP_DIFFUSE_MATERIALIZED: rc = orc_reg_inversion_materialized(r); break;
			case ORC_OP_MATRIX:
				rc = orc_reg_apply_gate(r, mat_pool + op->state, op->k, mat_qubits + op->q0);
				break;
			case ORC_OP_PHASE: {
				/* quantum_gate.cpp:162-169: [[1, 0], [0, complexExp(i theta)]] */
				const double m[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, cos(op->theta), sin(op->theta)};
				const unsigned q = op->q0;
				rc = orc_reg_apply_gate(r, m, 1, &q);
				break;
			}
			default: rc = ORC_INVALID_ARGUMENT; break;
		}
		if(rc) return rc;
	}
	return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Circuit builders: src/tmfqs/circuit/circuit.cpp. */

/* :87-100 appendQft */
size_t orc_qft_op_count(unsigned n) {
	return (size_t)n + (size_t)n * (n - (n ? 1u : 0u)) / 2u + n / 2u;
}

void orc_qft_ops(unsigned n, orc_op *out) {
	size_t k = 0;
	for(unsigned t = 0; t < n; ++t) {
		memset(&out[k], 0, sizeof(orc_op));
		out[k].kind = ORC_OP_H;
		out[k].q0 = t;
		++k;
		for(unsigned c = t + 1u; c < n; ++c) {
			memset(&out[k], 0, sizeof(orc_op));
			out[k].kind = ORC_OP_CPHASE;
			out[k].q0 = c;
			out[k].q1 = t;
			out[k].theta = ORC_PI / (double)((uint64_t)1 << (c - t));
			++k;
		}
	}
	for(unsigned l = 0; l < n / 2u; ++l) {
		memset(&out[k], 0, sizeof(orc_op));
		out[k].kind = ORC_OP_SWAP;
		out[k].q0 = l;
		out[k].q1 = n - l - 1u;
		++k;
	}
}

/* :108-116 groverIterationCount */
unsigned orc_grover_iterations(uint64_t state_count, size_t marked_count) {
	const double ideal = (ORC_PI / 4.0) * sqrt((double)state_count / (double)marked_count);
	return (unsigned)floor(ideal);
}

/* :118-141 appendGrover: sort+unique, H on all, iterations of (flips, diffusion);
 * the marked-range check runs inside the iteration loop. */
long long orc_grover_ops(unsigned n, const uint64_t *marked, size_t m, int materialized, orc_op *out, size_t cap) {
	if(m == 0) return ORC_INVALID_ARGUMENT;
	uint64_t *sel = (uint64_t *)malloc(sizeof(uint64_t) * m);
	memcpy(sel, marked, sizeof(uint64_t) * m);
	qsort(sel, m, sizeof(uint64_t), cmp_u64);
	size_t u = 0;
	for(size_t i = 0; i < m; ++i)
		if(i == 0 || sel[i] != sel[u - 1]) sel[u++] = sel[i];
	const uint64_t N = (uint64_t)1 << n;
	const unsigned iters = orc_grover_iterations(N, u);
	const long long total = (long long)n + (long long)iters * (long long)(u + 1);
	if(iters > 0 && sel[u - 1] >= N) {
		free(sel);
		return ORC_INVALID_ARGUMENT;
	}
	if(out) {
		if((size_t)total > cap) {
			free(sel);
			return ORC_INVALID_ARGUMENT;
		}
		size_t k = 0;
		for(unsigned q = 0; q < n; ++q) {
			memset(&out[k], 0, sizeof(orc_op));
			out[k].kind = ORC_OP_H;
			out[k].q0 = q;
			++k;
		}
		for(unsigned it = 0; it < iters; ++it) {
			for(size_t i = 0; i < u; ++i) {
				memset(&out[k], 0, sizeof(orc_op));
				out[k].kind = ORC_OP_PHASE_FLIP;
				out[k].state = sel[i];
				++k;
			}
			memset(&out[k], 0, sizeof(orc_op));
			out[k].kind = materialized ? ORC_OP_DIFFUSE_MATERIALIZED : ORC_OP_DIFFUSE_LAZY;
			++k;
		}
	}
	free(sel);
	return total;
}

/* ------------------------------------------------------------------------- */
/* CPU-baseline timing of the restatement: ops applied to a dense n-qubit
 * register, real amplitudes 2^(-n/2) (1 + d_i) with distinct small d_i (so no
 * zero-skip or equal-pair skip fires, as on a random-phase input); wall
 * seconds of the op loop only, and each operation's own seconds when per_op
 * is not null.
 * Runs on one thread, as the reference's dense backend does.  Used by the
 * benchmark when the reference build is absent. */
#include <omp.h>
#include <time.h>
static double orc_now(void) {
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int orc_time_ops_each(unsigned n, const orc_op *ops, size_t count, double *seconds, double *per_op) {
	orc_register *r = orc_reg_create(n);
	if(!r) return ORC_RUNTIME_ERROR;
	const double a = 1.0 / sqrt((double)r->N);
	for(uint64_t s = 0; s < r->N; ++s) {
		uint64_t x = s * 0x9E3779B97F4A7C15ull;
		x ^= x >> 29;
		x *= 0xBF58476D1CE4E5B9ull;
		x ^= x >> 32;
		r->amp[2 * s] = a * (1.0 + 0x1p-24 * (double)(x & 0xFFFFFull));
		r->amp[2 * s + 1] = 0.0;
	}
	const int threads = omp_get_max_threads();
	omp_set_num_threads(1);
	int rc = ORC_OK;
	const double t0 = orc_now();
	double prev = t0;
	for(size_t i = 0; i < count && rc == ORC_OK; ++i) {
		rc = orc_reg_run(r, ops + i, 1, NULL, NULL);
		if(per_op) {
			const double t = orc_now();
			per_op[i] = t - prev;
			prev = t;
		}
	}
	*seconds = orc_now() - t0;
	omp_set_num_threads(threads);
	orc_reg_free(r);
	return rc;
}

int orc_time_ops(unsigned n, const orc_op *ops, size_t count, double *seconds) {
	return orc_time_ops_each(n, ops, count, seconds, NULL);
}

/* ------------------------------------------------------------------------- */
/* Diagonal observable expectation and the layered ring circuit (no
 * counterpart in the reference; they define the adjoint benchmark). */
int orc_expectation(const double *a, unsigned n, const orc_diag_term *terms, size_t nterms, double *out) {
	if(n > 62) return ORC_INVALID_ARGUMENT;
	const uint64_t N = 1ull << n;
	for(size_t t = 0; t < nterms; ++t) {
		if(terms[t].kind == 0) {
			if(n < 64 && (terms[t].bits >> n) != 0) return ORC_OUT_OF_RANGE;
		} else if(terms[t].kind == 1) {
			if(terms[t].bits >= N) return ORC_OUT_OF_RANGE;
		} else {
			return ORC_INVALID_ARGUMENT;
		}
	}
	double e = 0.0;
	for(uint64_t x = 0; x < N; ++x) {
		double o = 0.0;
		for(size_t t = 0; t < nterms; ++t) {
			if(terms[t].kind == 0) {
				int parity = 0;
				for(unsigned q = 0; q < n; ++q)
					if((terms[t].bits >> q) & 1u) parity ^= (int)((x >> (n - 1 - q)) & 1u);
				o += parity ? -terms[t].coef : terms[t].coef;
			} else if(x == terms[t].bits) {
				o += terms[t].coef;
			}
		}
		const double re = a[2 * x], im = a[2 * x + 1];
		e += o * (re * re + im * im);
	}
	*out = e;
	return ORC_OK;
}

size_t orc_ring_ops(unsigned n, unsigned layers, uint32_t seed, orc_op *out) {
	orc_mt19937 g;
	orc_mt_seed(&g, seed);
	const double hi = 2.0 * acos(-1.0);
	size_t c = 0;
	for(unsigned l = 0; l < layers; ++l) {
		for(unsigned q = 0; q < n; ++q) {
			orc_op *o = &out[c++];
			memset(o, 0, sizeof *o);
			o->kind = ORC_OP_H;
			o->q0 = q;
		}
		for(unsigned i = 0; i < n; ++i) {
			orc_op *o = &out[c++];
			memset(o, 0, sizeof *o);
			o->kind = ORC_OP_CPHASE;
			o->q0 = i;
			o->q1 = (i + 1) % n;
			o->theta = orc_mt_uniform(&g, 0.0, hi);
		}
	}
	return c;
}
