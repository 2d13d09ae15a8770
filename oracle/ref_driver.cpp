// Driver that exposes the reference library (compiled from the reference's
// own sources by oracle/Makefile into oracle/_ref/) through a C ABI, so that
// tests can pin the oracle against the real reference and bench.py can time
// the real reference on the host as the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY: built against the reference's headers; never
// linked into the product.
#include "tmfqsfs.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "dense_oracle.h"

namespace {

int mapException() {
	try {
		throw;
	} catch(const std::out_of_range &) {
		return ORC_OUT_OF_RANGE;
	} catch(const std::invalid_argument &) {
		return ORC_INVALID_ARGUMENT;
	} catch(const std::logic_error &) {
		return ORC_LOGIC_ERROR;
	} catch(const std::runtime_error &) {
		return ORC_RUNTIME_ERROR;
	} catch(...) {
		return -99;
	}
}

tmfqs::circuit::Circuit buildCircuit(unsigned n, const orc_op *ops, size_t count, const double *matPool, const unsigned *matQubits) {
	using namespace tmfqs;
	circuit::Circuit c(n);
	for(size_t i = 0; i < count; ++i) {
		const orc_op &op = ops[i];
		switch(op.kind) {
			case ORC_OP_H: c.h(op.q0); break;
			case ORC_OP_X: c.x(op.q0); break;
			case ORC_OP_CX: c.cx(op.q0, op.q1); break;
			case ORC_OP_CPHASE: c.controlledPhase(op.q0, op.q1, op.theta); break;
			case ORC_OP_SWAP: c.swap(op.q0, op.q1); break;
			case ORC_OP_PHASE_FLIP: c.phaseFlip(op.state); break;
			case ORC_OP_DIFFUSE_LAZY: c.inversionAboutMean(false); break;
			case ORC_OP_DIFFUSE_MATERIALIZED: c.inversionAboutMean(true); break;
			case ORC_OP_MATRIX: {
				const unsigned dim = 1u << op.k;
				QuantumGate g(dim);
				const double *m = matPool + op.state;
				for(unsigned r = 0; r < dim; ++r)
					for(unsigned col = 0; col < dim; ++col)
						g[r][col] = Amplitude{m[2 * (r * dim + col)], m[2 * (r * dim + col) + 1]};
				QubitList targets;
				for(unsigned j = 0; j < op.k; ++j) targets.push_back(matQubits[op.q0 + j]);
				c.gate(g, targets);
				break;
			}
			default: throw std::invalid_argument("unknown op kind");
		}
	}
	return c;
}

// 20 pseudo-random bits of i (a 64-bit mix); only the timing states use it.
inline uint64_t timing_jitter(uint64_t i) {
	uint64_t x = i * 0x9E3779B97F4A7C15ull;
	x ^= x >> 29;
	x *= 0xBF58476D1CE4E5B9ull;
	x ^= x >> 32;
	return x & 0xFFFFFull;
}

} // namespace

extern "C" {

int ref_unit_draws(uint32_t seed, size_t count, double *out) {
	try {
		tmfqs::Mt19937RandomSource src(seed);
		for(size_t i = 0; i < count; ++i) out[i] = src.nextUnitDouble();
		return 0;
	} catch(...) {
		return mapException();
	}
}

int ref_random_phase(unsigned n, unsigned seed, double *out) {
	try {
		const tmfqs::AmplitudesVector v = tmfqs::state::QuantumState::randomPhase(n, seed).amplitudes();
		std::memcpy(out, v.data(), v.size() * sizeof(double));
		return 0;
	} catch(...) {
		return mapException();
	}
}

// Runs ops through sim::Simulator on the dense strategy.  init == nullptr
// starts from basis(0); otherwise the amplitudes are loaded unvalidated.
// measure_seed >= 0 draws one Mt19937RandomSource(measure_seed) measurement.
int ref_run(unsigned n, const double *init, const orc_op *ops, size_t count, const double *matPool,
            const unsigned *matQubits, double *outAmps, double *totalProb, long long measureSeed,
            uint64_t *measured, double *execSeconds) {
	try {
		using namespace tmfqs;
		const circuit::Circuit c = buildCircuit(n, ops, count, matPool, matQubits);
		state::QuantumState input;
		if(init) {
			AmplitudesVector v(init, init + 2 * (static_cast<size_t>(1) << n));
			state::StateValidation val;
			val.normalization = state::NormalizationPolicy::AllowUnnormalized;
			input = state::QuantumState::fromAmplitudes(n, std::move(v), val);
		} else {
			input = state::QuantumState::basis(n);
		}
		sim::ExecutionConfig cfg;
		cfg.backend.strategy = StorageStrategyKind::Dense;
		const sim::RunResult run = sim::Simulator(cfg).run(c, input);
		if(execSeconds) *execSeconds = run.report.executionSeconds;
		if(outAmps) {
			const AmplitudesVector v = run.state.amplitudes();
			std::memcpy(outAmps, v.data(), v.size() * sizeof(double));
		}
		if(totalProb) *totalProb = run.state.totalProbability();
		if(measureSeed >= 0 && measured) {
			Mt19937RandomSource rng(static_cast<uint32_t>(measureSeed));
			*measured = run.state.measure(rng);
		}
		return 0;
	} catch(...) {
		return mapException();
	}
}

// Register-level probe used by the overlay tests: loads amps into a dense
// register, applies inversion with an explicit mean (lazy or materialized),
// then exports, measures with a fixed rnd and reports total probability.
int ref_register_inversion(unsigned n, const double *amps, int materialized, double mre, double mim,
                           int useCachedMean, double rnd, double *outAmps, uint64_t *measured, double *totalProb) {
	try {
		using namespace tmfqs;
		RegisterConfig cfg;
		cfg.strategy = StorageStrategyKind::Dense;
		QuantumRegister reg(n, AmplitudesVector(amps, amps + 2 * (static_cast<size_t>(1) << n)), cfg);
		if(useCachedMean) {
			if(materialized) reg.applyInversionAboutMeanMaterialized();
			else reg.applyInversionAboutMean();
		} else {
			if(materialized) reg.applyInversionAboutMeanMaterialized(Amplitude{mre, mim});
			else reg.applyInversionAboutMean(Amplitude{mre, mim});
		}
		const AmplitudesVector v = reg.amplitudes();
		std::memcpy(outAmps, v.data(), v.size() * sizeof(double));
		*totalProb = reg.totalProbability();
		struct Fixed final : IRandomSource {
			double v;
			double nextUnitDouble() override { return v; }
		} src;
		src.v = rnd;
		*measured = reg.measure(src);
		return 0;
	} catch(...) {
		return mapException();
	}
}

// Times the reference dense backend on ops[0..count) over a dense n-qubit
// state: real amplitudes 2^(-n/2) (1 + d_i) with distinct small d_i, so that
// no zero-skip or equal-pair skip fires, as on a random-phase input.
// Returns the wall seconds of the op loop only, and, when perOp is not null,
// each operation's own seconds.
int ref_time_ops_each(unsigned n, const orc_op *ops, size_t count, double *seconds, double *perOp);

int ref_time_ops(unsigned n, const orc_op *ops, size_t count, double *seconds) {
	return ref_time_ops_each(n, ops, count, seconds, nullptr);
}

int ref_time_ops_each(unsigned n, const orc_op *ops, size_t count, double *seconds, double *perOp) {
	try {
		using namespace tmfqs;
		RegisterConfig cfg;
		cfg.strategy = StorageStrategyKind::Dense;
		const size_t states = static_cast<size_t>(1) << n;
		AmplitudesVector v(2 * states, 0.0);
		const double a = 1.0 / std::sqrt(static_cast<double>(states));
		for(size_t i = 0; i < states; ++i) v[2 * i] = a * (1.0 + 0x1p-24 * static_cast<double>(timing_jitter(i)));
		QuantumRegister reg(n, std::move(v), cfg);
		const auto t0 = std::chrono::steady_clock::now();
		auto tprev = t0;
		for(size_t i = 0; i < count; ++i) {
			const orc_op &op = ops[i];
			switch(op.kind) {
				case ORC_OP_H: reg.applyHadamard(op.q0); break;
				case ORC_OP_X: reg.applyPauliX(op.q0); break;
				case ORC_OP_CX: reg.applyControlledNot(op.q0, op.q1); break;
				case ORC_OP_CPHASE: reg.applyControlledPhaseShift(op.q0, op.q1, op.theta); break;
				case ORC_OP_SWAP: reg.applySwap(op.q0, op.q1); break;
				case ORC_OP_PHASE_FLIP: reg.applyPhaseFlipBasisState(op.state); break;
				case ORC_OP_DIFFUSE_LAZY: reg.applyInversionAboutMean(); break;
				case ORC_OP_DIFFUSE_MATERIALIZED: reg.applyInversionAboutMeanMaterialized(); break;
				default: throw std::invalid_argument("op kind not timed");
			}
			if(perOp) {
				const auto t = std::chrono::steady_clock::now();
				perOp[i] = std::chrono::duration<double>(t - tprev).count();
				tprev = t;
			}
		}
		*seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		return 0;
	} catch(...) {
		return mapException();
	}
}

// QuantumGate factories and products of the reference, in the index order of
// tmfqs_gate_matrix / tmfqs_gate_product (include/tmfqs_sim.h).
static tmfqs::QuantumGate refGateFactory(int which, double theta, unsigned identityDim) {
	using tmfqs::QuantumGate;
	switch(which) {
		case 0: return QuantumGate::Identity(identityDim);
		case 1: return QuantumGate::Hadamard();
		case 2: return QuantumGate::PauliX();
		case 3: return QuantumGate::PauliY();
		case 4: return QuantumGate::PauliZ();
		case 5: return QuantumGate::PhaseShift(theta);
		case 6: return QuantumGate::PiOverEight();
		case 7: return QuantumGate::ControlledNot();
		case 8: return QuantumGate::Toffoli();
		case 9: return QuantumGate::ControlledPhaseShift(theta);
		case 10: return QuantumGate::Swap();
		case 11: return QuantumGate::Ising(theta);
		default: throw std::invalid_argument("unknown gate factory");
	}
}

static tmfqs::QuantumGate refGateFrom(unsigned dim, const double *m) {
	tmfqs::QuantumGate g(dim);
	for(unsigned r = 0; r < dim; ++r)
		for(unsigned c = 0; c < dim; ++c) g[r][c] = tmfqs::Amplitude{m[2 * (r * dim + c)], m[2 * (r * dim + c) + 1]};
	return g;
}

static void refGateTo(const tmfqs::QuantumGate &g, double *out) {
	const unsigned d = g.dimension();
	for(unsigned r = 0; r < d; ++r)
		for(unsigned c = 0; c < d; ++c) {
			out[2 * (r * d + c)] = g[r][c].real;
			out[2 * (r * d + c) + 1] = g[r][c].imag;
		}
}

int ref_gate_matrix(int which, double theta, unsigned identityDim, double *out, size_t cap, unsigned *dim) {
	try {
		const tmfqs::QuantumGate g = refGateFactory(which, theta, identityDim);
		if(cap < 2 * static_cast<size_t>(g.dimension()) * g.dimension()) return ORC_INVALID_ARGUMENT;
		refGateTo(g, out);
		*dim = g.dimension();
		return 0;
	} catch(...) {
		return mapException();
	}
}

int ref_gate_product(unsigned dim, const double *a, const double *b, double sre, double sim, double *out) {
	try {
		const tmfqs::QuantumGate A = refGateFrom(dim, a);
		refGateTo(b ? A * refGateFrom(dim, b) : A * tmfqs::Amplitude{sre, sim}, out);
		return 0;
	} catch(...) {
		return mapException();
	}
}

} // extern "C"
