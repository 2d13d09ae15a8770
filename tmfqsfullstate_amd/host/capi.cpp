// extern "C" entry points declared in include/tmfqs_sim.h.
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <variant>

#include "tmfqs/storage/hip/hip_state_backend.h"
#include "tmfqs_sim.h"
#include "tmfqsfs.h"

struct tmfqs_reg {
	tmfqs::QuantumRegister reg;
};

namespace {

thread_local std::string g_err;

int mapException() {
	try {
		throw;
	} catch(const std::out_of_range &e) {
		g_err = e.what();
		return TMFQS_HIP_ERR_OUT_OF_RANGE;
	} catch(const std::invalid_argument &e) {
		g_err = e.what();
		return TMFQS_HIP_ERR_INVALID_ARGUMENT;
	} catch(const std::logic_error &e) {
		g_err = e.what();
		return TMFQS_HIP_ERR_LOGIC;
	} catch(const std::bad_alloc &) {
		g_err = "out of memory";
		return TMFQS_HIP_ERR_OUT_OF_MEMORY;
	} catch(const std::runtime_error &e) {
		g_err = e.what();
		return TMFQS_HIP_ERR_RUNTIME;
	} catch(...) {
		g_err = "unknown error";
		return TMFQS_HIP_ERR_RUNTIME;
	}
}

tmfqs::RegisterConfig regConfig(const tmfqs_hip_config *c) {
	tmfqs::RegisterConfig cfg;
	cfg.strategy = tmfqs::StorageStrategyKind::Hip;
	tmfqs_hip_config d;
	tmfqs_hip_config_default(&d);
	if(!c) c = &d;
	cfg.hip.precision = c->precision == TMFQS_HIP_COMPLEX64 ? tmfqs::HipPrecision::Complex64 : tmfqs::HipPrecision::Complex128;
	cfg.hip.device = c->device;
	cfg.hip.fusion = c->fusion != 0;
	cfg.hip.tileQubits = static_cast<unsigned>(c->tile_qubits);
	cfg.hip.coalesceQubits = static_cast<unsigned>(c->coalesce_qubits);
	cfg.hip.groverBlock = static_cast<unsigned>(c->grover_block);
	cfg.hip.exactPhases = c->exact_phases != 0;
	cfg.hip.fuseQubits = static_cast<unsigned>(c->fuse_qubits);
	cfg.hip.fuseMinGates = static_cast<unsigned>(c->fuse_min_gates);
	cfg.hip.qftPlan = static_cast<unsigned>(c->qft_plan);
	return cfg;
}

tmfqs::circuit::Circuit toCircuit(unsigned n, const tmfqs_hip_op *ops, size_t count, const double *pool, const unsigned *mq) {
	using namespace tmfqs;
	circuit::Circuit c(n);
	for(size_t i = 0; i < count; ++i) {
		const tmfqs_hip_op &op = ops[i];
		switch(op.kind) {
			case TMFQS_HIP_OP_H: c.h(op.q0); break;
			case TMFQS_HIP_OP_X: c.x(op.q0); break;
			case TMFQS_HIP_OP_CX: c.cx(op.q0, op.q1); break;
			case TMFQS_HIP_OP_CPHASE: c.controlledPhase(op.q0, op.q1, op.theta); break;
			case TMFQS_HIP_OP_SWAP: c.swap(op.q0, op.q1); break;
			case TMFQS_HIP_OP_PHASE_FLIP: c.phaseFlip(op.state); break;
			case TMFQS_HIP_OP_DIFFUSE_LAZY: c.inversionAboutMean(false); break;
			case TMFQS_HIP_OP_DIFFUSE_MATERIALIZED: c.inversionAboutMean(true); break;
			case TMFQS_HIP_OP_MATRIX: {
				if(!pool || !mq) throw std::invalid_argument("matrix op without a matrix pool");
				if(op.k > 20) throw std::invalid_argument("matrix op arity out of range");
				const unsigned dim = 1u << op.k;
				QuantumGate g(dim);
				const double *m = pool + op.state;
				for(unsigned r = 0; r < dim; ++r)
					for(unsigned col = 0; col < dim; ++col) g[r][col] = Amplitude{m[2 * (r * dim + col)], m[2 * (r * dim + col) + 1]};
				QubitList t;
				for(unsigned j = 0; j < op.k; ++j) t.push_back(mq[op.q0 + j]);
				c.gate(std::move(g), std::move(t));
				break;
			}
			default: throw std::invalid_argument("unknown op kind");
		}
	}
	return c;
}

tmfqs::sim::DiagonalObservable toObservable(const tmfqs_hip_diag_term *terms, size_t nterms) {
	if(nterms && !terms) throw std::invalid_argument("diagonal observable: null term list");
	tmfqs::sim::DiagonalObservable obs;
	for(size_t i = 0; i < nterms; ++i) {
		if(terms[i].kind == TMFQS_HIP_DIAG_Z) {
			tmfqs::sim::DiagonalObservable::ZString z;
			z.coef = terms[i].coef;
			for(unsigned q = 0; q < 64; ++q)
				if((terms[i].bits >> q) & 1ull) z.qubits.push_back(q);
			obs.zStrings.push_back(z);
		} else if(terms[i].kind == TMFQS_HIP_DIAG_PROJECTOR) {
			obs.projectors.push_back({terms[i].bits, terms[i].coef});
		} else {
			throw std::invalid_argument("diagonal observable: unknown term kind");
		}
	}
	return obs;
}

size_t toOps(const tmfqs::circuit::Circuit &c, tmfqs_hip_op *out, size_t cap) {
	using namespace tmfqs;
	const auto &ops = c.operations();
	if(!out) return ops.size();
	if(cap < ops.size()) throw std::invalid_argument("op buffer too small");
	for(size_t i = 0; i < ops.size(); ++i) {
		tmfqs_hip_op r;
		std::memset(&r, 0, sizeof(r));
		std::visit(
		    [&](const auto &op) {
			    using Op = std::decay_t<decltype(op)>;
			    if constexpr(std::is_same<Op, circuit::Hadamard>::value) {
				    r.kind = TMFQS_HIP_OP_H;
				    r.q0 = op.target;
			    } else if constexpr(std::is_same<Op, circuit::PauliX>::value) {
				    r.kind = TMFQS_HIP_OP_X;
				    r.q0 = op.target;
			    } else if constexpr(std::is_same<Op, circuit::ControlledNot>::value) {
				    r.kind = TMFQS_HIP_OP_CX;
				    r.q0 = op.control;
				    r.q1 = op.target;
			    } else if constexpr(std::is_same<Op, circuit::ControlledPhaseShift>::value) {
				    r.kind = TMFQS_HIP_OP_CPHASE;
				    r.q0 = op.control;
				    r.q1 = op.target;
				    r.theta = op.theta;
			    } else if constexpr(std::is_same<Op, circuit::Swap>::value) {
				    r.kind = TMFQS_HIP_OP_SWAP;
				    r.q0 = op.first;
				    r.q1 = op.second;
			    } else if constexpr(std::is_same<Op, circuit::PhaseFlipBasisState>::value) {
				    r.kind = TMFQS_HIP_OP_PHASE_FLIP;
				    r.state = op.state;
			    } else if constexpr(std::is_same<Op, circuit::InversionAboutMean>::value) {
				    r.kind = op.materialized ? TMFQS_HIP_OP_DIFFUSE_MATERIALIZED : TMFQS_HIP_OP_DIFFUSE_LAZY;
			    } else {
				    throw std::invalid_argument("matrix gates have no flat op record without a pool");
			    }
		    },
		    ops[i]);
		out[i] = r;
	}
	return ops.size();
}

class FixedRandomSource final : public tmfqs::IRandomSource {
	public:
		explicit FixedRandomSource(double v) : v_(v) {}
		double nextUnitDouble() override { return v_; }

	private:
		double v_;
};

}  // namespace

#define CAPI_BEGIN try {
#define CAPI_END             \
	return TMFQS_HIP_OK;     \
	}                        \
	catch(...) {             \
		return mapException(); \
	}

extern "C" {

const char *tmfqs_sim_last_error(void) { return g_err.c_str(); }

int tmfqs_reg_create_basis(unsigned n, uint64_t state, const tmfqs_hip_config *cfg, tmfqs_reg **out) {
	CAPI_BEGIN
	*out = new tmfqs_reg{tmfqs::QuantumRegister(n, state, regConfig(cfg))};
	CAPI_END
}

int tmfqs_reg_create_amplitudes(unsigned n, const double *amps, const tmfqs_hip_config *cfg, tmfqs_reg **out) {
	CAPI_BEGIN
	const size_t elems = tmfqs::checkedAmplitudeElementCount(n);
	*out = new tmfqs_reg{tmfqs::QuantumRegister(n, tmfqs::AmplitudesVector(amps, amps + elems), regConfig(cfg))};
	CAPI_END
}

int tmfqs_reg_create_random_phase(unsigned n, unsigned seed, const tmfqs_hip_config *cfg, tmfqs_reg **out) {
	CAPI_BEGIN
	// What QuantumState::randomPhase(n, seed) materializes to on a Hip
	// register, generated in device memory.
	tmfqs::QuantumRegister reg(n, regConfig(cfg));
	tmfqs_hip_state *h = tmfqs::hipBackendHandle(reg.backend());
	if(!h) throw std::logic_error("tmfqs_reg_create_random_phase: the register has no device backend");
	tmfqs::raiseHipStatus(tmfqs_hip_init_random_phase(h, n, seed));
	*out = new tmfqs_reg{std::move(reg)};
	CAPI_END
}

int tmfqs_reg_create_chunked(unsigned n, uint64_t state, const tmfqs_hip_config *cfg, uint64_t chunk_states,
                             uint64_t window_chunks, tmfqs_reg **out) {
	CAPI_BEGIN
	tmfqs::RegisterConfig rc = regConfig(cfg);
	rc.strategy = tmfqs::StorageStrategyKind::HipChunked;
	rc.hip.chunkStates = static_cast<size_t>(chunk_states);
	rc.hip.windowChunks = static_cast<size_t>(window_chunks);
	*out = new tmfqs_reg{tmfqs::QuantumRegister(n, state, rc)};
	CAPI_END
}

int tmfqs_reg_clone(tmfqs_reg *r, tmfqs_reg **out) {
	CAPI_BEGIN
	if(!r) throw std::invalid_argument("null register");
	*out = new tmfqs_reg{tmfqs::QuantumRegister(r->reg)};
	CAPI_END
}

int tmfqs_reg_free(tmfqs_reg *r) {
	CAPI_BEGIN
	delete r;
	CAPI_END
}

int tmfqs_reg_run(tmfqs_reg *r, const tmfqs_hip_op *ops, size_t count, const double *pool, const unsigned *mq, int trace, double *seconds) {
	CAPI_BEGIN
	if(!r) throw std::invalid_argument("null register");
	const tmfqs::circuit::Circuit c = toCircuit(r->reg.qubitCount(), ops, count, pool, mq);
	tmfqs::sim::ExecutionConfig ec;
	ec.observability.traceOperations = trace != 0;
	const tmfqs::experiment::RunReport rep = tmfqs::sim::Simulator(ec).runOn(c, r->reg);
	if(seconds) *seconds = rep.executionSeconds;
	CAPI_END
}

int tmfqs_reg_run_traced(tmfqs_reg *r, const tmfqs_hip_op *ops, size_t count, const double *pool, const unsigned *mq,
                         double *op_seconds, double *seconds) {
	CAPI_BEGIN
	if(!r || (count && !op_seconds)) throw std::invalid_argument("null register or output");
	const tmfqs::circuit::Circuit c = toCircuit(r->reg.qubitCount(), ops, count, pool, mq);
	tmfqs::sim::ExecutionConfig ec;
	ec.observability.traceOperations = true;
	const tmfqs::experiment::RunReport rep = tmfqs::sim::Simulator(ec).runOn(c, r->reg);
	for(size_t i = 0; i < count; ++i) op_seconds[i] = i < rep.operations.size() ? rep.operations[i].seconds : 0.0;
	if(seconds) *seconds = rep.executionSeconds;
	CAPI_END
}

int tmfqs_sim_run(unsigned n, uint64_t basis, const double *init, const tmfqs_hip_config *cfg, const tmfqs_hip_op *ops, size_t count,
                  const double *pool, const unsigned *mq, double *seconds, tmfqs_reg **result) {
	CAPI_BEGIN
	using namespace tmfqs;
	state::QuantumState input;
	if(init) {
		state::StateValidation v;
		v.normalization = state::NormalizationPolicy::AllowUnnormalized;
		input = state::QuantumState::fromAmplitudes(n, AmplitudesVector(init, init + checkedAmplitudeElementCount(n)), v);
	} else {
		input = state::QuantumState::basis(n, basis);
	}
	sim::ExecutionConfig ec;
	ec.backend = regConfig(cfg);
	sim::RunResult rr = sim::Simulator(ec).run(toCircuit(n, ops, count, pool, mq), input);
	if(seconds) *seconds = rr.report.executionSeconds;
	*result = new tmfqs_reg{*rr.state.registerPtr()};
	CAPI_END
}

namespace {
tmfqs::state::QuantumState hostState(unsigned n, const double *amps) {
	tmfqs::state::StateValidation v;
	v.normalization = tmfqs::state::NormalizationPolicy::AllowUnnormalized;
	return tmfqs::state::QuantumState::fromAmplitudes(n, tmfqs::AmplitudesVector(amps, amps + tmfqs::checkedAmplitudeElementCount(n)), v);
}

void fillComparison(const tmfqs::experiment::StateComparison &m, tmfqs_hip_comparison *out) {
	*out = tmfqs_hip_comparison{};
	out->bitwise_equal = m.bitwiseEqual ? 1 : 0;
	out->max_abs_amplitude_error = m.maxAbsAmplitudeError;
	out->max_abs_component_error = m.maxAbsComponentError;
	out->rmse_amplitude = m.rmseAmplitude;
	out->rel_l2 = m.relL2;
	out->max_abs_probability_error = m.maxAbsProbabilityError;
	out->total_probability_diff = m.totalProbabilityDiff;
	out->worst_state = m.worstState;
}
}  // namespace

int tmfqs_sim_compare(unsigned n, uint64_t basis, const double *init, const tmfqs_hip_config *ref_cfg,
                      const tmfqs_hip_config *cand_cfg, const tmfqs_hip_op *ops, size_t count, const double *pool,
                      const unsigned *mq, tmfqs_hip_comparison *out) {
	CAPI_BEGIN
	using namespace tmfqs;
	if(!out) throw std::invalid_argument("tmfqs_sim_compare: null output");
	const state::QuantumState input = init ? hostState(n, init) : state::QuantumState::basis(n, basis);
	const circuit::Circuit c = toCircuit(n, ops, count, pool, mq);
	sim::ExecutionConfig er, ec;
	er.backend = regConfig(ref_cfg);
	ec.backend = regConfig(cand_cfg);
	const sim::RunResult a = sim::Simulator(er).run(c, input);
	const sim::RunResult b = sim::Simulator(ec).run(c, input);
	fillComparison(experiment::compareStates(a.state, b.state), out);
	CAPI_END
}

int tmfqs_compare_amplitudes(unsigned n, const double *ref, const double *cand, tmfqs_hip_comparison *out) {
	CAPI_BEGIN
	if(!ref || !cand || !out) throw std::invalid_argument("tmfqs_compare_amplitudes: null pointer");
	fillComparison(tmfqs::experiment::compareStates(hostState(n, ref), hostState(n, cand)), out);
	CAPI_END
}

namespace {
tmfqs::QuantumGate gateFactory(int which, double theta, unsigned identityDim) {
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

tmfqs::QuantumGate gateFrom(unsigned dim, const double *m) {
	tmfqs::QuantumGate g(dim);
	for(unsigned r = 0; r < dim; ++r)
		for(unsigned c = 0; c < dim; ++c) g[r][c] = tmfqs::Amplitude{m[2 * (r * dim + c)], m[2 * (r * dim + c) + 1]};
	return g;
}

void gateTo(const tmfqs::QuantumGate &g, double *out) {
	const unsigned d = g.dimension();
	for(unsigned r = 0; r < d; ++r)
		for(unsigned c = 0; c < d; ++c) {
			out[2 * (r * d + c)] = g[r][c].real;
			out[2 * (r * d + c) + 1] = g[r][c].imag;
		}
}
}  // namespace

int tmfqs_gate_matrix(int which, double theta, unsigned identityDim, double *out, size_t cap, unsigned *dim) {
	CAPI_BEGIN
	const tmfqs::QuantumGate g = gateFactory(which, theta, identityDim);
	const size_t need = 2 * static_cast<size_t>(g.dimension()) * g.dimension();
	if(!out || !dim || cap < need) throw std::invalid_argument("tmfqs_gate_matrix: output too small");
	gateTo(g, out);
	*dim = g.dimension();
	CAPI_END
}

int tmfqs_gate_product(unsigned dim, const double *a, const double *b, double sre, double sim, double *out) {
	CAPI_BEGIN
	if(!a || !out) throw std::invalid_argument("tmfqs_gate_product: null matrix");
	const tmfqs::QuantumGate A = gateFrom(dim, a);
	gateTo(b ? A * gateFrom(dim, b) : A * tmfqs::Amplitude{sre, sim}, out);
	CAPI_END
}

int tmfqs_reg_qubit_count(tmfqs_reg *r, unsigned *out) {
	CAPI_BEGIN
	*out = r->reg.qubitCount();
	CAPI_END
}

int tmfqs_reg_export(tmfqs_reg *r, double *out) {
	CAPI_BEGIN
	const tmfqs::AmplitudesVector a = r->reg.amplitudes();
	std::memcpy(out, a.data(), a.size() * sizeof(double));
	CAPI_END
}

int tmfqs_reg_amplitude(tmfqs_reg *r, uint64_t s, double *re, double *im) {
	CAPI_BEGIN
	const tmfqs::Amplitude a = r->reg.amplitude(s);
	*re = a.real;
	*im = a.imag;
	CAPI_END
}

int tmfqs_reg_total_probability(tmfqs_reg *r, double *out) {
	CAPI_BEGIN
	*out = r->reg.totalProbability();
	CAPI_END
}

int tmfqs_reg_measure_seed(tmfqs_reg *r, uint32_t seed, uint64_t *out) {
	CAPI_BEGIN
	tmfqs::Mt19937RandomSource src(seed);
	*out = r->reg.measure(src);
	CAPI_END
}

int tmfqs_reg_measure_fixed(tmfqs_reg *r, double rnd, uint64_t *out) {
	CAPI_BEGIN
	FixedRandomSource src(rnd);
	*out = r->reg.measure(src);
	CAPI_END
}

int tmfqs_reg_hip_handle(tmfqs_reg *r, tmfqs_hip_state **out) {
	CAPI_BEGIN
	*out = tmfqs::hipBackendHandle(r->reg.backend());
	if(!*out) throw std::logic_error("register has no device backend");
	CAPI_END
}

int tmfqs_state_random_phase(unsigned n, unsigned seed, double *out) {
	CAPI_BEGIN
	const tmfqs::AmplitudesVector a = tmfqs::state::QuantumState::randomPhase(n, seed).amplitudes();
	std::memcpy(out, a.data(), a.size() * sizeof(double));
	CAPI_END
}

int tmfqs_circuit_qft(unsigned n, tmfqs_hip_op *out, size_t cap, size_t *count) {
	CAPI_BEGIN
	*count = toOps(tmfqs::circuit::makeQft(n), out, cap);
	CAPI_END
}

int tmfqs_circuit_grover(unsigned n, const uint64_t *marked, size_t m, int materialized, tmfqs_hip_op *out, size_t cap, size_t *count) {
	CAPI_BEGIN
	tmfqs::circuit::GroverCircuitOptions o;
	o.markedStates = tmfqs::BasisStateList(std::vector<tmfqs::StateIndex>(marked, marked + m));
	o.materializedDiffusion = materialized != 0;
	*count = toOps(tmfqs::circuit::makeGrover(n, o), out, cap);
	CAPI_END
}

int tmfqs_circuit_ring(unsigned n, unsigned layers, unsigned seed, tmfqs_hip_op *out, size_t cap, size_t *count) {
	CAPI_BEGIN
	*count = toOps(tmfqs::circuit::makeRingLayers(n, layers, seed), out, cap);
	CAPI_END
}

int tmfqs_reg_adjoint(tmfqs_reg *r, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *mat_qubits,
                      const tmfqs_hip_diag_term *terms, size_t nterms, unsigned checkpoint_gates, double *expval,
                      double *grads, size_t grad_count, double *seconds) {
	CAPI_BEGIN
	if(!r || !expval) throw std::invalid_argument("tmfqs_reg_adjoint: null argument");
	const tmfqs::sim::DiagonalObservable obs = toObservable(terms, nterms);
	const tmfqs::circuit::Circuit c = toCircuit(r->reg.qubitCount(), ops, count, mat_pool, mat_qubits);
	const tmfqs::sim::AdjointResult res = tmfqs::sim::adjointGradient(c, r->reg, obs, checkpoint_gates);
	if(res.gradients.size() != grad_count)
		throw std::invalid_argument("tmfqs_reg_adjoint: grad_count must equal the number of CPHASE ops (" +
		                            std::to_string(res.gradients.size()) + ")");
	for(size_t i = 0; i < grad_count; ++i) grads[i] = res.gradients[i];
	*expval = res.expectation;
	if(seconds) *seconds = res.executionSeconds;
	CAPI_END
}

int tmfqs_reg_expectation(tmfqs_reg *r, const tmfqs_hip_diag_term *terms, size_t nterms, double *out) {
	CAPI_BEGIN
	if(!r || !out) throw std::invalid_argument("tmfqs_reg_expectation: null argument");
	const tmfqs::sim::DiagonalObservable obs = toObservable(terms, nterms);
	*out = tmfqs::sim::expectation(r->reg, obs);
	CAPI_END
}

}  // extern "C"
