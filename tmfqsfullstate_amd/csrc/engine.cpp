// Host side of the gfx950 dense state-vector engine: handle, operation queue,
// pass planner, exact sequential-chain sampling, and the extern "C" ABI
// declared in include/tmfqs_hip.h.
//
// Compiled with -ffp-contract=off: the host replays elementwise steps and
// sequential probability chains with exactly the reference's operation order.
#include "tmfqs_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "engine.h"
#include "mt_jump.h"
#include "phase_math.h"

using namespace tmfqs_hip;

namespace {

constexpr double kPi = 3.14159265358979323846;  // include/tmfqs/core/constants.h:6 in the reference

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
	using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what) {
	if(e != hipSuccess) {
		if(e == hipErrorOutOfMemory) throw std::bad_alloc();
		throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
	}
}

// Fault injection for tests of the error paths: TMFQS_HIP_FAULT=<kind>@<k>
// with kind alloc (device and pinned allocations), launch (kernel launches)
// or sync (stream, event and device synchronizations) makes the k-th such
// call after each tmfqs_hip_create / tmfqs_hip_batch_create skip the HIP call
// and report hipErrorOutOfMemory, hipErrorLaunchFailure or hipErrorUnknown.
// Nothing reaches the device, so the state stays consistent; the handle is
// expected to be destroyed afterwards.  Not thread-safe (a test hook).
struct FaultPlan {
	int kind = 0;  // 0 off, 1 alloc, 2 launch, 3 sync
	long target = 0, count = 0;
};
FaultPlan g_fault;

void fault_arm() {
	g_fault = FaultPlan();
	const char *e = getenv("TMFQS_HIP_FAULT");
	if(!e || !*e) return;
	const char *at = strchr(e, '@');
	const std::string kind(e, at ? (size_t)(at - e) : strlen(e));
	const long k = at ? strtol(at + 1, nullptr, 10) : 1;
	if(k < 1) return;
	g_fault.kind = kind == "alloc" ? 1 : kind == "launch" ? 2 : kind == "sync" ? 3 : 0;
	g_fault.target = k;
}

int fault_class(const char *what) {
	if(strstr(what, "Malloc")) return 1;
	if(strncmp(what, "launch_", 7) == 0 || strstr(what, "GraphLaunch")) return 2;
	if(strstr(what, "Synchronize")) return 3;
	return 0;
}

bool fault_hit(const char *what) {
	if(!g_fault.kind || fault_class(what) != g_fault.kind) return false;
	return ++g_fault.count == g_fault.target;
}

hipError_t fault_error() {
	return g_fault.kind == 1 ? hipErrorOutOfMemory : g_fault.kind == 2 ? hipErrorLaunchFailure : hipErrorUnknown;
}

#define HIPC(expr) hip_check((g_fault.kind && fault_hit(#expr)) ? fault_error() : (expr), #expr)

int map_exception() {
	try {
		throw;
	} catch(const std::out_of_range &e) {
		g_last_error = e.what();
		return TMFQS_HIP_ERR_OUT_OF_RANGE;
	} catch(const std::invalid_argument &e) {
		g_last_error = e.what();
		return TMFQS_HIP_ERR_INVALID_ARGUMENT;
	} catch(const std::logic_error &e) {
		g_last_error = e.what();
		return TMFQS_HIP_ERR_LOGIC;
	} catch(const DeviceError &e) {
		g_last_error = e.what();
		return TMFQS_HIP_ERR_DEVICE;
	} catch(const std::bad_alloc &) {
		g_last_error = "out of memory";
		return TMFQS_HIP_ERR_OUT_OF_MEMORY;
	} catch(const std::runtime_error &e) {
		g_last_error = e.what();
		return TMFQS_HIP_ERR_RUNTIME;
	} catch(...) {
		g_last_error = "unknown error";
		return TMFQS_HIP_ERR_RUNTIME;
	}
}

// Queued operation; bit arguments are physical bit positions (p = n-1-q).
enum QKind { Q_H, Q_X, Q_CX, Q_CP, Q_SWAP, Q_U1, Q_UK, Q_FLIP, Q_SET, Q_DIFF, Q_AFFINE, Q_SCATTER,
             Q_PH,     // phase (cr, ci) on states with bit b0 set
             Q_GRAD };  // adjoint only: gradient reduction over states with bit b0 (and b1 when k == 2) set; slot = state

struct QOp {
	int kind;
	unsigned b0 = 0, b1 = 0;
	unsigned k = 0;
	unsigned bits[5] = {0, 0, 0, 0, 0};
	size_t mat = 0;  // offset into qmats
	double cr = 0, ci = 0, theta = 0;
	u64 state = 0;
	double ar = 0, ai = 0, br = 0, bi = 0;
	size_t sc0 = 0, scn = 0;  // Q_SCATTER: range in qscatter
};

struct Cplx {
	double re, im;
};

// Upload arena: pinned host and device buffers used as a bump allocator; a
// wrap-around synchronises the stream before reusing the space.
struct Arena {
	char *host = nullptr;
	char *dev = nullptr;
	size_t cap = 0, off = 0;
};

struct ProfAgg {
	u64 launches = 0;
	double ms = 0;
	double bytes = 0;
};

struct ProfRec {
	std::string name;
	double bytes;
	hipEvent_t start, stop;
};

}  // namespace

struct tmfqs_hip_state {
	tmfqs_hip_config cfg{};
	int prec = 0;
	size_t esz = 16;
	unsigned n = 0;
	u64 N = 0;
	bool init = false;
	void *amp = nullptr;
	size_t amp_bytes = 0;
	void *scratch = nullptr;  // lazily allocated second state buffer
	size_t scratch_bytes = 0;
	hipStream_t stream = nullptr;
	double *d_red = nullptr;    // reduction partials + outputs
	double *d_phase = nullptr;  // 5 x 256 complex phase tables
	double *d_cin = nullptr;    // e^{i pi v / 2^j} at 2^j - 1 + v, j = 0..5 (QFT in-block constants), FP64 then FP32
	double *d_sums = nullptr;   // sampling chunk sums
	ChunkTicks *d_ticks = nullptr;
	int32_t *d_ulog = nullptr;
	size_t chunk_cap = 0;
	Arena arena;
	char *h_stage = nullptr;  // pinned staging for bulk copies
	size_t stage_bytes = 0;
	std::vector<QOp> queue;
	std::vector<double> qmats;
	std::vector<Scatter> qscatter;
	size_t ew_steps_pending = 0;
	std::unordered_map<u64, Cplx> tracked;
	std::vector<u64> dirty;  // states whose tracked value must be scattered at the end of the run
	std::unordered_set<u64> dirty_set;
	bool profiling = false;
	std::string trace_path;  // TMFQS_HIP_TRACE: where the handle's kernel report goes when it is destroyed
	std::vector<ProfRec> recs;
	std::vector<hipEvent_t> event_pool;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> timed;  // tmfqs_hip_flush_timed brackets
	std::map<std::string, ProfAgg> prof;
	// adjoint: checkpoint/lambda buffers (kept between calls), gradient partials and outputs
	std::vector<void *> adj_bufs;
	size_t adj_buf_bytes = 0;
	double *d_gpart = nullptr;
	double *d_grads = nullptr;
	size_t grads_cap = 0;
	Launch L() const { return Launch{stream, prec}; }
};

namespace {

// ------------------------------------------------------------ utilities
void require_init(const tmfqs_hip_state *h, const char *op) {
	if(!h->init) throw std::logic_error(std::string("HipStateBackend is not initialized for ") + op);
}

void validate_state(const tmfqs_hip_state *h, u64 s, const char *scope) {
	if(s >= h->N) throw std::out_of_range(std::string(scope) + " state index out of range");
}

void validate_qubit(const tmfqs_hip_state *h, unsigned q, const char *scope) {
	if(q >= h->n) throw std::out_of_range(std::string(scope) + " qubit index out of range");
}

void validate_two(const tmfqs_hip_state *h, unsigned a, unsigned b, const char *scope) {
	validate_qubit(h, a, scope);
	validate_qubit(h, b, scope);
	if(a == b) throw std::invalid_argument(std::string(scope) + " requires distinct qubits");
}

unsigned pbit(const tmfqs_hip_state *h, unsigned q) { return h->n - 1u - q; }

hipEvent_t take_event(tmfqs_hip_state *h) {
	if(!h->event_pool.empty()) {
		hipEvent_t e = h->event_pool.back();
		h->event_pool.pop_back();
		return e;
	}
	hipEvent_t e;
	HIPC(hipEventCreate(&e));
	return e;
}

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
	tmfqs_hip_state *h;
	ProfRec rec;
	bool on;
	ProfScope(tmfqs_hip_state *hh, const char *name, double bytes) : h(hh), on(hh->profiling) {
		if(on) {
			rec.name = name;
			rec.bytes = bytes;
			rec.start = take_event(h);
			rec.stop = take_event(h);
			HIPC(hipEventRecord(rec.start, h->stream));
		}
	}
	~ProfScope() {
		if(on) {
			if(hipEventRecord(rec.stop, h->stream) == hipSuccess) h->recs.push_back(rec);
		}
	}
};

void collect_profile(tmfqs_hip_state *h) {
	if(h->recs.empty()) return;
	HIPC(hipStreamSynchronize(h->stream));
	for(auto &r : h->recs) {
		float ms = 0;
		HIPC(hipEventElapsedTime(&ms, r.start, r.stop));
		auto &a = h->prof[r.name];
		a.launches += 1;
		a.ms += ms;
		a.bytes += r.bytes;
		h->event_pool.push_back(r.start);
		h->event_pool.push_back(r.stop);
	}
	h->recs.clear();
}

void arena_ensure(tmfqs_hip_state *h, size_t need) {
	Arena &A = h->arena;
	if(need > A.cap) {
		HIPC(hipStreamSynchronize(h->stream));
		if(A.host) hipHostFree(A.host);
		if(A.dev) hipFree(A.dev);
		A.host = nullptr;
		A.dev = nullptr;
		A.cap = 0;  // stays empty if an allocation below fails
		A.off = 0;
		size_t cap = std::max<size_t>(need * 2, 8u << 20);
		HIPC(hipHostMalloc((void **)&A.host, cap, hipHostMallocDefault));
		HIPC(hipMalloc((void **)&A.dev, cap));
		A.cap = cap;
		A.off = 0;
	}
}

// Copies bytes into the upload arena and enqueues the H2D copy; returns the
// device pointer valid for kernels enqueued after this call.
void *arena_upload(tmfqs_hip_state *h, const void *src, size_t bytes) {
	if(bytes == 0) return nullptr;
	const size_t need = (bytes + 255) & ~size_t(255);
	arena_ensure(h, need);
	Arena &A = h->arena;
	if(A.off + need > A.cap) {
		HIPC(hipStreamSynchronize(h->stream));
		A.off = 0;
	}
	std::memcpy(A.host + A.off, src, bytes);
	void *d = A.dev + A.off;
	HIPC(hipMemcpyAsync(d, A.host + A.off, bytes, hipMemcpyHostToDevice, h->stream));
	A.off += need;
	return d;
}

void ensure_stage(tmfqs_hip_state *h, size_t bytes) {
	if(h->stage_bytes >= bytes) return;
	if(h->h_stage) {
		HIPC(hipStreamSynchronize(h->stream));
		hipHostFree(h->h_stage);
		h->h_stage = nullptr;
		h->stage_bytes = 0;
	}
	HIPC(hipHostMalloc((void **)&h->h_stage, bytes, hipHostMallocDefault));
	h->stage_bytes = bytes;
}

// A second state buffer for out-of-place passes, allocated when device
// memory keeps at least 1 GiB free beside it; false (nothing allocated) when
// there is no room.
bool reserve_scratch(tmfqs_hip_state *h) {
	if(h->scratch && h->scratch_bytes >= h->amp_bytes) return true;
	size_t freeb = 0, total = 0;
	if(hipMemGetInfo(&freeb, &total) != hipSuccess) {
		(void)hipGetLastError();
		return false;
	}
	const size_t reserve = (size_t)1 << 30;
	if(h->scratch) {
		freeb += h->scratch_bytes;
		hipFree(h->scratch);
		h->scratch = nullptr;
		h->scratch_bytes = 0;
	}
	if(freeb < h->amp_bytes + reserve) return false;
	if(hipMalloc(&h->scratch, h->amp_bytes) != hipSuccess) {
		(void)hipGetLastError();
		h->scratch = nullptr;
		return false;
	}
	h->scratch_bytes = h->amp_bytes;
	return true;
}

void ensure_scratch(tmfqs_hip_state *h) {
	if(h->scratch_bytes >= h->amp_bytes) return;
	if(h->scratch) hipFree(h->scratch);
	h->scratch = nullptr;
	h->scratch_bytes = 0;
	HIPC(hipMalloc(&h->scratch, h->amp_bytes));
	h->scratch_bytes = h->amp_bytes;
}

void ensure_chunk_buffers(tmfqs_hip_state *h, size_t nch) {
	if(h->chunk_cap >= nch) return;
	hipFree(h->d_sums);
	hipFree(h->d_ticks);
	hipFree(h->d_ulog);
	h->d_sums = nullptr;
	h->d_ticks = nullptr;
	h->d_ulog = nullptr;
	h->chunk_cap = 0;  // stays empty if an allocation below fails
	HIPC(hipMalloc((void **)&h->d_sums, nch * sizeof(double)));
	HIPC(hipMalloc((void **)&h->d_ticks, nch * sizeof(ChunkTicks)));
	HIPC(hipMalloc((void **)&h->d_ulog, nch * sizeof(int32_t)));
	h->chunk_cap = nch;
}

void build_phase_tables(tmfqs_hip_state *h) {
	std::vector<double> t(5 * 256 * 2);
	for(int j = 0; j < 5; ++j) {
		const double denom = std::ldexp(1.0, 8 * (j + 1));
		for(int v = 0; v < 256; ++v) {
			const double ang = 2.0 * kPi * (double)v / denom;
			t[2 * (j * 256 + v)] = std::cos(ang);
			t[2 * (j * 256 + v) + 1] = std::sin(ang);
		}
	}
	HIPC(hipMalloc((void **)&h->d_phase, t.size() * sizeof(double)));
	HIPC(hipMemcpy(h->d_phase, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
	std::vector<double> c;
	for(int j = 0; j < 6; ++j)
		for(int v = 0; v < (1 << j); ++v) {
			double s, co;
			__builtin_sincos(kPi * (double)v / (double)(1 << j), &s, &co);
			c.push_back(co);
			c.push_back(s);
		}
	// The same 63 constants rounded to FP32 follow the FP64 ones (complex64
	// kernels use them as uniform operands without converting per use).
	std::vector<float> cf(c.begin(), c.end());
	HIPC(hipMalloc((void **)&h->d_cin, c.size() * sizeof(double) + cf.size() * sizeof(float)));
	HIPC(hipMemcpy(h->d_cin, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
	HIPC(hipMemcpy(h->d_cin + c.size(), cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice));
}

double pass_bytes(const tmfqs_hip_state *h) { return 2.0 * (double)h->N * (double)h->esz; }

// ------------------------------------------------------------ allocation
void allocate_state(tmfqs_hip_state *h, unsigned n) {
	if(n > 63u) throw std::invalid_argument("Number of qubits exceeds supported range for state indexing");
	const u64 N = 1ull << n;
	const size_t bytes = (size_t)N * h->esz;
	h->queue.clear();
	h->qmats.clear();
	h->qscatter.clear();
	h->ew_steps_pending = 0;
	h->tracked.clear();
	h->dirty.clear();
	h->dirty_set.clear();
	if(h->amp_bytes != bytes) {
		HIPC(hipStreamSynchronize(h->stream));
		if(h->amp) hipFree(h->amp);
		h->amp = nullptr;
		h->amp_bytes = 0;
		h->init = false;
		HIPC(hipMalloc(&h->amp, bytes));
		h->amp_bytes = bytes;
		if(h->scratch) {
			hipFree(h->scratch);
			h->scratch = nullptr;
			h->scratch_bytes = 0;
		}
	}
	h->n = n;
	h->N = N;
	h->init = true;
}

// ------------------------------------------------------------ execution
void exec_op_unfused(tmfqs_hip_state *h, const QOp &op);
// Where exec_gates applies a gate run: the handle's state by default, or two
// states stepped through the same ops with gradient reductions (adjoint).
struct GateTarget {
	void *a = nullptr;
	void *b = nullptr;         // second state (dual passes), or null
	int tile_delta = 0;        // tile bits relative to the handle's default
	double *gpart = nullptr;   // dual: kMaxGradSlots * tile_dual_grid partials
	double *dgrads = nullptr;  // dual: gradient outputs (device), indexed by Q_GRAD's slot
	double grad_scale = 1.0;
};
void exec_gates(tmfqs_hip_state *h, const std::vector<QOp> &ops, size_t lo, size_t hi, const GateTarget *tgt = nullptr);
void exec_elementwise(tmfqs_hip_state *h, const std::vector<QOp> &ops, size_t lo, size_t hi);

bool is_elementwise(int kind) {
	return kind == Q_FLIP || kind == Q_SET || kind == Q_DIFF || kind == Q_AFFINE || kind == Q_SCATTER;
}

// Closes the current elementwise run: the exact values of every state flipped
// or set in the run become a scatter applied after the run's sweeps.
void close_ew_run(tmfqs_hip_state *h) {
	if(h->dirty.empty()) return;
	QOp sc;
	sc.kind = Q_SCATTER;
	sc.sc0 = h->qscatter.size();
	for(u64 s : h->dirty) {
		const Cplx v = h->tracked.at(s);
		h->qscatter.push_back(Scatter{s, v.re, v.im});
	}
	sc.scn = h->dirty.size();
	h->queue.push_back(sc);
	h->dirty.clear();
	h->dirty_set.clear();
}

void flush(tmfqs_hip_state *h) {
	if(h->queue.empty()) return;
	close_ew_run(h);
	std::vector<QOp> ops;
	ops.swap(h->queue);
	size_t i = 0;
	while(i < ops.size()) {
		size_t j = i;
		const bool ew = is_elementwise(ops[i].kind);
		while(j < ops.size() && is_elementwise(ops[j].kind) == ew) ++j;
		if(ew)
			exec_elementwise(h, ops, i, j);
		else
			exec_gates(h, ops, i, j);
		i = j;
	}
	h->qmats.clear();
	h->qscatter.clear();
	h->ew_steps_pending = 0;
	if(h->prec == 1) {
		// complex64 storage rounds every tracked value once it is flushed.
		for(auto &kv : h->tracked) {
			kv.second.re = (double)(float)kv.second.re;
			kv.second.im = (double)(float)kv.second.im;
		}
	}
}

void sync_all(tmfqs_hip_state *h) {
	flush(h);
	HIPC(hipStreamSynchronize(h->stream));
}

// Default most elementwise steps per flush: complex128 replays each step's
// rounding (bit-identical to the dense backend), so a sweep carries 1024
// steps (one FP64 add per amplitude and step for real states, two otherwise:
// past ~100 steps a sweep is bound by FP64 issue, not HBM, and longer sweeps
// only save passes); complex64 composes a whole run into one affine map.
unsigned ew_block(const tmfqs_hip_state *h) {
	if(h->cfg.grover_block > 0) return (unsigned)h->cfg.grover_block;
	return h->prec == 0 ? 1024u : 65536u;
}

void exec_elementwise(tmfqs_hip_state *h, const std::vector<QOp> &ops, size_t lo, size_t hi) {
	std::vector<EwStep> steps;
	const unsigned block = ew_block(h);
	auto run_steps = [&]() {
		if(h->prec == 1 && !steps.empty()) {
			// complex64: a -> 2m - a and a -> A a + B are affine, so the run
			// collapses (in FP64) to a -> s a + b and stored amplitudes are
			// rounded once per sweep instead of once per step
			double sr = 1, si = 0, br = 0, bi = 0;
			for(const EwStep &st : steps) {
				if(st.kind == EW_DIFFUSE) {
					sr = -sr;
					si = -si;
					br = st.ar - br;
					bi = st.ai - bi;
				} else {
					const double nsr = st.ar * sr - st.ai * si, nsi = st.ar * si + st.ai * sr;
					const double nbr = st.ar * br - st.ai * bi + st.br, nbi = st.ar * bi + st.ai * br + st.bi;
					sr = nsr;
					si = nsi;
					br = nbr;
					bi = nbi;
				}
			}
			steps.assign(1, EwStep{EW_AFFINE, 0, sr, si, br, bi});
		}
		size_t s0 = 0;
		while(s0 < steps.size()) {
			const size_t cnt = std::min<size_t>(block, steps.size() - s0);
			bool pure = true;
			for(size_t i = s0; i < s0 + cnt; ++i) pure = pure && steps[i].kind == EW_DIFFUSE;
			if(pure) {
				std::vector<double> m2(2 * cnt);
				// the chain applied to an imaginary part +0 and -0, valid when every Im(2m) is +-0
				bool zero_ok = true;
				double zp = 0.0, zm = -0.0;
				for(size_t i = 0; i < cnt; ++i) {
					m2[2 * i] = steps[s0 + i].ar;
					m2[2 * i + 1] = steps[s0 + i].ai;
					zero_ok = zero_ok && steps[s0 + i].ai == 0.0;
					zp = steps[s0 + i].ai - zp;
					zm = steps[s0 + i].ai - zm;
				}
				const double *d = (const double *)arena_upload(h, m2.data(), m2.size() * sizeof(double));
				ProfScope ps(h, "k_diffuse_chain", pass_bytes(h));
				HIPC(launch_diffuse_chain(h->L(), h->amp, h->n, d, (unsigned)cnt, zp, zm, zero_ok ? 1 : 0));
			} else {
				const EwStep *d = (const EwStep *)arena_upload(h, steps.data() + s0, cnt * sizeof(EwStep));
				ProfScope ps(h, "k_ew_chain", pass_bytes(h));
				HIPC(launch_ew_chain(h->L(), h->amp, h->n, d, (unsigned)cnt));
			}
			s0 += cnt;
		}
		steps.clear();
	};
	for(size_t i = lo; i < hi; ++i) {
		const QOp &op = ops[i];
		if(op.kind == Q_DIFF) {
			EwStep s{EW_DIFFUSE, 0, 2.0 * op.ar, 2.0 * op.ai, 0, 0};
			steps.push_back(s);
		} else if(op.kind == Q_AFFINE) {
			EwStep s{EW_AFFINE, 0, op.ar, op.ai, op.br, op.bi};
			steps.push_back(s);
		} else if(op.kind == Q_SCATTER) {
			run_steps();
			const Scatter *d = (const Scatter *)arena_upload(h, h->qscatter.data() + op.sc0, op.scn * sizeof(Scatter));
			HIPC(launch_scatter(h->L(), h->amp, d, (unsigned)op.scn));
		}
		// Q_FLIP / Q_SET carry no sweep: their exact values arrive via Q_SCATTER.
	}
	run_steps();
}

// Host replay of one elementwise step on a tracked value (same arithmetic as
// k_ew_chain, i.e. the reference's dense_state_backend.cpp:230-252).
void replay(Cplx &v, const QOp &op) {
	if(op.kind == Q_DIFF) {
		v.re = 2.0 * op.ar - v.re;
		v.im = 2.0 * op.ai - v.im;
	} else if(op.kind == Q_AFFINE) {
		const double nr = op.ar * v.re - op.ai * v.im + op.br;
		const double ni = op.ar * v.im + op.ai * v.re + op.bi;
		v.re = nr;
		v.im = ni;
	}
}

Cplx read_device(tmfqs_hip_state *h, u64 s) {
	sync_all(h);
	Cplx v;
	if(h->prec == 0) {
		double buf[2];
		HIPC(hipMemcpy(buf, (char *)h->amp + s * 16, 16, hipMemcpyDeviceToHost));
		v.re = buf[0];
		v.im = buf[1];
	} else {
		float buf[2];
		HIPC(hipMemcpy(buf, (char *)h->amp + s * 8, 8, hipMemcpyDeviceToHost));
		v.re = buf[0];
		v.im = buf[1];
	}
	return v;
}

// A value as the state's storage holds it: complex64 keeps FP32 parts.
Cplx storage_value(const tmfqs_hip_state *h, Cplx v) {
	if(h->prec == 1) {
		v.re = (double)(float)v.re;
		v.im = (double)(float)v.im;
	}
	return v;
}

// Tracked values replay pending elementwise steps in FP64 (complex64 rounds
// them when a flush stores them); queries report them through storage_value,
// as the storage holds them once the steps are flushed.
Cplx current_value(tmfqs_hip_state *h, u64 s) {
	auto it = h->tracked.find(s);
	if(it != h->tracked.end()) return it->second;
	const Cplx v = read_device(h, s);
	h->tracked[s] = v;
	return v;
}

void enqueue_gate(tmfqs_hip_state *h, const QOp &op) {
	if(!h->cfg.fusion) {
		flush(h);
		h->tracked.clear();
		exec_op_unfused(h, op);
		h->qmats.clear();
		return;
	}
	close_ew_run(h);
	h->tracked.clear();
	h->queue.push_back(op);
	if(h->queue.size() >= 8192) flush(h);
}

void enqueue_ew(tmfqs_hip_state *h, const QOp &op) {
	for(auto &kv : h->tracked) replay(kv.second, op);
	h->queue.push_back(op);
	if(op.kind == Q_DIFF || op.kind == Q_AFFINE) {
		++h->ew_steps_pending;
		if(!h->cfg.fusion || h->ew_steps_pending >= ew_block(h)) flush(h);
	}
}

void mark_dirty(tmfqs_hip_state *h, u64 s) {
	if(h->dirty_set.insert(s).second) h->dirty.push_back(s);
}

// ----------------------------------------------------- unfused execution
void exec_op_unfused(tmfqs_hip_state *h, const QOp &op) {
	const Launch L = h->L();
	const double pb = pass_bytes(h);
	switch(op.kind) {
		case Q_H: {
			ProfScope ps(h, "k_hadamard", pb);
			HIPC(launch_hadamard(L, h->amp, h->n, op.b0));
			break;
		}
		case Q_X: {
			ProfScope ps(h, "k_pauli_x", pb);
			HIPC(launch_pauli_x(L, h->amp, h->n, op.b0, 0));
			break;
		}
		case Q_CX: {
			ProfScope ps(h, "k_pauli_x", pb / 2);
			HIPC(launch_pauli_x(L, h->amp, h->n, op.b1, 1ull << op.b0));
			break;
		}
		case Q_CP: {
			ProfScope ps(h, "k_cphase", pb / 4);
			HIPC(launch_cphase(L, h->amp, h->n, op.b1, 1ull << op.b0, op.cr, op.ci));
			break;
		}
		case Q_SWAP: {
			ProfScope ps(h, "k_bitswap", pb / 2);
			HIPC(launch_bitswap(L, h->amp, h->n, op.b0, op.b1));
			break;
		}
		case Q_U1: {
			ProfScope ps(h, "k_u1", pb);
			HIPC(launch_u1(L, h->amp, h->n, op.b0, h->qmats.data() + op.mat));
			break;
		}
		case Q_UK: {
			const size_t dim = 1ull << op.k;
			const double *d = (const double *)arena_upload(h, h->qmats.data() + op.mat, 2 * dim * dim * sizeof(double));
			ProfScope ps(h, "k_uk", pb);
			HIPC(launch_uk(L, h->amp, h->n, op.k, op.bits, d));
			break;
		}
		case Q_PH: {
			const double m[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, op.cr, op.ci};
			ProfScope ps(h, "k_u1", pb);
			HIPC(launch_u1(L, h->amp, h->n, op.b0, m));
			break;
		}
		default: break;
	}
}

// ------------------------------------------------------------ planner
struct PassPlan {
	std::vector<size_t> ops;  // indices into the mapped op list
	u64 active = 0;           // bits needing locality
};

unsigned tile_bits(const tmfqs_hip_state *h) {
	unsigned T = h->cfg.tile_qubits > 0 ? (unsigned)h->cfg.tile_qubits : (h->prec == 0 ? 12u : 13u);
	if(T > h->n) T = h->n;
	return T;
}

unsigned coalesce_bits(const tmfqs_hip_state *h, unsigned T) {
	unsigned C = h->cfg.coalesce_qubits > 0 ? (unsigned)h->cfg.coalesce_qubits : (h->prec == 0 ? 4u : 5u);
	if(C > T) C = T;
	if(T > (unsigned)kMaxHighBits && C < T - (unsigned)kMaxHighBits) C = T - (unsigned)kMaxHighBits;
	return C;
}

u64 nondiag_bits(const QOp &op) {
	switch(op.kind) {
		case Q_H:
		case Q_X:
		case Q_U1: return 1ull << op.b0;
		case Q_CX: return 1ull << op.b1;
		case Q_UK: {
			u64 m = 0;
			for(unsigned i = 0; i < op.k; ++i) m |= 1ull << op.bits[i];
			return m;
		}
		default: return 0;
	}
}

bool is_qft_ladder_cp(const QOp &op, unsigned &hi, unsigned &lo) {
	if(op.kind != Q_CP) return false;
	hi = std::max(op.b0, op.b1);
	lo = std::min(op.b0, op.b1);
	if(hi > 39) return false;
	const double want = kPi / (double)(1ull << (hi - lo));
	return op.theta == want;
}

// Phase-table chunks (chunk j holds F bits 56-8j .. 63-8j, five chunks) that an
// F with bits only in [lo, hi] can touch; bits below the table are dropped by
// the lookup anyway.
static void phase_chunks(unsigned lo, unsigned hi, uint32_t &j0, uint32_t &j1) {
	j0 = std::min(4u, (63u - std::min(hi, 63u)) / 8u);
	j1 = std::min(4u, (63u - std::min(lo, 63u)) / 8u);
}

// A pass that is exactly K <= 8 consecutive QFT targets t = top .. top-K+1,
// each an H on t followed by the complete ladder of controlled phases
// pi/2^(t-c) from every lower bit c (in any order).  Such a pass runs as one
// k_qft_block launch over the block bits top .. top-7 (pbot = top - 7 >= C),
// whose top K bits are its targets.
bool match_qft_block(const std::vector<QOp> &ops, const std::vector<size_t> &idx, unsigned C, unsigned n, unsigned &pbot,
                     unsigned &K) {
	size_t k = 0;
	int targets = 0;
	unsigned prev = 0;
	while(k < idx.size()) {
		const QOp &hop = ops[idx[k]];
		if(hop.kind != Q_H) return false;
		const unsigned t = hop.b0;
		if(targets > 0 && t + 1 != prev) return false;
		++k;
		u64 seen = 0;
		while(k < idx.size() && ops[idx[k]].kind == Q_CP) {
			unsigned hb, lb;
			if(!is_qft_ladder_cp(ops[idx[k]], hb, lb) || hb != t || ((seen >> lb) & 1ull)) return false;
			seen |= 1ull << lb;
			++k;
		}
		if(seen != ((1ull << t) - 1ull)) return false;
		prev = t;
		++targets;
	}
	if(targets < 1 || targets > 8) return false;
	// The targets are the top `targets` bits of the block pbot+7 .. pbot.
	const unsigned top = prev + (unsigned)targets - 1;
	if(top < 7 + C || top >= n) return false;
	pbot = top - 7;
	K = (unsigned)targets;
	return true;
}

// Recognises a complete QFT as makeQft emits it (circuit.cpp:87-106): for
// t = n-1 .. 0, H on t and then the ladder from every lower bit, with the
// swap network leaving the full bit reversal in `map`.  Plans it as an
// optional leading pass of r = (n - 6) mod 8 targets, passes of eight targets
// (k_qft_block), and the last six targets fused with the reversal.
bool plan_qft(const std::vector<QOp> &ops, const std::vector<unsigned> &map, unsigned n, std::vector<PassPlan> &passes) {
	if(n < 12 || n > 40) return false;
	for(unsigned b = 0; b < n; ++b)
		if(map[b] != n - 1 - b) return false;
	std::vector<size_t> start(n + 1);  // start[i]: first op of target n-1-i
	size_t k = 0;
	for(unsigned i = 0; i < n; ++i) {
		const unsigned t = n - 1 - i;
		if(k >= ops.size() || ops[k].kind != Q_H || ops[k].b0 != t) return false;
		start[i] = k++;
		u64 seen = 0;
		while(k < ops.size() && ops[k].kind == Q_CP) {
			unsigned hb, lb;
			if(!is_qft_ladder_cp(ops[k], hb, lb) || hb != t || ((seen >> lb) & 1ull)) return false;
			seen |= 1ull << lb;
			++k;
		}
		if(seen != ((1ull << t) - 1ull)) return false;
	}
	if(k != ops.size()) return false;
	start[n] = ops.size();
	passes.clear();
	auto add = [&](unsigned i0, unsigned i1) {  // targets n-1-i0 down to n-1-(i1-1)
		PassPlan P;
		for(size_t j = start[i0]; j < start[i1]; ++j) {
			P.ops.push_back(j);
			P.active |= nondiag_bits(ops[j]);
		}
		passes.push_back(P);
	};
	const unsigned r = (n - 6) % 8;
	unsigned i = 0;
	if(r) {
		add(0, r);
		i = r;
	}
	for(; i + 6 < n; i += 8) add(i, i + 8);
	add(n - 6, n);
	return true;
}

void run_permutation(tmfqs_hip_state *h, std::vector<unsigned> map, void *buf) {
	// Data of nominal bit b sits at physical map[b]; move it back to b.
	const unsigned n = h->n;
	std::vector<unsigned> pi(n);
	bool ident = true;
	for(unsigned b = 0; b < n; ++b) {
		pi[map[b]] = b;
		if(map[b] != b) ident = false;
	}
	if(ident) return;
	bool invol = true;
	for(unsigned b = 0; b < n; ++b)
		if(pi[pi[b]] != b) invol = false;
	// runs of 2^C amplitudes (512 B) and tiles of up to 2^(2C) per pass
	const unsigned C = h->prec == 0 ? 5u : 6u;
	if(invol && n >= 2 * C + 1) {
		ProfScope ps(h, "k_involution", pass_bytes(h));
		HIPC(launch_involution(h->L(), buf, n, pi.data(), C));
		return;
	}
	// General permutation: transpositions, one in-place pass each.
	std::vector<unsigned> cur = map;  // cur[b] = physical position of nominal b
	std::vector<unsigned> who(n);     // who[p] = nominal bit at physical p
	for(unsigned b = 0; b < n; ++b) who[cur[b]] = b;
	for(unsigned b = 0; b < n; ++b) {
		if(cur[b] == b) continue;
		const unsigned p = cur[b];
		const unsigned other = who[b];
		ProfScope ps(h, "k_bitswap", pass_bytes(h) / 2);
		HIPC(launch_bitswap(h->L(), buf, n, p, b));
		who[p] = other;
		cur[other] = p;
		who[b] = b;
		cur[b] = b;
	}
}

// Groups a launch's tile ops into register stages of W local bits (ops that
// touch a bit outside the stage start a new one; diagonal ops join any stage)
// and LDS steps for k-qubit matrices.  Sets each op's position sj among its
// stage's bits.
std::vector<TileStage> plan_stages(TileOp *ops, size_t cnt, unsigned T, unsigned W) {
	std::vector<TileStage> out;
	TileStage cur{};
	bool open = false;
	uint32_t mask = 0;
	auto close = [&]() {
		if(!open) return;
		// pad with the highest free local bits: lanes then step through low
		// (contiguous) LDS addresses
		uint32_t m = mask;
		for(int b = (int)T - 1; b >= 0 && (unsigned)__builtin_popcount(m) < W; --b) m |= 1u << b;
		unsigned j = 0;
		for(unsigned b = 0; b < T; ++b)
			if((m >> b) & 1u) cur.bits[j++] = b;
		for(uint32_t o = cur.op0; o < cur.op0 + cur.nops; ++o) {
			TileOp &t = ops[o];
			if(t.kind == TOP_H || t.kind == TOP_X || t.kind == TOP_U1)
				for(unsigned q = 0; q < W; ++q)
					if(cur.bits[q] == t.j0) t.sj = q;
		}
		out.push_back(cur);
		open = false;
	};
	for(size_t o = 0; o < cnt; ++o) {
		const TileOp &t = ops[o];
		if(t.kind == TOP_UK) {
			close();
			TileStage l{};
			l.op0 = (uint32_t)o;
			l.nops = 1;
			l.lds = 1;
			out.push_back(l);
			continue;
		}
		const uint32_t need = (t.kind == TOP_H || t.kind == TOP_X || t.kind == TOP_U1) ? (1u << t.j0) : 0u;
		if(open && (unsigned)__builtin_popcount(mask | need) > W) close();
		if(!open) {
			cur = TileStage{};
			cur.op0 = (uint32_t)o;
			mask = 0;
			open = true;
		}
		mask |= need;
		cur.nops++;
	}
	close();
	return out;
}

// One launch of a planned gate run.
struct PassLaunch {
	enum Kind { TILE, QFT_BLOCK, QFT_TAIL, QFT_WBLOCK, QFT_WTAIL } kind = TILE;
	// QFT_BLOCK: block bits pbot+7 .. pbot, the top qk of them targets;
	// QFT_WBLOCK: block bits pbot+9 .. pbot, the top qk of them targets
	unsigned pbot = 0, qc = 0, qk = 0;
	TilePass tp{};              // TILE
	unsigned W = 0;
	std::vector<TileOp> ops;
	std::vector<TileStage> stages;
	std::vector<double> mats;
	std::vector<TwiddleTerm> terms;
	std::vector<uint32_t> gidx;  // dual: global gradient slot of each TOP_GRAD slot
	int ldsops = 0;  // kernel build, see launch_tile_staged
};

struct GatePlan {
	std::vector<PassLaunch> launches;
	std::vector<unsigned> map;  // data of nominal bit b sits at physical map[b] after the launches
	bool permute = false;
};

struct PlanCfg {
	int prec = 0;
	bool dual = false;
	bool ladders = true;  // fuse QFT ladders (not bit-identical to one product per gate)
	unsigned T = 12, C = 4, qft_c = 5;
	unsigned W = 0;  // stage width of tile passes; 0 picks tile_stage_width(prec, T, dual)
	bool mfma_blocks = true;  // complex128 five-qubit blocks on FP64 MFMA (not the reference's accumulation order)
	bool phase_runs = false;  // dual passes: a run of controlled phases as one phase per amplitude (not bit-identical)
	unsigned wide_qft_min = 0;  // complete QFTs of at least this many qubits take the ten-target plan (0: never)
};

GatePlan plan_gates(const PlanCfg &cfg, unsigned n, const std::vector<QOp> &in, size_t lo, size_t hi,
                    const std::vector<double> &qmats) {
	GatePlan plan;
	const bool dual = cfg.dual;
	const bool ladders = cfg.ladders && !dual;
	// 1. relabel swaps into a bit map; op bits become physical positions.
	std::vector<unsigned> map(n);
	for(unsigned b = 0; b < n; ++b) map[b] = b;
	std::vector<QOp> ops;
	ops.reserve(hi - lo);
	for(size_t i = lo; i < hi; ++i) {
		QOp op = in[i];
		if(op.kind == Q_SWAP) {
			std::swap(map[op.b0], map[op.b1]);
			continue;
		}
		op.b0 = map[op.b0];
		op.b1 = map[op.b1];
		for(unsigned j = 0; j < op.k; ++j) op.bits[j] = map[op.bits[j]];
		ops.push_back(op);
	}
	// 2. greedy grouping into passes.
	const unsigned T = std::max(1u, std::min(n, cfg.T));
	const unsigned C = std::min(cfg.C, T);
	const u64 lowC = (C >= 64) ? ~0ull : ((1ull << C) - 1ull);
	std::vector<PassPlan> passes;
	PassPlan cur;
	for(size_t i = 0; i < ops.size(); ++i) {
		const u64 need = nondiag_bits(ops[i]);
		const u64 na = cur.active | need;
		if(!cur.ops.empty() && (unsigned)__builtin_popcountll(na | lowC) > T) {
			passes.push_back(cur);
			cur = PassPlan();
		}
		cur.active |= need;
		cur.ops.push_back(i);
	}
	if(!cur.ops.empty()) passes.push_back(cur);
	// a whole QFT: eight-target block passes and the tail fused with the reversal
	const bool qft_fused = ladders && plan_qft(ops, map, n, passes);
	if(qft_fused && cfg.wide_qft_min && n >= cfg.wide_qft_min && n <= 40) {
		// ten-target plan: a leading pass of (n - 10) mod 10 targets, passes of
		// ten targets down to bit 10, the last ten targets with the reversal
		unsigned top = n - 1;
		const unsigned r = (n - 10) % 10;
		auto block = [&](unsigned pbot, unsigned K) {
			PassLaunch L;
			L.kind = PassLaunch::QFT_WBLOCK;
			L.pbot = pbot;
			L.qk = K;
			plan.launches.push_back(std::move(L));
		};
		if(r) {
			block(n - 10, r);
			top -= r;
		}
		for(; top > 9; top -= 10) block(top - 9, 10);
		PassLaunch L;
		L.kind = PassLaunch::QFT_WTAIL;
		plan.launches.push_back(std::move(L));
		plan.map = map;
		for(unsigned b = 0; b < n; ++b) plan.map[b] = b;
		plan.permute = false;
		return plan;
	}

	// 3. emit each pass as a tile kernel.
	for(size_t pi = 0; pi < passes.size(); ++pi) {
		const PassPlan &P = passes[pi];
		if(qft_fused && pi + 1 == passes.size()) {
			PassLaunch L;
			L.kind = PassLaunch::QFT_TAIL;
			plan.launches.push_back(std::move(L));
			continue;
		}
		unsigned pbot = 0, qk = 0;
		const unsigned qc = cfg.qft_c;
		if(ladders && n <= 40 && match_qft_block(ops, P.ops, qc, n, pbot, qk)) {
			PassLaunch L;
			L.kind = PassLaunch::QFT_BLOCK;
			L.pbot = pbot;
			L.qc = qc;
			L.qk = qk;
			plan.launches.push_back(std::move(L));
			continue;
		}
		unsigned Cp = T;
		std::vector<unsigned> high;
		for(;;) {
			high.clear();
			for(unsigned b = Cp; b < n; ++b)
				if((P.active >> b) & 1ull) high.push_back(b);
			if(Cp + high.size() <= T) break;
			--Cp;
		}
		const unsigned Tp = Cp + (unsigned)high.size();
		auto local = [&](unsigned b) -> int {
			if(b < Cp) return (int)b;
			for(size_t i = 0; i < high.size(); ++i)
				if(high[i] == b) return (int)(Cp + i);
			return -1;
		};
		std::vector<TileOp> tops;
		std::vector<TwiddleTerm> terms;
		std::vector<double> mats;
		// bits touched non-diagonally by ops at or after position k in this pass
		std::vector<u64> later(P.ops.size() + 1, 0);
		for(size_t k = P.ops.size(); k-- > 0;) later[k] = later[k + 1] | nondiag_bits(ops[P.ops[k]]);
		auto add_cond = [&](TileOp &t, unsigned b) {
			const int j = local(b);
			if(j >= 0)
				t.lcond |= 1ull << j;
			else
				t.gcond |= 1ull << b;
		};
		size_t k = 0;
		while(k < P.ops.size()) {
			const QOp &op = ops[P.ops[k]];
			TileOp t{};
			switch(op.kind) {
				case Q_H:
					t.kind = TOP_H;
					t.j0 = (unsigned)local(op.b0);
					tops.push_back(t);
					++k;
					break;
				case Q_X:
					t.kind = TOP_X;
					t.j0 = (unsigned)local(op.b0);
					tops.push_back(t);
					++k;
					break;
				case Q_CX:
					t.kind = TOP_X;
					t.j0 = (unsigned)local(op.b1);
					add_cond(t, op.b0);
					tops.push_back(t);
					++k;
					break;
				case Q_U1:
					t.kind = TOP_U1;
					t.j0 = (unsigned)local(op.b0);
					t.mat = (unsigned)mats.size();
					mats.insert(mats.end(), qmats.begin() + op.mat, qmats.begin() + op.mat + 8);
					tops.push_back(t);
					++k;
					break;
				case Q_UK: {
					t.kind = TOP_UK;
					t.k = op.k;
					for(unsigned j = 0; j < op.k; ++j) t.jl[j] = (unsigned)local(op.bits[j]);
					t.mat = (unsigned)mats.size();
					const size_t dim = 1ull << op.k;
					mats.insert(mats.end(), qmats.begin() + op.mat, qmats.begin() + op.mat + 2 * dim * dim);
					if(cfg.prec == 1 && op.k == 5) {
						// complex64 five-qubit blocks run on FP32 MFMA: the A operands of
						// every lane, already in FP32 and in MFMA order (see tile_mfma5),
						// 16 B aligned after the FP64 matrix.
						if(mats.size() & 1) mats.push_back(0.0);
						t.twid = (unsigned)mats.size();
						std::vector<float> af(2 * 64 * 32);
						const double *g = qmats.data() + op.mat;
						for(unsigned half = 0; half < 2; ++half)
							for(unsigned lane = 0; lane < 64; ++lane)
								for(unsigned s2 = 0; s2 < 32; ++s2) {
									const unsigned rr = lane & 31u, k2 = 2u * s2 + (lane >> 5), kk = k2 & 31u;
									const bool right = k2 >= 32u;
									const double re = g[2 * (rr * 32 + kk)], im = g[2 * (rr * 32 + kk) + 1];
									af[(half * 64 + lane) * 32 + s2] = (float)(half ? (right ? re : im) : (right ? -im : re));
								}
						const size_t at = mats.size();
						mats.resize(at + af.size() / 2);
						std::memcpy(mats.data() + at, af.data(), af.size() * sizeof(float));
					} else if(cfg.prec == 0 && op.k == 5) {
						t.twid = kNoMfmaOperands;
						if(cfg.mfma_blocks) {
							// complex128 five-qubit blocks on FP64 MFMA: lane l of wavefront w
							// takes A[s] = Mbig[16 w + l % 16][4 s + l / 16] (see
							// tile_mfma5_f64), 16 B aligned after the matrix.
							if(mats.size() & 1) mats.push_back(0.0);
							t.twid = (unsigned)mats.size();
							const double *g = qmats.data() + op.mat;
							mats.resize(mats.size() + 4 * 64 * 16);
							double *ad = mats.data() + t.twid;
							for(unsigned w = 0; w < 4; ++w)
								for(unsigned lane = 0; lane < 64; ++lane)
									for(unsigned st = 0; st < 16; ++st) {
										const unsigned R = 16 * w + (lane & 15u), K = 4 * st + (lane >> 4);
										const unsigned r = R & 31u, c = K & 31u;
										const double re = g[2 * (r * 32 + c)], im = g[2 * (r * 32 + c) + 1];
										// [[Mr, -Mi], [Mi, Mr]]
										ad[(w * 64 + lane) * 16 + st] = R < 32 ? (K < 32 ? re : -im) : (K < 32 ? im : re);
									}
						}
					}
					tops.push_back(t);
					++k;
					break;
				}
				case Q_CP: {
					unsigned hb, lb;
					if(ladders && is_qft_ladder_cp(op, hb, lb)) {
						// maximal run of ladder CPs sharing the high bit
						u64 amask = 0;
						size_t e = k;
						while(e < P.ops.size()) {
							unsigned h2, l2;
							if(!is_qft_ladder_cp(ops[P.ops[e]], h2, l2) || h2 != hb) break;
							amask |= 1ull << l2;
							++e;
						}
						const u64 after = later[e];
						const bool hb_free = !((after >> hb) & 1ull);
						const u64 defer = hb_free ? (amask & ~after) : 0;
						const u64 keep = amask & ~defer;
						if(keep) {
							t.kind = TOP_LADDER;
							add_cond(t, hb);
							t.amask = keep;
							t.shift = 63u - hb;
							phase_chunks((unsigned)__builtin_ctzll(keep) + t.shift, 63u - (unsigned)__builtin_clzll(keep) + t.shift,
							             t.cj0, t.cj1);
							tops.push_back(t);
						}
						if(defer) terms.push_back(TwiddleTerm{1ull << hb, defer, 63u - hb, 0});
						k = e;
					} else {
						t.kind = TOP_CP;
						add_cond(t, op.b0);
						add_cond(t, op.b1);
						t.amask = (1ull << op.b0) | (1ull << op.b1);  // its global bits (phase runs)
						t.cr = op.cr;
						t.ci = op.ci;
						tops.push_back(t);
						++k;
					}
					break;
				}
				case Q_PH:
					t.kind = TOP_CP;
					add_cond(t, op.b0);
					t.amask = 1ull << op.b0;
					t.cr = op.cr;
					t.ci = op.ci;
					tops.push_back(t);
					++k;
					break;
				case Q_GRAD:
					if(dual) {
						t.kind = TOP_GRAD;
						add_cond(t, op.b0);
						if(op.k == 2) add_cond(t, op.b1);
						t.twid = (uint32_t)op.state;  // global slot; t.mat gets the launch-local slot
						tops.push_back(t);
					}
					++k;
					break;
				default: ++k; break;
			}
		}
		// Phase runs (dual passes): a run of controlled phases and gradient
		// reductions becomes the reductions followed by one TOP_PHASES op per
		// kMaxPhaseTerms phases.  Diagonal gates commute with the reductions'
		// diagonal projectors and are applied to both states, so each
		// reduction's value does not depend on where in the run it is taken;
		// the amplitudes get one rounding of e^{i sum theta} instead of one
		// per gate.
		std::vector<TwiddleTerm> pterms;
		if(cfg.phase_runs && dual) {
			std::vector<TileOp> out;
			out.reserve(tops.size());
			for(size_t a = 0; a < tops.size();) {
				if(tops[a].kind != TOP_CP && tops[a].kind != TOP_GRAD) {
					out.push_back(tops[a++]);
					continue;
				}
				size_t e = a;
				unsigned ncp = 0;
				for(; e < tops.size() && (tops[e].kind == TOP_CP || tops[e].kind == TOP_GRAD); ++e) ncp += tops[e].kind == TOP_CP;
				if(ncp < 2) {
					out.insert(out.end(), tops.begin() + a, tops.begin() + e);
					a = e;
					continue;
				}
				for(size_t q = a; q < e; ++q)
					if(tops[q].kind == TOP_GRAD) out.push_back(tops[q]);
				TileOp ph{};
				for(size_t q = a; q < e; ++q) {
					if(tops[q].kind != TOP_CP) continue;
					if(ph.nterms == 0) {
						ph = TileOp{};
						ph.kind = TOP_PHASES;
						ph.twid = (uint32_t)pterms.size();
					}
					const double th = std::atan2(tops[q].ci, tops[q].cr);
					u64 bits;
					std::memcpy(&bits, &th, sizeof bits);
					pterms.push_back(TwiddleTerm{tops[q].amask, bits, 0, 0});
					if(++ph.nterms == kMaxPhaseTerms) {
						out.push_back(ph);
						ph.nterms = 0;
					}
				}
				if(ph.nterms) out.push_back(ph);
				a = e;
			}
			tops.swap(out);
		}
		// deferred ladder factors: one twiddle per kMaxTwiddleTerms terms, at the end
		std::vector<TwiddleTerm> allterms;
		for(size_t s = 0; s < terms.size(); s += kMaxTwiddleTerms) {
			TileOp t{};
			t.kind = TOP_TWIDDLE;
			t.twid = (unsigned)allterms.size();
			t.nterms = (unsigned)std::min<size_t>(kMaxTwiddleTerms, terms.size() - s);
			// F is a sum of the terms: its lowest bit is the lowest term bit, its
			// highest is bounded by the sum of the term maxima (wrapping mod 2^64).
			unsigned lo = 63;
			unsigned __int128 fmax = 0;
			for(unsigned q = 0; q < t.nterms; ++q) {
				const TwiddleTerm &w = terms[s + q];
				allterms.push_back(w);
				lo = std::min(lo, (unsigned)__builtin_ctzll(w.mask) + w.shift);
				fmax += (unsigned __int128)w.mask << w.shift;
			}
			const unsigned hi = (fmax >> 64) ? 63u : 63u - (unsigned)__builtin_clzll((u64)fmax);
			phase_chunks(lo, hi, t.cj0, t.cj1);
			tops.push_back(t);
		}
		if(!pterms.empty()) {
			const uint32_t off = (uint32_t)allterms.size();
			allterms.insert(allterms.end(), pterms.begin(), pterms.end());
			for(TileOp &o : tops)
				if(o.kind == TOP_PHASES) o.twid += off;
		}
		if(tops.empty()) continue;
		TilePass tp{};
		tp.n = n;
		tp.T = Tp;
		tp.C = Cp;
		tp.nh = (unsigned)high.size();
		for(size_t i = 0; i < high.size(); ++i) tp.hbits[i] = high[i];
		tp.ntiles = 1ull << (n - Tp);
		const unsigned W = std::min(Tp, cfg.W ? cfg.W : tile_stage_width(cfg.prec, cfg.T, dual));
		// op lists longer than kMaxTileOps (or with more than kMaxGradSlots
		// gradient reductions) are split across launches.
		size_t s0 = 0;
		while(s0 < tops.size()) {
			PassLaunch L;
			L.kind = PassLaunch::TILE;
			size_t e = s0;
			while(e < tops.size() && e - s0 < (size_t)kMaxTileOps) {
				if(tops[e].kind == TOP_GRAD) {
					if(L.gidx.size() == kMaxGradSlots) break;
					tops[e].mat = (uint32_t)L.gidx.size();
					L.gidx.push_back(tops[e].twid);
				}
				++e;
			}
			L.ops.assign(tops.begin() + s0, tops.begin() + e);
			L.stages = plan_stages(L.ops.data(), L.ops.size(), Tp, W);
			for(const TileStage &st : L.stages)
				if(st.lds) L.ldsops = 1;
			unsigned Wl = W;
			if(cfg.prec == 0 && !L.ldsops && !cfg.W && W > 1 && Tp == W + 8) {
				// complex128, register stages only: 512 lanes with stages one bit
				// narrower hold 127 (dual: 103) VGPRs, so four wavefronts share a
				// SIMD at the same LDS tile (measured: Hadamard layers 8-9% and the
				// n = 24 adjoint 7% faster; the complex64 dual pass was slower)
				Wl = W - 1;
				L.stages = plan_stages(L.ops.data(), L.ops.size(), Tp, Wl);
			}
			if(L.ldsops && cfg.prec == 0 && Tp >= 11) {
				// complex128 five-qubit blocks, all with MFMA operands: the MFMA build
				bool k5 = false, all = true;
				for(const TileOp &o : L.ops)
					if(o.kind == TOP_UK && o.k == 5) {
						k5 = true;
						all = all && o.twid != kNoMfmaOperands;
					}
				if(k5 && all) L.ldsops = 2;
			}
			L.mats = mats;
			L.terms = allterms;
			L.tp = tp;
			L.tp.nops = (unsigned)L.ops.size();
			L.tp.op0 = 0;
			L.W = Wl;
			plan.launches.push_back(std::move(L));
			s0 = e;
		}
	}
	// 4. the final bit layout (the fused QFT tail already restores it).
	plan.map = map;
	plan.permute = false;
	if(!qft_fused)
		for(unsigned b = 0; b < n; ++b) plan.permute = plan.permute || map[b] != b;
	return plan;
}

// A complete QFT in makeQft's op order (circuit.cpp:87-106) starting at
// in[s]: for t = n-1 .. 0 an H on t and then the ladder pi/2^(t-c) from every
// lower bit c (in any order), then floor(n/2) swaps composing to the full bit
// reversal.  Returns the end of the unit, or s when there is none.
static size_t qft_unit_end(const std::vector<QOp> &in, size_t s, size_t hi, unsigned n) {
	size_t k = s;
	for(unsigned i = 0; i < n; ++i) {
		const unsigned t = n - 1 - i;
		if(k >= hi || in[k].kind != Q_H || in[k].b0 != t) return s;
		++k;
		const u64 full = (1ull << t) - 1ull;
		u64 seen = 0;
		while(seen != full) {
			unsigned hb, lb;
			if(k >= hi || !is_qft_ladder_cp(in[k], hb, lb) || hb != t || ((seen >> lb) & 1ull)) return s;
			seen |= 1ull << lb;
			++k;
		}
	}
	std::vector<unsigned> m(n);
	for(unsigned b = 0; b < n; ++b) m[b] = b;
	for(unsigned i = 0; i < n / 2; ++i, ++k) {
		if(k >= hi || in[k].kind != Q_SWAP) return s;
		std::swap(m[in[k].b0], m[in[k].b1]);
	}
	for(unsigned b = 0; b < n; ++b)
		if(m[b] != n - 1 - b) return s;
	return k;
}

// ------------------------------------------------------------ gate-run composition
// Physical bits a gate acts on; false for operations a composed run cannot
// hold (swaps are relabelings handled by the planner, elementwise steps touch
// every amplitude).
bool gate_support(const QOp &op, u64 &support) {
	switch(op.kind) {
		case Q_H:
		case Q_X:
		case Q_U1:
		case Q_PH: support = 1ull << op.b0; return true;
		case Q_CX:
		case Q_CP: support = (1ull << op.b0) | (1ull << op.b1); return true;
		case Q_UK:
			support = 0;
			for(unsigned j = 0; j < op.k; ++j) support |= 1ull << op.bits[j];
			return true;
		default: return false;
	}
}

// Applies gate op to v, a vector over the 2^k local indices of bits
// (bits[0] is the index MSB), in FP64: the columns of a composed matrix.
void apply_local(const QOp &op, const std::vector<double> &qmats, const unsigned *bits, unsigned k,
                 std::complex<double> *v) {
	using C = std::complex<double>;
	auto lm = [&](unsigned p) -> unsigned {
		for(unsigned i = 0; i < k; ++i)
			if(bits[i] == p) return 1u << (k - 1 - i);
		return 0;
	};
	const unsigned D = 1u << k;
	switch(op.kind) {
		case Q_H: {
			const unsigned m = lm(op.b0);
			const double c = 1.0 / std::sqrt(2.0);
			for(unsigned r = 0; r < D; ++r)
				if(!(r & m)) {
					const C a = v[r], b = v[r | m];
					v[r] = (a + b) * c;
					v[r | m] = (a - b) * c;
				}
			break;
		}
		case Q_X:
		case Q_CX: {
			const unsigned m = lm(op.kind == Q_X ? op.b0 : op.b1), cm = op.kind == Q_X ? 0u : lm(op.b0);
			for(unsigned r = 0; r < D; ++r)
				if(!(r & m) && (r & cm) == cm) std::swap(v[r], v[r | m]);
			break;
		}
		case Q_CP:
		case Q_PH: {
			const unsigned m = op.kind == Q_PH ? lm(op.b0) : (lm(op.b0) | lm(op.b1));
			const C f(op.cr, op.ci);
			for(unsigned r = 0; r < D; ++r)
				if((r & m) == m) v[r] *= f;
			break;
		}
		case Q_U1:
		case Q_UK: {
			const unsigned kg = op.kind == Q_U1 ? 1u : op.k;
			const unsigned Dg = 1u << kg;
			unsigned gm[5];
			for(unsigned j = 0; j < kg; ++j) gm[j] = lm(op.kind == Q_U1 ? op.b0 : op.bits[j]);
			unsigned all = 0;
			for(unsigned j = 0; j < kg; ++j) all |= gm[j];
			const double *g = qmats.data() + op.mat;
			std::vector<C> x(Dg);
			std::vector<unsigned> idx(Dg);
			for(unsigned base = 0; base < D; ++base) {
				if(base & all) continue;
				for(unsigned r = 0; r < Dg; ++r) {
					unsigned off = 0;
					for(unsigned j = 0; j < kg; ++j)
						if((r >> (kg - 1 - j)) & 1u) off |= gm[j];
					idx[r] = base | off;
					x[r] = v[idx[r]];
				}
				for(unsigned r = 0; r < Dg; ++r) {
					C acc(0.0, 0.0);
					for(unsigned c = 0; c < Dg; ++c) acc += C(g[2 * (r * Dg + c)], g[2 * (r * Dg + c) + 1]) * x[c];
					v[idx[r]] = acc;
				}
			}
			break;
		}
		default: break;
	}
}

// Replaces every run of consecutive gates that acts on at most
// cfg.fuse_qubits qubits and holds at least cfg.fuse_min_gates gates (one of
// them not diagonal) by one dense block: the product of the run's matrices,
// formed column by column in FP64 on the host (the matrix product of
// src/tmfqs/gates/quantum_gate.cpp:59-77, gate after gate).  Five-qubit
// composition pads smaller runs with low bits up to five qubits, so blocks
// take the MFMA path.  Returns false when no run qualifies.
bool compose_gate_runs(tmfqs_hip_state *h, const std::vector<QOp> &in, size_t lo, size_t hi, std::vector<QOp> &out) {
	const unsigned kmax = (unsigned)h->cfg.fuse_qubits;
	const size_t minGates = h->cfg.fuse_min_gates > 0 ? (size_t)h->cfg.fuse_min_gates : 24;
	const bool pad5 = kmax == 5 && h->n >= (h->prec == 0 ? 11u : 10u);
	out.clear();
	bool any = false;
	size_t s = lo;
	while(s < hi) {
		u64 sup = 0, Q = 0;
		if(!gate_support(in[s], sup)) {
			out.push_back(in[s++]);
			continue;
		}
		size_t e = s;
		bool nondiag = false;
		while(e < hi && gate_support(in[e], sup) && (unsigned)__builtin_popcountll(Q | sup) <= kmax) {
			Q |= sup;
			nondiag = nondiag || !(in[e].kind == Q_CP || in[e].kind == Q_PH);
			++e;
		}
		if(e - s < minGates || !nondiag) {
			out.insert(out.end(), in.begin() + s, in.begin() + e);
			s = e;
			continue;
		}
		for(unsigned b = 0; pad5 && __builtin_popcountll(Q) < 5 && b < h->n; ++b) Q |= 1ull << b;
		unsigned bits[5], k = 0;
		for(int b = 63; b >= 0; --b)
			if((Q >> b) & 1ull) bits[k++] = (unsigned)b;
		const unsigned D = 1u << k;
		std::vector<std::complex<double>> col(D);
		std::vector<double> M(2 * (size_t)D * D);
		for(unsigned c = 0; c < D; ++c) {
			std::fill(col.begin(), col.end(), std::complex<double>(0.0, 0.0));
			col[c] = 1.0;
			for(size_t i = s; i < e; ++i) apply_local(in[i], h->qmats, bits, k, col.data());
			for(unsigned r = 0; r < D; ++r) {
				M[2 * ((size_t)r * D + c)] = col[r].real();
				M[2 * ((size_t)r * D + c) + 1] = col[r].imag();
			}
		}
		QOp op;
		op.mat = h->qmats.size();
		h->qmats.insert(h->qmats.end(), M.begin(), M.end());
		if(k == 1) {
			op.kind = Q_U1;
			op.b0 = bits[0];
		} else {
			op.kind = Q_UK;
			op.k = k;
			for(unsigned j = 0; j < k; ++j) op.bits[j] = bits[j];
		}
		out.push_back(op);
		any = true;
		if(h->profiling) h->prof["gate_runs_composed"].launches += 1;  // a host step: no device time or bytes
		s = e;
	}
	return any;
}

void exec_planned(tmfqs_hip_state *h, const std::vector<QOp> &in, size_t lo, size_t hi, const GateTarget *tgt);

void exec_gates(tmfqs_hip_state *h, const std::vector<QOp> &in, size_t lo, size_t hi, const GateTarget *tgt) {
	const unsigned n = h->n;
	if(!tgt && !h->cfg.exact_phases && n >= 12 && n <= 40) {
		// A complete QFT inside a longer gate run (other gates around it, or
		// several QFTs) is planned on its own, so it runs on the QFT kernels.
		const size_t len = (size_t)n + (size_t)n * (n - 1) / 2 + n / 2;
		if(hi - lo > len)
			for(size_t s = lo; s + len <= hi; ++s) {
				if(in[s].kind != Q_H || in[s].b0 != n - 1) continue;
				const size_t e = qft_unit_end(in, s, hi, n);
				if(e == s) continue;
				if(s > lo) exec_gates(h, in, lo, s, tgt);
				exec_gates(h, in, s, e, tgt);
				if(e < hi) exec_gates(h, in, e, hi, tgt);
				return;
			}
	}
	if((!tgt || !tgt->b) && !h->cfg.exact_phases && h->cfg.fuse_qubits >= 2 &&
	   !(n >= 12 && n <= 40 && qft_unit_end(in, lo, hi, n) == hi)) {
		std::vector<QOp> fused;
		if(compose_gate_runs(h, in, lo, hi, fused)) {
			exec_planned(h, fused, 0, fused.size(), tgt);
			return;
		}
	}
	exec_planned(h, in, lo, hi, tgt);
}

// Plans a gate run (no complete QFT cut out, runs already composed) into
// tiled passes and launches them.
void exec_planned(tmfqs_hip_state *h, const std::vector<QOp> &in, size_t lo, size_t hi, const GateTarget *tgt) {
	const unsigned n = h->n;
	void *const A = tgt ? tgt->a : h->amp;
	void *const B = tgt ? tgt->b : nullptr;
	const bool dual = B != nullptr;
	PlanCfg cfg;
	cfg.prec = h->prec;
	cfg.dual = dual;
	cfg.ladders = !h->cfg.exact_phases;
	cfg.mfma_blocks = !h->cfg.exact_phases;
	// complex128 only: for complex64 the FP64 angle sum and sincos cost more
	// than the FP32 phase products they replace (measured 10.2 -> 11.7 ms)
	cfg.phase_runs = dual && !h->cfg.exact_phases && h->prec == 0;
	cfg.T = tile_bits(h);
	if(tgt && tgt->tile_delta) cfg.T = (unsigned)std::max<int>(1, std::min<int>((int)n, (int)cfg.T + tgt->tile_delta));
	cfg.C = coalesce_bits(h, cfg.T);
	cfg.qft_c = qft_block_coalesce_bits(h->prec, h->cfg.coalesce_qubits);
	if(!tgt && h->cfg.qft_plan == 0) cfg.wide_qft_min = qft_wide_min_qubits(h->prec);
	GatePlan plan = plan_gates(cfg, n, in, lo, hi, h->qmats);
	if(!plan.launches.empty() && plan.launches.back().kind == PassLaunch::QFT_WTAIL && !reserve_scratch(h)) {
		cfg.wide_qft_min = 0;  // no room for the second state buffer: in-place plan
		plan = plan_gates(cfg, n, in, lo, hi, h->qmats);
	}
	if(!dual && !tgt && h->cfg.tile_qubits == 0 && cfg.T > 8 && cfg.T == (h->prec == 0 ? 12u : 13u)) {
		// One bit narrower tiles run about a fifth faster per pass (four
		// workgroups per CU, three-bit stages); take them when they need no
		// more passes.
		PlanCfg c2 = cfg;
		c2.T = cfg.T - 1;
		c2.C = coalesce_bits(h, c2.T);
		GatePlan p2 = plan_gates(c2, n, in, lo, hi, h->qmats);
		if(p2.launches.size() <= plan.launches.size() && (plan.permute || !p2.permute)) plan = std::move(p2);
	}
	for(const PassLaunch &L : plan.launches) {
		if(L.kind == PassLaunch::QFT_WBLOCK) {
			ProfScope ps(h, "k_qft_wide_block", pass_bytes(h));
			HIPC(launch_qft_wide_block(h->L(), A, n, L.pbot, L.qk, h->d_phase, h->d_cin, 0));
			continue;
		}
		if(L.kind == PassLaunch::QFT_WTAIL) {
			// the last launch of the plan: the state moves into the second buffer,
			// which becomes the state buffer
			{
				ProfScope ps(h, "k_qft_wide_tail", pass_bytes(h));
				HIPC(launch_qft_wide_tail(h->L(), A, h->scratch, n, h->d_phase, h->d_cin, 0));
			}
			std::swap(h->amp, h->scratch);
			std::swap(h->amp_bytes, h->scratch_bytes);
			continue;
		}
		if(L.kind == PassLaunch::QFT_TAIL) {
			ProfScope ps(h, "k_qft_tail_reverse", pass_bytes(h));
			HIPC(launch_qft_tail_reverse(h->L(), A, n, h->d_cin));
			continue;
		}
		if(L.kind == PassLaunch::QFT_BLOCK) {
			ProfScope ps(h, "k_qft_block", pass_bytes(h));
			HIPC(launch_qft_block(h->L(), A, n, L.pbot, h->d_phase, h->d_cin, L.qc, L.qk));
			continue;
		}
		const TileOp *dops = (const TileOp *)arena_upload(h, L.ops.data(), L.ops.size() * sizeof(TileOp));
		const TileStage *dst = (const TileStage *)arena_upload(h, L.stages.data(), L.stages.size() * sizeof(TileStage));
		const double *dm = L.mats.empty() ? nullptr : (const double *)arena_upload(h, L.mats.data(), L.mats.size() * sizeof(double));
		const TwiddleTerm *dt =
		    L.terms.empty() ? nullptr : (const TwiddleTerm *)arena_upload(h, L.terms.data(), L.terms.size() * sizeof(TwiddleTerm));
		{
			ProfScope ps(h, dual ? "k_tile_dual" : "k_tile", (dual ? 2 : 1) * pass_bytes(h));
			HIPC(launch_tile_staged(h->L(), A, B, L.tp, L.W, dops, dst, (unsigned)L.stages.size(), dm, dt, h->d_phase,
			                        dual ? tgt->gpart : nullptr, (unsigned)L.gidx.size(), L.ldsops));
		}
		if(dual && !L.gidx.empty()) {
			const uint32_t *dg = (const uint32_t *)arena_upload(h, L.gidx.data(), L.gidx.size() * sizeof(uint32_t));
			HIPC(launch_grad_finish(h->L(), tgt->gpart, tile_grid(L.tp, true, h->prec), (unsigned)L.gidx.size(), dg, tgt->grad_scale,
			                        tgt->dgrads));
		}
	}
	if(plan.permute) {
		run_permutation(h, plan.map, A);
		if(dual) run_permutation(h, plan.map, B);
	}
}

// ------------------------------------------------------------ op records
// tmfqs_hip_op -> queued ops on physical bits, with the reference's argument
// checks; records the index of every parameterized op (CPHASE, PHASE).
std::vector<QOp> convert_ops(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool,
                             const unsigned *qubit_pool, bool allow_elementwise, std::vector<size_t> *params) {
	std::vector<QOp> out;
	out.reserve(count);
	for(size_t i = 0; i < count; ++i) {
		const tmfqs_hip_op &o = ops[i];
		QOp q;
		switch(o.kind) {
			case TMFQS_HIP_OP_H:
				validate_qubit(h, o.q0, "HipStateBackend::applyHadamard");
				q.kind = Q_H;
				q.b0 = pbit(h, o.q0);
				break;
			case TMFQS_HIP_OP_X:
				validate_qubit(h, o.q0, "HipStateBackend::applyPauliX");
				q.kind = Q_X;
				q.b0 = pbit(h, o.q0);
				break;
			case TMFQS_HIP_OP_CX:
				validate_two(h, o.q0, o.q1, "HipStateBackend::applyControlledNot");
				q.kind = Q_CX;
				q.b0 = pbit(h, o.q0);
				q.b1 = pbit(h, o.q1);
				break;
			case TMFQS_HIP_OP_CPHASE:
				validate_two(h, o.q0, o.q1, "HipStateBackend::applyControlledPhaseShift");
				q.kind = Q_CP;
				q.b0 = pbit(h, o.q0);
				q.b1 = pbit(h, o.q1);
				q.theta = o.theta;
				__builtin_sincos(o.theta, &q.ci, &q.cr);
				if(params) params->push_back(out.size());
				break;
			case TMFQS_HIP_OP_PHASE:
				validate_qubit(h, o.q0, "phase gate");
				q.kind = Q_PH;
				q.b0 = pbit(h, o.q0);
				q.theta = o.theta;
				__builtin_sincos(o.theta, &q.ci, &q.cr);
				if(params) params->push_back(out.size());
				break;
			case TMFQS_HIP_OP_SWAP:
				validate_two(h, o.q0, o.q1, "QuantumRegister::applySwap");
				q.kind = Q_SWAP;
				q.b0 = pbit(h, o.q0);
				q.b1 = pbit(h, o.q1);
				break;
			case TMFQS_HIP_OP_MATRIX: {
				const char *scope = "HipStateBackend::applyGate";
				if(!mat_pool || !qubit_pool) throw std::invalid_argument(std::string(scope) + " needs the matrix and qubit pools");
				if(o.k < 1 || o.k > 5 || o.k > h->n) throw std::invalid_argument(std::string(scope) + " supports gates on 1 to 5 qubits");
				const unsigned *qs = qubit_pool + o.q0;
				u64 seen = 0;
				for(unsigned j = 0; j < o.k; ++j) {
					validate_qubit(h, qs[j], scope);
					if((seen >> qs[j]) & 1ull) throw std::invalid_argument(std::string(scope) + " duplicate qubit index");
					seen |= 1ull << qs[j];
				}
				const size_t dim = 1ull << o.k;
				q.mat = h->qmats.size();
				h->qmats.insert(h->qmats.end(), mat_pool + o.state, mat_pool + o.state + 2 * dim * dim);
				if(o.k == 1) {
					q.kind = Q_U1;
					q.b0 = pbit(h, qs[0]);
				} else {
					q.kind = Q_UK;
					q.k = o.k;
					for(unsigned j = 0; j < o.k; ++j) q.bits[j] = pbit(h, qs[j]);
				}
				break;
			}
			case TMFQS_HIP_OP_PHASE_FLIP:
			case TMFQS_HIP_OP_DIFFUSE_LAZY:
			case TMFQS_HIP_OP_DIFFUSE_MATERIALIZED:
				if(!allow_elementwise) throw std::invalid_argument("adjoint: only gate operations can be differentiated");
				q.kind = -1 - o.kind;  // dispatched by the caller
				q.state = o.state;
				break;
			default: throw std::invalid_argument("unknown operation kind " + std::to_string(o.kind));
		}
		out.push_back(q);
	}
	return out;
}

// ------------------------------------------------------------ adjoint
// The observable's device form on n qubits (qubit q is bit n-1-q).
ObsParams make_observable(unsigned n, const tmfqs_hip_diag_term *terms, size_t nterms) {
	ObsParams O{};
	if(nterms && !terms) throw std::invalid_argument("diagonal observable: null term list");
	for(size_t i = 0; i < nterms; ++i) {
		const tmfqs_hip_diag_term &t = terms[i];
		if(t.kind == TMFQS_HIP_DIAG_Z) {
			if(O.nz == kMaxObsTerms) throw std::invalid_argument("diagonal observable: too many Z terms");
			u64 m = 0;
			for(unsigned q = 0; q < 64; ++q)
				if((t.bits >> q) & 1ull) {
					if(q >= n) throw std::out_of_range("diagonal observable qubit index out of range");
					m |= 1ull << (n - 1u - q);
				}
			O.zmask[O.nz] = m;
			O.zcoef[O.nz++] = t.coef;
		} else if(t.kind == TMFQS_HIP_DIAG_PROJECTOR) {
			if(O.np == kMaxObsTerms) throw std::invalid_argument("diagonal observable: too many projector terms");
			if(n < 64 && t.bits >= (1ull << n)) throw std::out_of_range("diagonal observable state index out of range");
			O.pstate[O.np] = t.bits;
			O.pcoef[O.np++] = t.coef;
		} else {
			throw std::invalid_argument("diagonal observable: unknown term kind");
		}
	}
	return O;
}

ObsParams make_observable(const tmfqs_hip_state *h, const tmfqs_hip_diag_term *terms, size_t nterms) {
	return make_observable(h->n, terms, nterms);
}

double device_expectation(tmfqs_hip_state *h, const void *a, void *lam, const ObsParams &O) {
	{
		ProfScope ps(h, "k_observable", (lam ? 2.0 : 1.0) * (double)h->N * (double)h->esz);
		HIPC(launch_observable(h->L(), a, lam, h->n, O, h->d_red, h->d_red + 2 * kReduceBlocks));
	}
	double e = 0;
	HIPC(hipMemcpyAsync(&e, h->d_red + 2 * kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIPC(hipStreamSynchronize(h->stream));
	return e;
}

// The inverse of a gate op (matrices are conjugate-transposed into qmats).
QOp inverse_op(tmfqs_hip_state *h, const QOp &op) {
	QOp r = op;
	switch(op.kind) {
		case Q_CP:
		case Q_PH:
			r.ci = -op.ci;
			r.theta = -op.theta;
			break;
		case Q_U1:
		case Q_UK: {
			const unsigned k = op.kind == Q_U1 ? 1u : op.k;
			const size_t dim = 1ull << k;
			r.mat = h->qmats.size();
			h->qmats.resize(r.mat + 2 * dim * dim);
			for(size_t i = 0; i < dim; ++i)
				for(size_t j = 0; j < dim; ++j) {
					h->qmats[r.mat + 2 * (i * dim + j)] = h->qmats[op.mat + 2 * (j * dim + i)];
					h->qmats[r.mat + 2 * (i * dim + j) + 1] = -h->qmats[op.mat + 2 * (j * dim + i) + 1];
				}
			break;
		}
		default: break;  // H, X, CX, swap are involutions
	}
	return r;
}

// Segment boundaries for forward checkpoints: the first pass boundary at
// least `seg` ops after the previous checkpoint (forced after 4 * seg ops),
// so checkpointing adds no pass to the fused forward sweep.
std::vector<size_t> checkpoint_bounds(const tmfqs_hip_state *h, const std::vector<QOp> &ops, size_t seg) {
	const unsigned n = h->n;
	const unsigned T = tile_bits(h);
	const unsigned C = coalesce_bits(h, T);
	const u64 lowC = (C >= 64) ? ~0ull : ((1ull << C) - 1ull);
	std::vector<size_t> b{0};
	std::vector<unsigned> map(n);
	for(unsigned i = 0; i < n; ++i) map[i] = i;
	u64 active = 0;
	bool empty = true;
	size_t last = 0;
	for(size_t i = 0; i < ops.size(); ++i) {
		QOp op = ops[i];
		if(op.kind == Q_SWAP) {
			std::swap(map[op.b0], map[op.b1]);
			continue;
		}
		op.b0 = map[op.b0];
		op.b1 = map[op.b1];
		for(unsigned j = 0; j < op.k; ++j) op.bits[j] = map[op.bits[j]];
		const u64 need = nondiag_bits(op);
		const bool newpass = !empty && (unsigned)__builtin_popcountll(active | need | lowC) > T;
		if((newpass && i - last >= seg) || i - last >= 4 * seg) {
			b.push_back(i);
			last = i;
			for(unsigned j = 0; j < n; ++j) map[j] = j;
			op = ops[i];
			active = 0;
		} else if(newpass) {
			active = 0;
		}
		active |= nondiag_bits(op);
		empty = false;
	}
	b.push_back(ops.size());
	return b;
}

// Enqueues the whole adjoint sweep on the handle's stream without waiting
// for it: E goes to dev_e and the gradients of the P parameters (in params
// order) to dev_grads, both device memory.
void adjoint_enqueue(tmfqs_hip_state *h, std::vector<QOp> &qops, const std::vector<size_t> &params, const ObsParams &O,
                     unsigned checkpoint_gates, double *dev_e, double *dev_grads) {
	const size_t G = qops.size();
	size_t seg = checkpoint_gates ? checkpoint_gates : (size_t)std::ceil(std::sqrt((double)std::max<size_t>(G, 1)));
	std::vector<size_t> bounds = checkpoint_bounds(h, qops, seg);
	// one buffer per segment end plus lambda; fewer checkpoints when HBM is short
	size_t freeb = 0, totalb = 0;
	HIPC(hipMemGetInfo(&freeb, &totalb));
	freeb += h->adj_bufs.size() * h->adj_buf_bytes;  // reusable
	const size_t budget = (size_t)(0.9 * (double)freeb);
	while(bounds.size() > 2 && bounds.size() * h->amp_bytes > budget) {
		seg *= 2;
		bounds = checkpoint_bounds(h, qops, seg);
	}
	const size_t S = bounds.size() - 1;
	const size_t nbuf = S + 1;
	if(h->adj_buf_bytes != h->amp_bytes) {
		HIPC(hipStreamSynchronize(h->stream));
		for(void *p : h->adj_bufs) hipFree(p);
		h->adj_bufs.clear();
		h->adj_buf_bytes = h->amp_bytes;
	}
	while(h->adj_bufs.size() < nbuf) {
		void *p = nullptr;
		HIPC(hipMalloc(&p, h->adj_buf_bytes));
		h->adj_bufs.push_back(p);
	}
	void *lam = h->adj_bufs[S];
	const size_t P = params.size();
	if(!h->d_gpart) HIPC(hipMalloc((void **)&h->d_gpart, (size_t)kMaxGradSlots * kMaxDualGrid * sizeof(double)));
	// forward sweep with a checkpoint at every segment end
	for(size_t s = 0; s < S; ++s) {
		exec_gates(h, qops, bounds[s], bounds[s + 1]);
		ProfScope ps(h, "k_copy", pass_bytes(h));
		HIPC(launch_copy(h->L(), h->amp, h->adj_bufs[s], h->n));
	}
	{
		ProfScope ps(h, "k_observable", 2.0 * (double)h->N * (double)h->esz);
		HIPC(launch_observable(h->L(), h->adj_bufs[S - 1], lam, h->n, O, h->d_red, h->d_red + 2 * kReduceBlocks));
	}
	HIPC(hipMemcpyAsync(dev_e, h->d_red + 2 * kReduceBlocks, sizeof(double), hipMemcpyDeviceToDevice, h->stream));
	if(P == 0) return;
	HIPC(hipMemsetAsync(dev_grads, 0, P * sizeof(double), h->stream));
	std::vector<int64_t> slot(G, -1);
	for(size_t p = 0; p < P; ++p) slot[params[p]] = (int64_t)p;
	// backward sweep: psi restarts from each segment's checkpoint, lambda runs through
	GateTarget tg;
	tg.b = lam;
	// dual tiles of 11 bits in both precisions (two LDS tiles per workgroup;
	// measured fastest: 12-bit complex64 dual tiles are 30% slower, 10-bit
	// complex128 ones 70% slower)
	tg.tile_delta = h->prec == 0 ? -1 : -2;
	tg.gpart = h->d_gpart;
	tg.dgrads = dev_grads;
	tg.grad_scale = -2.0;  // dE/dtheta = -2 Im <lambda| P |psi>
	for(size_t s = S; s-- > 0;) {
		std::vector<QOp> rev;
		for(size_t i = bounds[s + 1]; i-- > bounds[s];) {
			const QOp &op = qops[i];
			if(slot[i] >= 0) {
				QOp g;
				g.kind = Q_GRAD;
				g.b0 = op.b0;
				g.b1 = op.b1;
				g.k = op.kind == Q_CP ? 2u : 1u;
				g.state = (u64)slot[i];
				rev.push_back(g);
			}
			rev.push_back(inverse_op(h, op));
		}
		if(s == 0) {  // nothing before the first parameter needs undoing
			size_t keep = 0;
			for(size_t i = 0; i < rev.size(); ++i)
				if(rev[i].kind == Q_GRAD) keep = i + 1;
			rev.resize(keep);
		}
		tg.a = h->adj_bufs[s];
		exec_gates(h, rev, 0, rev.size(), &tg);
	}
}

void adjoint_run(tmfqs_hip_state *h, std::vector<QOp> &qops, const std::vector<size_t> &params, const ObsParams &O,
                 unsigned checkpoint_gates, double *expval, double *grads) {
	const size_t P = params.size();
	if(h->grads_cap < P + 1) {
		HIPC(hipStreamSynchronize(h->stream));
		hipFree(h->d_grads);
		h->d_grads = nullptr;
		h->grads_cap = 0;
		HIPC(hipMalloc((void **)&h->d_grads, (P + 1) * sizeof(double)));
		h->grads_cap = P + 1;
	}
	adjoint_enqueue(h, qops, params, O, checkpoint_gates, h->d_grads + P, h->d_grads);
	if(P) HIPC(hipMemcpyAsync(grads, h->d_grads, P * sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIPC(hipMemcpyAsync(expval, h->d_grads + P, sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIPC(hipStreamSynchronize(h->stream));
}

// ------------------------------------------------------------ sampling
// Exact replay of the reference's sequential FP64 chain S = fl(S + p_i)
// (dense_state_backend.cpp:173-187, quantum_register.cpp:289-309).  Chunk
// sums locate each chunk's binade; inside one binade every step adds
// rint(p_i / ulp) ulps exactly, so whole chunks advance by an integer count
// computed on the device.  Chunks that cross a binade, contain a tie, or start
// at S == 0 are replayed element by element on the host.
struct ChainResult {
	bool found = false;
	u64 index = 0;
	double total = 0;
};

inline double host_prob(double re, double im, const double *af) {
	if(af) {
		const double mr = af[0] * re - af[1] * im;
		const double mi = af[0] * im + af[1] * re;
		re = mr + af[2];
		im = mi + af[3];
	}
	return re * re + im * im;
}

void two_sum(double a, double b, double &s, double &e) {
	s = a + b;
	const double bb = s - a;
	e = (a - (s - bb)) + (b - bb);
}

inline double host_chain_value(double re, double im, int sel, const double *af) {
	if(sel == 0) return host_prob(re, im, af);
	return sel == 1 ? re : im;
}

// sel: 0 = probabilities of the (optionally affine-mapped) amplitudes, with
// find/rnd locating the first state whose cumulative sum reaches rnd;
// 1 / 2 = real / imaginary parts (the reference's sequential amplitude sums).
ChainResult chain_walk(tmfqs_hip_state *h, int sel, bool find, double rnd, const double *af) {
	sync_all(h);
	const unsigned cl = h->n < kSampleChunkLog ? h->n : kSampleChunkLog;
	const u64 chunk = 1ull << cl;
	const u64 nch = 1ull << (h->n - cl);
	ensure_chunk_buffers(h, nch);
	const double *daf = af ? (const double *)arena_upload(h, af, 4 * sizeof(double)) : nullptr;
	std::vector<double> sums(nch);
	{
		ProfScope ps(h, "k_chunk_sum", (double)h->N * h->esz);
		HIPC(launch_chunk_sum(h->L(), h->amp, h->n, daf, sel, h->d_sums));
	}
	HIPC(hipMemcpyAsync(sums.data(), h->d_sums, nch * sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIPC(hipStreamSynchronize(h->stream));
	// approximate exact prefix (double-double) -> predicted binade per chunk
	std::vector<int32_t> ulog(nch, INT32_MIN);
	double eh = 0, el = 0;
	for(u64 c = 0; c < nch; ++c) {
		const double start = eh + el;
		double s, e;
		two_sum(eh, sums[c], s, e);
		el += e;
		two_sum(s, el, eh, el);
		const double end = eh + el;
		const double a0 = std::fabs(start), a1 = std::fabs(end);
		if(a0 > 0x1p-1000 && a1 > 0x1p-1000 && (start > 0) == (end > 0)) {
			const int e0 = std::ilogb(std::min(a0, a1) * (1.0 - 1e-6));
			const int e1 = std::ilogb(std::max(a0, a1) * (1.0 + 1e-6));
			if(e0 == e1) ulog[c] = e0 - 52;
		}
	}
	std::vector<ChunkTicks> tk(nch);
	HIPC(hipMemcpyAsync(h->d_ulog, ulog.data(), nch * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
	{
		ProfScope ps(h, "k_chunk_ticks", (double)h->N * h->esz);
		HIPC(launch_chunk_ticks(h->L(), h->amp, h->n, daf, sel, h->d_ulog, h->d_ticks));
	}
	HIPC(hipMemcpyAsync(tk.data(), h->d_ticks, nch * sizeof(ChunkTicks), hipMemcpyDeviceToHost, h->stream));
	HIPC(hipStreamSynchronize(h->stream));

	std::vector<double> buf(2 * chunk);
	std::vector<float> fbuf(h->prec == 1 ? 2 * chunk : 0);
	auto load_chunk = [&](u64 c) {
		if(h->prec == 0) {
			HIPC(hipMemcpy(buf.data(), (char *)h->amp + c * chunk * 16, chunk * 16, hipMemcpyDeviceToHost));
		} else {
			HIPC(hipMemcpy(fbuf.data(), (char *)h->amp + c * chunk * 8, chunk * 8, hipMemcpyDeviceToHost));
			for(u64 i = 0; i < 2 * chunk; ++i) buf[i] = fbuf[i];
		}
	};
	// Inside one binade of |S| (ulp u) every partial sum S_i = u * (K + prefix_i)
	// is exact; it is the reference's value as long as all exact partial sums
	// stay at least half an ulp inside the binade, which the integer prefix
	// extremes certify with a one-ulp margin.
	constexpr long long lo = (1ll << 52) + 1, hi = (1ll << 53) - 1;
	ChainResult r;
	double S = 0.0;
	for(u64 c = 0; c < nch; ++c) {
		const ChunkTicks &t = tk[c];
		if(!t.nonzero) {  // every value is +-0: S unchanged
			// the chunk's first state already satisfies rnd <= S (rnd == 0 on
			// a leading run of zero probabilities): the reference stops there
			if(find && rnd <= S) {
				r.found = true;
				r.index = c * chunk;
				r.total = S;
				return r;
			}
			continue;
		}
		bool fast = false;
		if(ulog[c] != INT32_MIN && !t.tie && std::fabs(S) > 0x1p-1000 && std::ilogb(S) - 52 == ulog[c]) {
			const long long K = (long long)std::ldexp(S, -ulog[c]);  // exact integer, |K| < 2^53
			const long long kmin = K + t.pmin, kmax = K + t.pmax, kend = K + t.ticks;
			const bool ok = K > 0 ? (kmin >= lo && kmax <= hi) : (kmax <= -lo && kmin >= -hi);
			if(ok) {
				const double Send = std::ldexp((double)kend, ulog[c]);
				if(!(find && rnd <= Send)) {
					S = Send;
					fast = true;
				}
			}
		}
		if(fast) continue;
		load_chunk(c);
		for(u64 i = 0; i < chunk; ++i) {
			S += host_chain_value(buf[2 * i], buf[2 * i + 1], sel, af);
			if(find && rnd <= S) {
				r.found = true;
				r.index = c * chunk + i;
				r.total = S;
				return r;
			}
		}
	}
	r.total = S;
	return r;
}

// ------------------------------------------------------------ bulk copies
// The two staging slots' completion events, released on every exit path.
struct SlotEvents {
	hipEvent_t ev[2] = {nullptr, nullptr};
	void create() {
		HIPC(hipEventCreate(&ev[0]));
		HIPC(hipEventCreate(&ev[1]));
	}
	~SlotEvents() {
		for(hipEvent_t e : ev)
			if(e) hipEventDestroy(e);
	}
};

void load_host(tmfqs_hip_state *h, const double *src) {
	const size_t stage = 64u << 20;
	ensure_stage(h, 2 * stage);
	const u64 total_states = h->N;
	const u64 per = stage / h->esz;
	SlotEvents slots;
	slots.create();
	hipEvent_t *ev = slots.ev;
	bool used[2] = {false, false};
	u64 s0 = 0;
	int slot = 0;
	while(s0 < total_states) {
		const u64 cnt = std::min<u64>(per, total_states - s0);
		if(used[slot]) HIPC(hipEventSynchronize(ev[slot]));
		char *dst = h->h_stage + slot * stage;
		if(h->prec == 0) {
			std::memcpy(dst, src + 2 * s0, cnt * 16);
		} else {
			float *f = (float *)dst;
			for(u64 i = 0; i < 2 * cnt; ++i) f[i] = (float)src[2 * s0 + i];
		}
		HIPC(hipMemcpyAsync((char *)h->amp + s0 * h->esz, dst, cnt * h->esz, hipMemcpyHostToDevice, h->stream));
		HIPC(hipEventRecord(ev[slot], h->stream));
		used[slot] = true;
		s0 += cnt;
		slot ^= 1;
	}
	HIPC(hipStreamSynchronize(h->stream));
}

// Device -> host through two pinned slots: the copy of chunk i+1 is in
// flight while chunk i is copied (or widened, complex64) into `out`.
void export_host(tmfqs_hip_state *h, u64 first, u64 count, double *out) {
	sync_all(h);  // exported values include every queued exact scatter
	const size_t stage = 64u << 20;
	ensure_stage(h, 2 * stage);
	const u64 per = stage / h->esz;
	SlotEvents slots;
	slots.create();
	auto enqueue = [&](u64 s0, int slot) {
		const u64 cnt = std::min<u64>(per, count - s0);
		HIPC(hipMemcpyAsync(h->h_stage + slot * stage, (char *)h->amp + (first + s0) * h->esz, cnt * h->esz, hipMemcpyDeviceToHost,
		                    h->stream));
		HIPC(hipEventRecord(slots.ev[slot], h->stream));
	};
	if(count) enqueue(0, 0);
	u64 s0 = 0;
	int slot = 0;
	while(s0 < count) {
		const u64 cnt = std::min<u64>(per, count - s0);
		if(s0 + cnt < count) enqueue(s0 + cnt, slot ^ 1);  // that slot's previous chunk is already copied out
		HIPC(hipEventSynchronize(slots.ev[slot]));
		const char *src = h->h_stage + slot * stage;
		if(h->prec == 0) {
			std::memcpy(out + 2 * s0, src, cnt * 16);
		} else {
			const float *f = (const float *)src;
			for(u64 i = 0; i < 2 * cnt; ++i) out[2 * s0 + i] = f[i];
		}
		s0 += cnt;
		slot ^= 1;
	}
}

void device_zero(tmfqs_hip_state *h) { HIPC(hipMemsetAsync(h->amp, 0, h->amp_bytes, h->stream)); }

void device_set(tmfqs_hip_state *h, u64 s, double re, double im) {
	Scatter sc{s, re, im};
	const Scatter *d = (const Scatter *)arena_upload(h, &sc, sizeof(sc));
	HIPC(launch_scatter(h->L(), h->amp, d, 1));
}

tmfqs_hip_state *checked(tmfqs_hip_state *h) {
	if(!h) throw std::invalid_argument("null tmfqs_hip_state handle");
	return h;
}

// One JSON line per handle: every kernel's launches, device time,
// algorithmic bytes, GB/s and the fraction of the 8 TB/s HBM peak.  "1" or
// "stderr" prints it; any other value names a file the line is appended to.
void write_trace(tmfqs_hip_state *h) {
	try {
		collect_profile(h);
	} catch(...) {
		return;
	}
	std::string out = "{\"tmfqs_hip_trace\": 1, \"qubits\": " + std::to_string(h->init ? h->n : 0) + ", \"precision\": \"" +
	                  (h->prec == 0 ? "complex128" : "complex64") + "\", \"hbm_peak_gb_per_s\": 8000, \"kernels\": [";
	double total = 0;
	bool first = true;
	for(const auto &kv : h->prof) {
		const ProfAgg &a = kv.second;
		const double gbs = a.ms > 0 ? a.bytes / (a.ms * 1e6) : 0.0;
		char buf[320];
		snprintf(buf, sizeof buf,
		         "%s{\"name\": \"%s\", \"launches\": %llu, \"ms\": %.6f, \"bytes\": %.0f, \"gb_per_s\": %.1f, "
		         "\"fraction_of_hbm_peak\": %.6f}",
		         first ? "" : ", ", kv.first.c_str(), (unsigned long long)a.launches, a.ms, a.bytes, gbs, gbs / 8000.0);
		out += buf;
		first = false;
		total += a.ms;
	}
	char tail[64];
	snprintf(tail, sizeof tail, "], \"total_ms\": %.6f}\n", total);
	out += tail;
	if(h->trace_path == "1" || h->trace_path == "stderr") {
		fputs(out.c_str(), stderr);
	} else if(FILE *f = fopen(h->trace_path.c_str(), "a")) {
		fputs(out.c_str(), f);
		fclose(f);
	}
}

void destroy_handle(tmfqs_hip_state *h) {
	if(!h) return;
	hipSetDevice(h->cfg.device);
	if(h->stream) hipStreamSynchronize(h->stream);
	if(!h->trace_path.empty() && h->stream) write_trace(h);
	for(auto &r : h->recs) {
		hipEventDestroy(r.start);
		hipEventDestroy(r.stop);
	}
	for(auto &p : h->timed) {
		hipEventDestroy(p.first);
		hipEventDestroy(p.second);
	}
	for(auto e : h->event_pool) hipEventDestroy(e);
	hipFree(h->amp);
	hipFree(h->scratch);
	hipFree(h->d_red);
	hipFree(h->d_phase);
	hipFree(h->d_cin);
	hipFree(h->d_sums);
	hipFree(h->d_ticks);
	hipFree(h->d_ulog);
	for(void *p : h->adj_bufs) hipFree(p);
	hipFree(h->d_gpart);
	hipFree(h->d_grads);
	if(h->arena.host) hipHostFree(h->arena.host);
	if(h->arena.dev) hipFree(h->arena.dev);
	if(h->h_stage) hipHostFree(h->h_stage);
	if(h->stream) hipStreamDestroy(h->stream);
	delete h;
}

}  // namespace

// Helpers other translation units of the library share (engine.h).
namespace tmfqs_hip {
ObsParams observable_params(unsigned n, const tmfqs_hip_diag_term *terms, size_t nterms) {
	return make_observable(n, terms, nterms);
}
void check_device(hipError_t e, const char *what) { hip_check(e, what); }
int status_of_current_exception() { return map_exception(); }
void *window_reset(tmfqs_hip_state *h, hipStream_t *stream) {
	require_init(h, "chunk window");
	HIPC(hipSetDevice(h->cfg.device));
	flush(h);
	h->queue.clear();
	h->qmats.clear();
	h->qscatter.clear();
	h->ew_steps_pending = 0;
	h->tracked.clear();
	h->dirty.clear();
	h->dirty_set.clear();
	*stream = h->stream;
	return h->amp;
}
void *window_flush(tmfqs_hip_state *h, hipStream_t *stream) {
	require_init(h, "chunk window");
	HIPC(hipSetDevice(h->cfg.device));
	flush(h);
	*stream = h->stream;
	return h->amp;
}
}  // namespace tmfqs_hip

// =========================================================================
// extern "C" ABI
static void throw_status(int rc) {
	switch(rc) {
		case TMFQS_HIP_OK: return;
		case TMFQS_HIP_ERR_OUT_OF_RANGE: throw std::out_of_range(g_last_error);
		case TMFQS_HIP_ERR_INVALID_ARGUMENT: throw std::invalid_argument(g_last_error);
		case TMFQS_HIP_ERR_LOGIC: throw std::logic_error(g_last_error);
		case TMFQS_HIP_ERR_OUT_OF_MEMORY: throw std::bad_alloc();
		case TMFQS_HIP_ERR_DEVICE: throw DeviceError(g_last_error);
		default: throw std::runtime_error(g_last_error);
	}
}

#define ABI_BEGIN try {
#define ABI_END                 \
	return TMFQS_HIP_OK;        \
	}                           \
	catch(...) {                \
		return map_exception(); \
	}

extern "C" {

const char *tmfqs_hip_last_error(void) { return g_last_error.c_str(); }

int tmfqs_hip_abi_version(void) { return TMFQS_HIP_ABI_VERSION; }

int tmfqs_hip_device_count(int *count) {
	ABI_BEGIN
	int c = 0;
	if(hipGetDeviceCount(&c) != hipSuccess) c = 0;
	*count = c;
	ABI_END
}

void tmfqs_hip_config_default(tmfqs_hip_config *cfg) {
	if(!cfg) return;
	std::memset(cfg, 0, sizeof(*cfg));
	cfg->precision = TMFQS_HIP_COMPLEX128;
	cfg->device = 0;
	cfg->fusion = 1;
	cfg->exact_phases = 0;
}

int tmfqs_hip_create(const tmfqs_hip_config *cfg, tmfqs_hip_state **out) {
	tmfqs_hip_state *h = nullptr;
	try {
		if(!out) throw std::invalid_argument("tmfqs_hip_create: null output pointer");
		*out = nullptr;
		fault_arm();
		h = new tmfqs_hip_state();
		if(cfg)
			h->cfg = *cfg;
		else
			tmfqs_hip_config_default(&h->cfg);
		if(h->cfg.precision != TMFQS_HIP_COMPLEX128 && h->cfg.precision != TMFQS_HIP_COMPLEX64)
			throw std::invalid_argument("tmfqs_hip_create: unknown precision");
		// a tile of 2^T amplitudes plus its index table must fit the 160 KiB LDS of a gfx950 CU
		const int maxT = h->cfg.precision == TMFQS_HIP_COMPLEX128 ? 13 : 14;
		if(h->cfg.tile_qubits < 0 || h->cfg.tile_qubits > maxT) throw std::invalid_argument("tmfqs_hip_create: tile_qubits out of range");
		if(h->cfg.coalesce_qubits < 0 || h->cfg.coalesce_qubits > 14)
			throw std::invalid_argument("tmfqs_hip_create: coalesce_qubits out of range");
		if(h->cfg.fuse_qubits < 0 || h->cfg.fuse_qubits > 5 || h->cfg.fuse_qubits == 1)
			throw std::invalid_argument("tmfqs_hip_create: fuse_qubits must be 0 or 2..5");
		if(h->cfg.fuse_min_gates < 0) throw std::invalid_argument("tmfqs_hip_create: fuse_min_gates out of range");
		if(h->cfg.qft_plan < 0 || h->cfg.qft_plan > 1) throw std::invalid_argument("tmfqs_hip_create: qft_plan must be 0 or 1");
		h->prec = h->cfg.precision;
		h->esz = h->prec == 0 ? 16 : 8;
		int count = 0;
		if(hipGetDeviceCount(&count) != hipSuccess || count == 0) throw DeviceError("tmfqs_hip_create: no HIP device available");
		if(h->cfg.device < 0 || h->cfg.device >= count) throw std::invalid_argument("tmfqs_hip_create: device ordinal out of range");
		HIPC(hipSetDevice(h->cfg.device));
		HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
		if(const char *tr = getenv("TMFQS_HIP_TRACE"))
			if(*tr && strcmp(tr, "0") != 0) {
				h->trace_path = tr;
				h->profiling = true;
			}
		HIPC(hipMalloc((void **)&h->d_red, (2 * kReduceBlocks + 16) * sizeof(double)));
		build_phase_tables(h);
		arena_ensure(h, 8u << 20);
		*out = h;
		return TMFQS_HIP_OK;
	} catch(...) {
		const int rc = map_exception();
		destroy_handle(h);
		return rc;
	}
}

int tmfqs_hip_destroy(tmfqs_hip_state *h) {
	ABI_BEGIN
	destroy_handle(h);
	ABI_END
}

int tmfqs_hip_clone(tmfqs_hip_state *src, tmfqs_hip_state **out) {
	tmfqs_hip_state *h = nullptr;
	try {
		checked(src);
		HIPC(hipSetDevice(src->cfg.device));
		sync_all(src);
		const int rc = tmfqs_hip_create(&src->cfg, &h);
		if(rc) return rc;
		if(src->init) {
			allocate_state(h, src->n);
			HIPC(hipMemcpyAsync(h->amp, src->amp, src->amp_bytes, hipMemcpyDeviceToDevice, h->stream));
			HIPC(hipStreamSynchronize(h->stream));
			h->tracked = src->tracked;
		}
		*out = h;
		return TMFQS_HIP_OK;
	} catch(...) {
		const int rc = map_exception();
		destroy_handle(h);
		return rc;
	}
}

int tmfqs_hip_qubit_count(tmfqs_hip_state *h, unsigned *out) {
	ABI_BEGIN
	*out = checked(h)->n;
	ABI_END
}

int tmfqs_hip_get_config(tmfqs_hip_state *h, tmfqs_hip_config *out) {
	ABI_BEGIN
	*out = checked(h)->cfg;
	ABI_END
}

int tmfqs_hip_init_zero(tmfqs_hip_state *h, unsigned n) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	allocate_state(h, n);
	device_zero(h);
	ABI_END
}

int tmfqs_hip_init_basis(tmfqs_hip_state *h, unsigned n, uint64_t state, double re, double im) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	allocate_state(h, n);
	device_zero(h);
	validate_state(h, state, "HipStateBackend::initBasis");
	device_set(h, state, re, im);
	ABI_END
}

int tmfqs_hip_init_uniform(tmfqs_hip_state *h, unsigned n, const uint64_t *states, size_t count) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	allocate_state(h, n);
	device_zero(h);
	std::vector<u64> sel;
	for(size_t i = 0; i < count; ++i) {
		validate_state(h, states[i], "HipStateBackend::initUniformSuperposition");
		sel.push_back(states[i]);
	}
	std::sort(sel.begin(), sel.end());
	sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
	if(sel.empty()) throw std::invalid_argument("HipStateBackend::initUniformSuperposition requires at least one state");
	const double amp = 1.0 / std::sqrt((double)sel.size());
	std::vector<Scatter> sc;
	sc.reserve(sel.size());
	for(u64 s : sel) sc.push_back(Scatter{s, amp, 0.0});
	for(size_t s0 = 0; s0 < sc.size(); s0 += 1u << 20) {
		const size_t cnt = std::min<size_t>(1u << 20, sc.size() - s0);
		const Scatter *d = (const Scatter *)arena_upload(h, sc.data() + s0, cnt * sizeof(Scatter));
		HIPC(launch_scatter(h->L(), h->amp, d, (unsigned)cnt));
	}
	ABI_END
}

int tmfqs_hip_load(tmfqs_hip_state *h, unsigned n, const double *amps, uint64_t elem_count) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	if(n > 63u) throw std::invalid_argument("Number of qubits exceeds supported range for state indexing");
	if(elem_count != 2ull * (1ull << n)) throw std::invalid_argument("HipStateBackend: amplitudes size mismatch");
	allocate_state(h, n);
	load_host(h, amps);
	ABI_END
}

// ------------------------------------------------------------ randomPhase
// QuantumState::randomPhase(n, seed) (quantum_state.cpp:264-277) generated
// in device memory.  The seed's 2^(n+1) words are split into P = 2^levels
// streams of L words; stream 0 starts at the seeded state and level i's
// kernel jumps streams [0, 2^i) by L 2^i words into [2^i, 2^(i+1)).  Each
// workgroup then generates its stream's states; values whose double-double
// cos/sin lie near a rounding midpoint are listed, recomputed on the host
// with glibc (the call the reference's compiler emits for its cos/sin pair)
// and patched back.
namespace {

struct DeviceBlock {
	void *p = nullptr;
	explicit DeviceBlock(size_t bytes) { HIPC(hipMalloc(&p, std::max<size_t>(bytes, 256))); }
	~DeviceBlock() {
		if(p) hipFree(p);
	}
	DeviceBlock(const DeviceBlock &) = delete;
	DeviceBlock &operator=(const DeviceBlock &) = delete;
};

// glibc's value of each flagged entry, in place: x -> sin x or cos x.
void glibc_phase_values(PhaseFlag *f, u64 count) {
	auto work = [f](u64 lo, u64 hi) {
		for(u64 i = lo; i < hi; ++i) {
			double sn, cs;
			__builtin_sincos(f[i].x, &sn, &cs);
			f[i].x = (f[i].key & 1u) ? sn : cs;
		}
	};
	const unsigned hw = std::thread::hardware_concurrency();
	const unsigned workers = std::max(1u, std::min(hw ? hw : 1u, 16u));
	if(workers == 1 || count < (1u << 16)) {
		work(0, count);
		return;
	}
	std::vector<std::thread> pool;
	const u64 per = (count + workers - 1) / workers;
	for(u64 lo = 0; lo < count; lo += per) pool.emplace_back(work, lo, std::min(count, lo + per));
	for(auto &t : pool) t.join();
}

// Round trip of `count` flag entries through the two pinned staging slots:
// the copy down of chunk c + 1 runs while the host evaluates chunk c.
void patch_flags_via_host(tmfqs_hip_state *h, PhaseFlag *dflags, u64 count) {
	if(count == 0) return;
	const size_t stage = 64u << 20;
	ensure_stage(h, 2 * stage);
	const u64 per = stage / sizeof(PhaseFlag);
	const u64 chunks = (count + per - 1) / per;
	SlotEvents slots;
	slots.create();
	auto down = [&](u64 c) {
		const u64 k = std::min(per, count - c * per);
		char *slot = h->h_stage + (c & 1) * stage;
		HIPC(hipMemcpyAsync(slot, dflags + c * per, k * sizeof(PhaseFlag), hipMemcpyDeviceToHost, h->stream));
		HIPC(hipEventRecord(slots.ev[c & 1], h->stream));
	};
	down(0);
	for(u64 c = 0; c < chunks; ++c) {
		if(c + 1 < chunks) down(c + 1);  // after the copy up of chunk c - 1 from that slot, in stream order
		HIPC(hipEventSynchronize(slots.ev[c & 1]));
		const u64 k = std::min(per, count - c * per);
		PhaseFlag *slot = (PhaseFlag *)(h->h_stage + (c & 1) * stage);
		glibc_phase_values(slot, k);
		HIPC(hipMemcpyAsync(dflags + c * per, slot, k * sizeof(PhaseFlag), hipMemcpyHostToDevice, h->stream));
	}
	HIPC(hipStreamSynchronize(h->stream));
}

void random_phase_device(tmfqs_hip_state *h, unsigned seed) {
	using tmfqs_phase::kMtN;
	using tmfqs_phase::kPolyWords;
	const u64 words = 2 * h->N;
	const u64 L = words <= (1u << 12) ? words : std::max<u64>(1u << 12, words >> 11);
	const unsigned P = (unsigned)(words / L);
	unsigned logL = 0, levels = 0;
	while((u64{1} << logL) < L) ++logL;
	while((1u << levels) < P) ++levels;
	// Stream states and the jump polynomials of the levels.
	std::vector<uint32_t> st0(kMtN);
	tmfqs_mt::seed_state(seed, st0.data());
	std::vector<uint64_t> polys((size_t)std::max(levels, 1u) * kPolyWords, 0);
	for(unsigned i = 0; i < levels; ++i) {
		const tmfqs_mt::Poly &g = tmfqs_mt::pow2_mod(logL + i);
		// The jump kernel reads x_{j+m} for set bits j < 19937 only.
		if(g.size() != (size_t)kPolyWords || (g.back() >> (tmfqs_phase::kMtDegree - 64 * (kPolyWords - 1))) != 0)
			throw std::logic_error("randomPhase: jump polynomial of unexpected degree");
		std::copy(g.begin(), g.end(), polys.begin() + (size_t)i * kPolyWords);
	}
	DeviceBlock dstates((size_t)P * kMtN * sizeof(uint32_t)), dpoly(polys.size() * sizeof(uint64_t));
	HIPC(hipMemcpy(dstates.p, st0.data(), kMtN * sizeof(uint32_t), hipMemcpyHostToDevice));
	HIPC(hipMemcpy(dpoly.p, polys.data(), polys.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
	for(unsigned i = 0; i < levels; ++i)
		HIPC(launch_mt_jump(h->stream, (uint32_t *)dstates.p, 0, 1u << i, 1u << i, (const uint64_t *)dpoly.p + (size_t)i * kPolyWords));
	// About 9% of values fall in the flag bands: a round of R streams lists at
	// most a quarter of its values (plus slack for tiny states), 8 GiB at most.
	const u64 per_stream = L / 2;
	const u64 slack = 1u << 16, max_entries = u64{1} << 29;
	unsigned R = P;
	while(R > 1 && (u64)R * per_stream / 2 > max_entries) R /= 2;
	const u64 cap = (u64)R * per_stream / 2 + slack;
	DeviceBlock dflags(cap * sizeof(PhaseFlag)), dcount(sizeof(unsigned long long));
	const double norm = 1.0 / std::sqrt((double)h->N);
	const double range = 2.0 * std::acos(-1.0);
	for(unsigned s0 = 0; s0 < P; s0 += R) {
		const unsigned cnt = std::min(R, P - s0);
		HIPC(hipMemsetAsync(dcount.p, 0, sizeof(unsigned long long), h->stream));
		HIPC(launch_random_phase(h->L(), (const uint32_t *)dstates.p + (size_t)s0 * kMtN, s0, cnt, L, norm, range, h->amp,
		                         (PhaseFlag *)dflags.p, (unsigned long long *)dcount.p, cap));
		unsigned long long flagged = 0;
		HIPC(hipMemcpyAsync(&flagged, dcount.p, sizeof(flagged), hipMemcpyDeviceToHost, h->stream));
		HIPC(hipStreamSynchronize(h->stream));
		if(flagged > cap) throw std::runtime_error("randomPhase: more near-midpoint values than the flag list holds");
		patch_flags_via_host(h, (PhaseFlag *)dflags.p, flagged);
		HIPC(launch_phase_patch(h->L(), (const PhaseFlag *)dflags.p, flagged, norm, h->amp));
	}
	HIPC(hipStreamSynchronize(h->stream));
}

}  // namespace

int tmfqs_hip_init_random_phase(tmfqs_hip_state *h, unsigned n, unsigned seed) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	if(n > 63u) throw std::invalid_argument("Number of qubits exceeds supported range for state indexing");
	allocate_state(h, n);
	try {
		random_phase_device(h, seed);
		// fromAmplitudes' default validation: the probability mass, summed in
		// state order, within 1e-9 of one.  The reference checks the FP64
		// vector before storage rounds it; complex64 holds only the rounded
		// values, whose mass lies within a factor (1 +- 2^-24)^2 of it, so
		// the bound widens by that much (it never rejects an accepted vector).
		const double tol = h->prec == 0 ? 1e-9 : 1e-9 + 0x1p-22;
		if(std::abs(chain_walk(h, 0, false, 0.0, nullptr).total - 1.0) > tol)
			throw std::invalid_argument("Amplitude vector is not normalized");
	} catch(...) {
		h->init = false;
		throw;
	}
	ABI_END
}

int tmfqs_hip_mt_jump_outputs(uint32_t seed, uint64_t jump_words, size_t count, uint32_t *out) {
	ABI_BEGIN
	if(count && !out) throw std::invalid_argument("tmfqs_hip_mt_jump_outputs: null output");
	uint32_t st[tmfqs_phase::kMtN];
	tmfqs_mt::seed_state(seed, st);
	for(unsigned m = 0; m < 64; ++m)
		if((jump_words >> m) & 1u) tmfqs_mt::jump_host(st, tmfqs_mt::pow2_mod(m));
	std::vector<uint32_t> x(st, st + tmfqs_phase::kMtN);
	for(size_t k = 0; k < count; ++k) {
		x.push_back(tmfqs_phase::mt_next(x[k], x[k + 1], x[k + tmfqs_phase::kMtM]));
		out[k] = tmfqs_phase::mt_temper(x.back());
	}
	ABI_END
}

int tmfqs_hip_phase_eval(double x, double out[4], int flags[2]) {
	ABI_BEGIN
	if(!out || !flags) throw std::invalid_argument("tmfqs_hip_phase_eval: null output");
	if(!(x >= 0.0 && x < 7.0)) throw std::invalid_argument("tmfqs_hip_phase_eval: x outside [0, 2 pi)");
	tmfqs_phase::DD s, c;
	tmfqs_phase::dd_sincos(x, s, c);
	bool fc, fs;
	tmfqs_phase::round_flagged(c, &fc);
	tmfqs_phase::round_flagged(s, &fs);
	out[0] = c.hi;
	out[1] = c.lo;
	out[2] = s.hi;
	out[3] = s.lo;
	flags[0] = fc;
	flags[1] = fs;
	ABI_END
}

int tmfqs_hip_export(tmfqs_hip_state *h, uint64_t first, uint64_t count, double *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "amplitude export");
	HIPC(hipSetDevice(h->cfg.device));
	if(first > h->N || count > h->N - first) throw std::out_of_range("HipStateBackend::export range out of bounds");
	export_host(h, first, count, out);
	ABI_END
}

int tmfqs_hip_amplitude(tmfqs_hip_state *h, uint64_t state, double *re, double *im) {
	ABI_BEGIN
	checked(h);
	require_init(h, "amplitude query");
	validate_state(h, state, "HipStateBackend::amplitude");
	HIPC(hipSetDevice(h->cfg.device));
	const Cplx v = storage_value(h, current_value(h, state));
	*re = v.re;
	*im = v.im;
	ABI_END
}

int tmfqs_hip_set_amplitude(tmfqs_hip_state *h, uint64_t state, double re, double im) {
	ABI_BEGIN
	checked(h);
	require_init(h, "state update");
	validate_state(h, state, "HipStateBackend::setAmplitude");
	HIPC(hipSetDevice(h->cfg.device));
	const Cplx v = storage_value(h, Cplx{re, im});
	if(!h->cfg.fusion) {
		sync_all(h);
		device_set(h, state, v.re, v.im);
		h->tracked[state] = v;
		return TMFQS_HIP_OK;
	}
	h->tracked[state] = v;
	mark_dirty(h, state);
	QOp op;
	op.kind = Q_SET;
	op.state = state;
	h->queue.push_back(op);
	ABI_END
}

int tmfqs_hip_probability(tmfqs_hip_state *h, uint64_t state, double *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "amplitude query");
	validate_state(h, state, "HipStateBackend::amplitude");
	HIPC(hipSetDevice(h->cfg.device));
	const Cplx v = storage_value(h, current_value(h, state));
	*out = v.re * v.re + v.im * v.im;
	ABI_END
}

int tmfqs_hip_total_probability(tmfqs_hip_state *h, double *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "total probability query");
	HIPC(hipSetDevice(h->cfg.device));
	*out = chain_walk(h, 0, false, 0.0, nullptr).total;
	ABI_END
}

int tmfqs_hip_sample(tmfqs_hip_state *h, double rnd, uint64_t *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "measurement");
	if(rnd < 0.0 || rnd >= 1.0) throw std::invalid_argument("HipStateBackend::sampleMeasurement requires rnd in [0,1)");
	HIPC(hipSetDevice(h->cfg.device));
	const ChainResult r = chain_walk(h, 0, true, rnd, nullptr);
	if(!r.found) throw std::runtime_error("HipStateBackend::sampleMeasurement cumulative probability did not reach sample");
	*out = r.index;
	ABI_END
}

int tmfqs_hip_amplitude_sum(tmfqs_hip_state *h, double *re, double *im) {
	ABI_BEGIN
	checked(h);
	require_init(h, "amplitude chunk iteration");
	HIPC(hipSetDevice(h->cfg.device));
	// the reference sums real and imaginary parts sequentially in state order
	*re = chain_walk(h, 1, false, 0.0, nullptr).total;
	*im = chain_walk(h, 2, false, 0.0, nullptr).total;
	ABI_END
}

int tmfqs_hip_total_probability_affine(tmfqs_hip_state *h, double s_re, double s_im, double b_re, double b_im, double *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "probability summation");
	HIPC(hipSetDevice(h->cfg.device));
	const double af[4] = {s_re, s_im, b_re, b_im};
	*out = chain_walk(h, 0, false, 0.0, af).total;
	ABI_END
}

int tmfqs_hip_sample_affine(tmfqs_hip_state *h, double rnd, double s_re, double s_im, double b_re, double b_im, uint64_t *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "measurement");
	HIPC(hipSetDevice(h->cfg.device));
	const double af[4] = {s_re, s_im, b_re, b_im};
	const ChainResult r = chain_walk(h, 0, true, rnd, af);
	if(!r.found) throw std::runtime_error("QuantumRegister::measure cumulative probability did not reach sample");
	*out = r.index;
	ABI_END
}

int tmfqs_hip_phase_flip(tmfqs_hip_state *h, uint64_t state) {
	ABI_BEGIN
	checked(h);
	require_init(h, "basis-state phase flip");
	validate_state(h, state, "HipStateBackend::phaseFlipBasisState");
	HIPC(hipSetDevice(h->cfg.device));
	Cplx v = current_value(h, state);
	v.re = -v.re;
	v.im = -v.im;
	h->tracked[state] = v;
	if(!h->cfg.fusion) {
		device_set(h, state, v.re, v.im);
		return TMFQS_HIP_OK;
	}
	mark_dirty(h, state);
	QOp op;
	op.kind = Q_FLIP;
	op.state = state;
	h->queue.push_back(op);
	ABI_END
}

int tmfqs_hip_inversion_about_mean(tmfqs_hip_state *h) {
	ABI_BEGIN
	checked(h);
	require_init(h, "inversion about mean");
	double re, im;
	const int rc = tmfqs_hip_amplitude_sum(h, &re, &im);
	if(rc) return rc;
	return tmfqs_hip_inversion_about_mean_value(h, re / (double)h->N, im / (double)h->N);
	ABI_END
}

int tmfqs_hip_inversion_about_mean_value(tmfqs_hip_state *h, double mean_re, double mean_im) {
	ABI_BEGIN
	checked(h);
	require_init(h, "inversion about mean");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_DIFF;
	op.ar = mean_re;
	op.ai = mean_im;
	enqueue_ew(h, op);
	ABI_END
}

int tmfqs_hip_affine(tmfqs_hip_state *h, double s_re, double s_im, double b_re, double b_im) {
	ABI_BEGIN
	checked(h);
	require_init(h, "affine transform");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_AFFINE;
	op.ar = s_re;
	op.ai = s_im;
	op.br = b_re;
	op.bi = b_im;
	enqueue_ew(h, op);
	ABI_END
}

int tmfqs_hip_hadamard(tmfqs_hip_state *h, unsigned qubit) {
	ABI_BEGIN
	checked(h);
	require_init(h, "HipStateBackend::applyHadamard");
	validate_qubit(h, qubit, "HipStateBackend::applyHadamard");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_H;
	op.b0 = pbit(h, qubit);
	enqueue_gate(h, op);
	ABI_END
}

int tmfqs_hip_pauli_x(tmfqs_hip_state *h, unsigned qubit) {
	ABI_BEGIN
	checked(h);
	require_init(h, "HipStateBackend::applyPauliX");
	validate_qubit(h, qubit, "HipStateBackend::applyPauliX");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_X;
	op.b0 = pbit(h, qubit);
	enqueue_gate(h, op);
	ABI_END
}

int tmfqs_hip_controlled_phase(tmfqs_hip_state *h, unsigned control, unsigned target, double theta) {
	ABI_BEGIN
	checked(h);
	require_init(h, "HipStateBackend::applyControlledPhaseShift");
	validate_two(h, control, target, "HipStateBackend::applyControlledPhaseShift");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_CP;
	op.b0 = pbit(h, control);
	op.b1 = pbit(h, target);
	op.theta = theta;
	// dense_state_backend.cpp:296-297 computes cos and sin of theta, which the
	// reference's g++ build turns into one sincos call; glibc's sincos can
	// differ from separate cos/sin calls in the last bit.
	__builtin_sincos(theta, &op.ci, &op.cr);
	enqueue_gate(h, op);
	ABI_END
}

int tmfqs_hip_controlled_not(tmfqs_hip_state *h, unsigned control, unsigned target) {
	ABI_BEGIN
	checked(h);
	require_init(h, "HipStateBackend::applyControlledNot");
	validate_two(h, control, target, "HipStateBackend::applyControlledNot");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_CX;
	op.b0 = pbit(h, control);
	op.b1 = pbit(h, target);
	enqueue_gate(h, op);
	ABI_END
}

int tmfqs_hip_swap(tmfqs_hip_state *h, unsigned q1, unsigned q2) {
	ABI_BEGIN
	checked(h);
	require_init(h, "swap");
	validate_two(h, q1, q2, "QuantumRegister::applySwap");
	HIPC(hipSetDevice(h->cfg.device));
	QOp op;
	op.kind = Q_SWAP;
	op.b0 = pbit(h, q1);
	op.b1 = pbit(h, q2);
	enqueue_gate(h, op);
	ABI_END
}

int tmfqs_hip_apply_gate(tmfqs_hip_state *h, const double *mat, unsigned k, const unsigned *qubits) {
	ABI_BEGIN
	checked(h);
	require_init(h, "gate application");
	const char *scope = "HipStateBackend::applyGate";
	// validation.h:56-78 of the reference
	if(k > 63u) throw std::invalid_argument(std::string(scope) + " qubit list exceeds supported index width");
	if(k > h->n) throw std::invalid_argument(std::string(scope) + " gate acts on more qubits than available");
	std::vector<bool> seen(h->n, false);
	for(unsigned i = 0; i < k; ++i) {
		validate_qubit(h, qubits[i], scope);
		if(seen[qubits[i]]) throw std::invalid_argument(std::string(scope) + " duplicate qubit index");
		seen[qubits[i]] = true;
	}
	if(k > 5) throw std::invalid_argument(std::string(scope) + " supports gates on at most 5 qubits");
	HIPC(hipSetDevice(h->cfg.device));
	const size_t dim = 1ull << k;
	QOp op;
	op.mat = h->qmats.size();
	h->qmats.insert(h->qmats.end(), mat, mat + 2 * dim * dim);
	if(k == 0) {
		// 1x1 matrix: a global scalar, an affine step with zero bias.
		QOp a;
		a.kind = Q_AFFINE;
		a.ar = mat[0];
		a.ai = mat[1];
		h->qmats.resize(op.mat);
		enqueue_ew(h, a);
		return TMFQS_HIP_OK;
	}
	if(k == 1) {
		op.kind = Q_U1;
		op.b0 = pbit(h, qubits[0]);
	} else {
		op.kind = Q_UK;
		op.k = k;
		for(unsigned i = 0; i < k; ++i) op.bits[i] = pbit(h, qubits[i]);
	}
	enqueue_gate(h, op);
	ABI_END
}

int tmfqs_hip_begin_batch(tmfqs_hip_state *h) {
	ABI_BEGIN
	checked(h);
	ABI_END
}

int tmfqs_hip_end_batch(tmfqs_hip_state *h) {
	ABI_BEGIN
	checked(h);
	if(h->init) {
		HIPC(hipSetDevice(h->cfg.device));
		flush(h);
	}
	ABI_END
}

int tmfqs_hip_synchronize(tmfqs_hip_state *h) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	if(h->init) sync_all(h);
	HIPC(hipStreamSynchronize(h->stream));
	ABI_END
}

int tmfqs_hip_profile_enable(tmfqs_hip_state *h, int on) {
	ABI_BEGIN
	checked(h)->profiling = on != 0;
	ABI_END
}

int tmfqs_hip_profile_reset(tmfqs_hip_state *h) {
	ABI_BEGIN
	checked(h);
	collect_profile(h);
	h->prof.clear();
	ABI_END
}

int tmfqs_hip_profile_read(tmfqs_hip_state *h, char (*names)[64], uint64_t *launches, double *ms, double *bytes, size_t *count) {
	ABI_BEGIN
	checked(h);
	collect_profile(h);
	size_t i = 0;
	for(auto &kv : h->prof) {
		if(i >= *count) break;
		std::snprintf(names[i], 64, "%s", kv.first.c_str());
		launches[i] = kv.second.launches;
		ms[i] = kv.second.ms;
		bytes[i] = kv.second.bytes;
		++i;
	}
	*count = i;
	ABI_END
}

int tmfqs_hip_compare(tmfqs_hip_state *ref, tmfqs_hip_state *cand, tmfqs_hip_comparison *out) {
	ABI_BEGIN
	checked(ref);
	checked(cand);
	require_init(ref, "comparison");
	require_init(cand, "comparison");
	if(ref->n != cand->n) throw std::invalid_argument("Compared states must have the same state count");
	if(ref->cfg.device != cand->cfg.device) throw std::invalid_argument("tmfqs_hip_compare: states live on different devices");
	HIPC(hipSetDevice(ref->cfg.device));
	sync_all(cand);
	sync_all(ref);
	CompareBlock *d = nullptr;
	HIPC(hipMalloc((void **)&d, kReduceBlocks * sizeof(CompareBlock)));
	std::vector<CompareBlock> blocks(kReduceBlocks);
	unsigned nb = 0;
	hipError_t e = launch_compare(ref->L(), ref->prec, ref->amp, cand->prec, cand->amp, ref->n, d, &nb);
	if(e == hipSuccess) e = hipMemcpyAsync(blocks.data(), d, nb * sizeof(CompareBlock), hipMemcpyDeviceToHost, ref->stream);
	if(e == hipSuccess) e = hipStreamSynchronize(ref->stream);
	hipFree(d);
	HIPC(e);
	double d2 = 0, r2 = 0, c2 = 0, ovr = 0, ovi = 0, neq = 0, maxd = -1, maxcomp = 0, maxdp = 0;
	u64 worst = 0;
	for(unsigned i = 0; i < nb; ++i) {
		const CompareBlock &b = blocks[i];
		d2 += b.d2;
		r2 += b.r2;
		c2 += b.c2;
		ovr += b.ovr;
		ovi += b.ovi;
		neq += b.neq;
		if(b.maxd > maxd || (b.maxd == maxd && b.worst < worst)) {
			maxd = b.maxd;
			worst = b.worst;
		}
		maxcomp = std::max(maxcomp, b.maxcomp);
		maxdp = std::max(maxdp, b.maxdp);
	}
	out->bitwise_equal = neq == 0;
	out->max_abs_amplitude_error = maxd > 0 ? maxd : 0.0;
	out->max_abs_component_error = maxcomp;
	out->rmse_amplitude = std::sqrt(d2 / (double)ref->N);
	out->rel_l2 = r2 == 0 ? 0.0 : std::sqrt(d2 / r2);
	out->max_abs_probability_error = maxdp;
	out->total_probability_diff = std::abs(c2 - r2);
	out->worst_state = maxd > 0 ? worst : 0;
	out->overlap_re = ovr;
	out->overlap_im = ovi;
	out->fidelity = (r2 > 0 && c2 > 0) ? (ovr * ovr + ovi * ovi) / (r2 * c2) : 0.0;
	ABI_END
}

int tmfqs_hip_copy_probe(tmfqs_hip_state *h, int reps, double *ms) {
	ABI_BEGIN
	checked(h);
	require_init(h, "copy probe");
	if(reps < 1) throw std::invalid_argument("tmfqs_hip_copy_probe: reps must be positive");
	HIPC(hipSetDevice(h->cfg.device));
	sync_all(h);
	ensure_scratch(h);
	hipEvent_t e0 = take_event(h), e1 = take_event(h);
	HIPC(hipEventRecord(e0, h->stream));
	for(int i = 0; i < reps; ++i) {
		ProfScope ps(h, "k_copy", pass_bytes(h));
		HIPC(launch_copy(h->L(), h->amp, h->scratch, h->n));
	}
	HIPC(hipEventRecord(e1, h->stream));
	HIPC(hipEventSynchronize(e1));
	float t = 0;
	HIPC(hipEventElapsedTime(&t, e0, e1));
	h->event_pool.push_back(e0);
	h->event_pool.push_back(e1);
	*ms = t / reps;
	ABI_END
}

int tmfqs_hip_stream(tmfqs_hip_state *h, void **stream) {
	ABI_BEGIN
	*stream = (void *)checked(h)->stream;
	ABI_END
}

int tmfqs_hip_run(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *qubit_pool) {
	ABI_BEGIN
	checked(h);
	require_init(h, "operation list");
	if(count && !ops) throw std::invalid_argument("tmfqs_hip_run: null operation list");
	HIPC(hipSetDevice(h->cfg.device));
	for(size_t i = 0; i < count; ++i) {
		// one record at a time: unfused execution clears the matrix pool after each gate
		const std::vector<QOp> q = convert_ops(h, ops + i, 1, mat_pool, qubit_pool, true, nullptr);
		int rc = TMFQS_HIP_OK;
		switch(ops[i].kind) {
			case TMFQS_HIP_OP_PHASE_FLIP: rc = tmfqs_hip_phase_flip(h, ops[i].state); break;
			case TMFQS_HIP_OP_DIFFUSE_LAZY:
			case TMFQS_HIP_OP_DIFFUSE_MATERIALIZED: rc = tmfqs_hip_inversion_about_mean(h); break;
			default: enqueue_gate(h, q[0]); break;
		}
		throw_status(rc);
	}
	ABI_END
}

int tmfqs_hip_tune(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *qubit_pool,
                   int *tile_qubits, int *fuse_qubits, double *best_ms) {
	ABI_BEGIN
	checked(h);
	require_init(h, "tuning");
	if(count && !ops) throw std::invalid_argument("tmfqs_hip_tune: null operation list");
	if(!tile_qubits || !fuse_qubits) throw std::invalid_argument("tmfqs_hip_tune: null output pointer");
	HIPC(hipSetDevice(h->cfg.device));
	sync_all(h);
	h->qmats.clear();
	const std::vector<QOp> q = convert_ops(h, ops, count, mat_pool, qubit_pool, false, nullptr);
	const size_t qm = h->qmats.size();
	ensure_scratch(h);
	const tmfqs_hip_config saved = h->cfg;
	const int defT = h->prec == 0 ? 12 : 13;
	GateTarget tg;
	tg.a = h->scratch;
	hipEvent_t ev[2] = {nullptr, nullptr};
	HIPC(hipEventCreate(&ev[0]));
	HIPC(hipEventCreate(&ev[1]));
	float best = 0;
	int bestT = saved.tile_qubits, bestF = saved.fuse_qubits;
	bool have = false;
	try {
		// tile width: automatic (default, or one bit narrower when that needs no
		// more passes), or either width forced; composition off or on 3..5 qubits
		for(int T : {0, defT - 1, defT})
			for(int F : {0, 3, 4, 5}) {
				if(T > (int)h->n) continue;
				h->cfg.tile_qubits = T;
				h->cfg.fuse_qubits = F;
				h->cfg.exact_phases = 0;
				float cand = 0;
				for(int rep = 0; rep < 2; ++rep) {
					HIPC(launch_copy(h->L(), h->amp, h->scratch, h->n));
					HIPC(hipEventRecord(ev[0], h->stream));
					exec_gates(h, q, 0, q.size(), &tg);
					HIPC(hipEventRecord(ev[1], h->stream));
					HIPC(hipEventSynchronize(ev[1]));
					float ms = 0;
					HIPC(hipEventElapsedTime(&ms, ev[0], ev[1]));
					cand = rep == 0 ? ms : std::min(cand, ms);
					h->qmats.resize(qm);  // composed blocks of this run
				}
				if(!have || cand < best) {
					best = cand;
					bestT = T;
					bestF = F;
					have = true;
				}
			}
	} catch(...) {
		h->cfg = saved;
		h->qmats.clear();
		hipEventDestroy(ev[0]);
		hipEventDestroy(ev[1]);
		throw;
	}
	hipEventDestroy(ev[0]);
	hipEventDestroy(ev[1]);
	h->cfg = saved;
	h->cfg.tile_qubits = bestT;
	h->cfg.fuse_qubits = saved.exact_phases ? saved.fuse_qubits : bestF;
	h->qmats.clear();
	*tile_qubits = bestT;
	*fuse_qubits = h->cfg.fuse_qubits;
	if(best_ms) *best_ms = best;
	ABI_END
}

int tmfqs_hip_set_tuning(tmfqs_hip_state *h, int tile_qubits, int fuse_qubits) {
	ABI_BEGIN
	checked(h);
	const int maxT = h->prec == 0 ? 13 : 14;
	if(tile_qubits < 0 || tile_qubits > maxT) throw std::invalid_argument("tmfqs_hip_set_tuning: tile_qubits out of range");
	if(fuse_qubits < 0 || fuse_qubits > 5 || fuse_qubits == 1)
		throw std::invalid_argument("tmfqs_hip_set_tuning: fuse_qubits must be 0 or 2..5");
	flush(h);
	h->cfg.tile_qubits = tile_qubits;
	h->cfg.fuse_qubits = fuse_qubits;
	ABI_END
}

int tmfqs_hip_flush_timed(tmfqs_hip_state *h) {
	ABI_BEGIN
	checked(h);
	HIPC(hipSetDevice(h->cfg.device));
	const hipEvent_t a = take_event(h), b = take_event(h);
	h->timed.emplace_back(a, b);
	HIPC(hipEventRecord(a, h->stream));
	flush(h);
	HIPC(hipEventRecord(b, h->stream));
	ABI_END
}

int tmfqs_hip_timed_read(tmfqs_hip_state *h, double *ms, size_t cap, size_t *count) {
	ABI_BEGIN
	checked(h);
	if(!count || (cap && !ms)) throw std::invalid_argument("tmfqs_hip_timed_read: null output");
	HIPC(hipSetDevice(h->cfg.device));
	HIPC(hipStreamSynchronize(h->stream));
	const size_t n = std::min(cap, h->timed.size());
	for(size_t i = 0; i < h->timed.size(); ++i) {
		if(i < n) {
			float t = 0;
			HIPC(hipEventElapsedTime(&t, h->timed[i].first, h->timed[i].second));
			ms[i] = t;
		}
		h->event_pool.push_back(h->timed[i].first);
		h->event_pool.push_back(h->timed[i].second);
	}
	*count = h->timed.size();
	h->timed.clear();
	ABI_END
}

int tmfqs_hip_expectation(tmfqs_hip_state *h, const tmfqs_hip_diag_term *terms, size_t nterms, double *out) {
	ABI_BEGIN
	checked(h);
	require_init(h, "expectation");
	if(!out) throw std::invalid_argument("tmfqs_hip_expectation: null output");
	const ObsParams O = make_observable(h, terms, nterms);
	HIPC(hipSetDevice(h->cfg.device));
	flush(h);
	*out = device_expectation(h, h->amp, nullptr, O);
	ABI_END
}

int tmfqs_hip_adjoint(tmfqs_hip_state *h, const tmfqs_hip_op *ops, size_t count, const double *mat_pool, const unsigned *qubit_pool,
                      const tmfqs_hip_diag_term *terms, size_t nterms, unsigned checkpoint_gates, double *expval, double *grads,
                      size_t grad_count) {
	ABI_BEGIN
	checked(h);
	require_init(h, "adjoint differentiation");
	if(count && !ops) throw std::invalid_argument("tmfqs_hip_adjoint: null operation list");
	if(!expval) throw std::invalid_argument("tmfqs_hip_adjoint: null expectation output");
	const ObsParams O = make_observable(h, terms, nterms);
	HIPC(hipSetDevice(h->cfg.device));
	flush(h);
	try {
		std::vector<size_t> params;
		std::vector<QOp> q = convert_ops(h, ops, count, mat_pool, qubit_pool, false, &params);
		if(params.size() != grad_count)
			throw std::invalid_argument("tmfqs_hip_adjoint: grad_count must equal the number of CPHASE and PHASE operations (" +
			                            std::to_string(params.size()) + ")");
		if(grad_count && !grads) throw std::invalid_argument("tmfqs_hip_adjoint: null gradient output");
		h->tracked.clear();
		adjoint_run(h, q, params, O, checkpoint_gates, expval, grads);
	} catch(...) {
		h->qmats.clear();  // the queue is empty: no queued op refers to the pool
		throw;
	}
	h->qmats.clear();
	ABI_END
}

}  // extern "C"

// ------------------------------------------------------------ batched circuits
// Many independent circuits of uneven size: each is planned like a single
// register's gate run (tiled passes, swaps relabeled), then round r launches
// pass r of every circuit as one persistent work-list kernel.  The rounds,
// the basis-state initialisation and the final permutations are captured into
// one HIP graph, so a batch run is a single graph launch.
struct tmfqs_hip_batch {
	tmfqs_hip_config cfg{};
	int prec = 0;
	size_t esz = 16;
	hipStream_t stream = nullptr;
	std::vector<unsigned> n;
	std::vector<u64> init;
	std::vector<std::vector<QOp>> ops;  // physical bit positions
	char *dmem = nullptr;                // every circuit's two state buffers
	std::vector<size_t> offA, offB;
	unsigned T = 0;                      // tile bits of the current plan
	std::vector<char> in_b;              // result left in the second buffer
	struct Round {
		size_t pass0, item0, nitems;
	};
	std::vector<Round> rounds;
	char *dtab = nullptr;  // device tables of the plan
	BatchPass *d_passes = nullptr;
	BatchItem *d_items = nullptr;
	TileOp *d_ops = nullptr;
	TileStage *d_stages = nullptr;
	BatchChunk *d_init = nullptr, *d_perm = nullptr;
	unsigned ninit = 0, nperm = 0;
	size_t npasses = 0, nitems = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	bool use_graph = true;
	Launch L() const { return Launch{stream, prec}; }
};

namespace {

void batch_free_plan(tmfqs_hip_batch *b) {
	if(b->exec) hipGraphExecDestroy(b->exec);
	if(b->graph) hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
	if(b->dtab) hipFree(b->dtab);
	b->dtab = nullptr;
	b->rounds.clear();
}

// Workgroups of a batch round's work-list kernel (the seed-42 workload:
// 12.8 ms at 8192, 13.4 ms at 2048).
constexpr unsigned kBatchGrid = 8192;

void batch_enqueue(tmfqs_hip_batch *b) {
	HIPC(launch_batch_init(b->L(), b->d_init, b->ninit));
	const unsigned W = tile_stage_width(b->prec, b->T, false);
	for(const auto &r : b->rounds)
		HIPC(launch_batch_tiles(b->L(), W, b->d_passes + r.pass0, b->d_items + r.item0, (unsigned)r.nitems, b->d_ops, b->d_stages,
		                        nullptr, kBatchGrid));
	HIPC(launch_batch_permute(b->L(), b->d_perm, b->nperm));
}

// Plans every circuit with tiles of T bits and builds the device tables and
// the graph of one batch run.
void batch_plan(tmfqs_hip_batch *b, unsigned T) {
	batch_free_plan(b);
	b->T = T;
	const size_t nc = b->n.size();
	std::vector<GatePlan> plans(nc);
	size_t R = 0;
	for(size_t c = 0; c < nc; ++c) {
		PlanCfg pc;
		pc.prec = b->prec;
		pc.dual = false;
		pc.ladders = false;  // one product per controlled phase: bit-identical to the dense backend
		pc.T = std::min(T, b->n[c]);
		pc.C = std::min(pc.T, b->prec == 0 ? 4u : 5u);
		pc.W = tile_stage_width(b->prec, T, false);  // one kernel width for the whole batch
		plans[c] = plan_gates(pc, b->n[c], b->ops[c], 0, b->ops[c].size(), std::vector<double>());
		for(const PassLaunch &l : plans[c].launches)
			if(l.kind != PassLaunch::TILE || l.ldsops || l.W != tile_stage_width(b->prec, T, false))
				throw std::invalid_argument("batched circuit needs a pass the batch kernel does not run");
		R = std::max(R, plans[c].launches.size());
	}
	std::vector<BatchPass> passes;
	std::vector<BatchItem> items;
	std::vector<TileOp> ops;
	std::vector<TileStage> stages;
	for(size_t r = 0; r < R; ++r) {
		tmfqs_hip_batch::Round rd{passes.size(), items.size(), 0};
		for(size_t c = 0; c < nc; ++c) {
			if(r >= plans[c].launches.size()) continue;
			const PassLaunch &l = plans[c].launches[r];
			BatchPass bp{};
			bp.state = b->dmem + b->offA[c];
			bp.tp = l.tp;
			bp.op0 = (uint32_t)ops.size();
			bp.st0 = (uint32_t)stages.size();
			bp.nst = (uint32_t)l.stages.size();
			ops.insert(ops.end(), l.ops.begin(), l.ops.end());
			stages.insert(stages.end(), l.stages.begin(), l.stages.end());
			const uint32_t local = (uint32_t)(passes.size() - rd.pass0);
			passes.push_back(bp);
			const u64 per = 4;  // tiles per work item
			for(u64 t0 = 0; t0 < l.tp.ntiles; t0 += per) {
				BatchItem it{};
				it.pass = local;
				it.tile0 = t0;
				it.ntiles = (uint32_t)std::min<u64>(per, l.tp.ntiles - t0);
				items.push_back(it);
			}
		}
		rd.nitems = items.size() - rd.item0;
		b->rounds.push_back(rd);
	}
	std::vector<BatchChunk> init, perm;
	const u64 chunk = 1ull << 16;
	b->in_b.assign(nc, 0);
	for(size_t c = 0; c < nc; ++c) {
		const u64 N = 1ull << b->n[c];
		for(u64 f = 0; f < N; f += chunk) {
			BatchChunk ch{};
			ch.dst = b->dmem + b->offA[c];
			ch.first = f;
			ch.count = std::min(chunk, N - f);
			ch.init = b->init[c];
			ch.n = b->n[c];
			init.push_back(ch);
		}
		if(plans[c].permute) {
			b->in_b[c] = 1;
			for(u64 f = 0; f < N; f += chunk) {
				BatchChunk ch{};
				ch.src = b->dmem + b->offA[c];
				ch.dst = b->dmem + b->offB[c];
				ch.first = f;
				ch.count = std::min(chunk, N - f);
				ch.n = b->n[c];
				for(unsigned q = 0; q < b->n[c]; ++q) ch.map[q] = (uint8_t)plans[c].map[q];
				perm.push_back(ch);
			}
		}
	}
	// one device allocation for every table
	auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
	const size_t sp = al(passes.size() * sizeof(BatchPass)), si = al(items.size() * sizeof(BatchItem)),
	             so = al(ops.size() * sizeof(TileOp)), ss = al(stages.size() * sizeof(TileStage)),
	             sn = al(init.size() * sizeof(BatchChunk)), sm = al(perm.size() * sizeof(BatchChunk));
	const size_t total = std::max<size_t>(256, sp + si + so + ss + sn + sm);
	HIPC(hipMalloc((void **)&b->dtab, total));
	std::vector<char> host(total, 0);
	size_t off = 0;
	auto put = [&](const void *src, size_t bytes, size_t span) {
		if(bytes) std::memcpy(host.data() + off, src, bytes);
		char *d = b->dtab + off;
		off += span;
		return d;
	};
	b->d_passes = (BatchPass *)put(passes.data(), passes.size() * sizeof(BatchPass), sp);
	b->d_items = (BatchItem *)put(items.data(), items.size() * sizeof(BatchItem), si);
	b->d_ops = (TileOp *)put(ops.data(), ops.size() * sizeof(TileOp), so);
	b->d_stages = (TileStage *)put(stages.data(), stages.size() * sizeof(TileStage), ss);
	b->d_init = (BatchChunk *)put(init.data(), init.size() * sizeof(BatchChunk), sn);
	b->d_perm = (BatchChunk *)put(perm.data(), perm.size() * sizeof(BatchChunk), sm);
	HIPC(hipMemcpy(b->dtab, host.data(), total, hipMemcpyHostToDevice));
	b->ninit = (unsigned)init.size();
	b->nperm = (unsigned)perm.size();
	b->npasses = passes.size();
	b->nitems = items.size();
	// one eager run (kernel attributes are set outside capture), then the graph
	batch_enqueue(b);
	HIPC(hipStreamSynchronize(b->stream));
	if(b->use_graph) {
		HIPC(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
		try {
			batch_enqueue(b);
		} catch(...) {
			hipGraph_t g = nullptr;
			(void)hipStreamEndCapture(b->stream, &g);
			if(g) hipGraphDestroy(g);
			throw;
		}
		HIPC(hipStreamEndCapture(b->stream, &b->graph));
		HIPC(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
	}
}

double batch_run_once(tmfqs_hip_batch *b) {
	hipEvent_t e0, e1;
	HIPC(hipEventCreate(&e0));
	HIPC(hipEventCreate(&e1));
	hipError_t err = hipEventRecord(e0, b->stream);
	if(err == hipSuccess) {
		if(b->exec)
			err = hipGraphLaunch(b->exec, b->stream);
		else {
			try {
				batch_enqueue(b);
			} catch(...) {
				hipEventDestroy(e0);
				hipEventDestroy(e1);
				throw;
			}
		}
	}
	if(err == hipSuccess) err = hipEventRecord(e1, b->stream);
	if(err == hipSuccess) err = hipEventSynchronize(e1);
	float ms = 0;
	if(err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	HIPC(err);
	return ms * 1e-3;
}

void destroy_batch(tmfqs_hip_batch *b) {
	if(!b) return;
	hipSetDevice(b->cfg.device);
	if(b->stream) hipStreamSynchronize(b->stream);
	batch_free_plan(b);
	hipFree(b->dmem);
	if(b->stream) hipStreamDestroy(b->stream);
	delete b;
}

}  // namespace

extern "C" {

int tmfqs_hip_batch_create(const tmfqs_hip_config *cfg, const tmfqs_hip_circuit *circuits, size_t count, tmfqs_hip_batch **out) {
	tmfqs_hip_batch *b = nullptr;
	ABI_BEGIN
	if(!out) throw std::invalid_argument("tmfqs_hip_batch_create: null output");
	*out = nullptr;
	if(count && !circuits) throw std::invalid_argument("tmfqs_hip_batch_create: null circuit list");
	fault_arm();
	b = new tmfqs_hip_batch();
	tmfqs_hip_config_default(&b->cfg);
	if(cfg) b->cfg = *cfg;
	b->prec = b->cfg.precision == TMFQS_HIP_COMPLEX64 ? 1 : 0;
	b->esz = b->prec == 0 ? 16 : 8;
	HIPC(hipSetDevice(b->cfg.device));
	HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
	size_t total = 0;
	for(size_t c = 0; c < count; ++c) {
		const tmfqs_hip_circuit &ci = circuits[c];
		if(ci.qubits < 4 || ci.qubits > 34) throw std::invalid_argument("batched circuits need 4 to 34 qubits");
		if(ci.init_state >> ci.qubits) throw std::out_of_range("batched circuit initial state out of range");
		if(ci.count && !ci.ops) throw std::invalid_argument("batched circuit with a null operation list");
		tmfqs_hip_state tmp;  // validation and conversion context only
		tmp.n = ci.qubits;
		tmp.N = 1ull << ci.qubits;
		std::vector<QOp> q = convert_ops(&tmp, ci.ops, ci.count, nullptr, nullptr, false, nullptr);
		for(const QOp &o : q)
			if(o.kind == Q_U1 || o.kind == Q_UK) throw std::invalid_argument("batched circuits take H, X, CX, CPHASE, PHASE and SWAP");
		b->n.push_back(ci.qubits);
		b->init.push_back(ci.init_state);
		b->ops.push_back(std::move(q));
		const size_t bytes = ((size_t)1 << ci.qubits) * b->esz;
		b->offA.push_back(total);
		total += (bytes + 255) & ~size_t(255);
		b->offB.push_back(total);
		total += (bytes + 255) & ~size_t(255);
	}
	HIPC(hipMalloc((void **)&b->dmem, std::max<size_t>(total, 256)));
	const unsigned maxT = batch_max_tile_bits(b->prec);
	unsigned T = b->cfg.tile_qubits > 0 ? (unsigned)b->cfg.tile_qubits : maxT;
	T = std::max(4u, std::min(T, maxT));
	batch_plan(b, T);
	*out = b;
	b = nullptr;
	return TMFQS_HIP_OK;
	}
	catch(...) {
		destroy_batch(b);
		return map_exception();
	}
}

int tmfqs_hip_batch_tune(tmfqs_hip_batch *b, int *tile_qubits) {
	ABI_BEGIN
	if(!b) throw std::invalid_argument("null tmfqs_hip_batch handle");
	HIPC(hipSetDevice(b->cfg.device));
	if(b->cfg.tile_qubits <= 0) {
		const unsigned maxT = batch_max_tile_bits(b->prec);
		unsigned best = b->T;
		double best_s = 1e300;
		for(unsigned T = maxT - 3; T <= maxT; ++T) {
			batch_plan(b, T);
			batch_run_once(b);
			const double s = std::min(batch_run_once(b), batch_run_once(b));
			if(s < best_s) {
				best_s = s;
				best = T;
			}
		}
		if(best != b->T) batch_plan(b, best);
	}
	if(tile_qubits) *tile_qubits = (int)b->T;
	ABI_END
}

int tmfqs_hip_batch_run(tmfqs_hip_batch *b, double *seconds) {
	ABI_BEGIN
	if(!b) throw std::invalid_argument("null tmfqs_hip_batch handle");
	HIPC(hipSetDevice(b->cfg.device));
	const double s = batch_run_once(b);
	if(seconds) *seconds = s;
	ABI_END
}

int tmfqs_hip_batch_export(tmfqs_hip_batch *b, size_t index, double *out) {
	ABI_BEGIN
	if(!b) throw std::invalid_argument("null tmfqs_hip_batch handle");
	if(index >= b->n.size()) throw std::out_of_range("tmfqs_hip_batch_export: circuit index out of range");
	if(!out) throw std::invalid_argument("tmfqs_hip_batch_export: null output");
	HIPC(hipSetDevice(b->cfg.device));
	HIPC(hipStreamSynchronize(b->stream));
	const u64 N = 1ull << b->n[index];
	const char *src = b->dmem + (b->in_b[index] ? b->offB[index] : b->offA[index]);
	if(b->prec == 0) {
		HIPC(hipMemcpy(out, src, N * 16, hipMemcpyDeviceToHost));
	} else {
		std::vector<float> tmp(2 * N);
		HIPC(hipMemcpy(tmp.data(), src, N * 8, hipMemcpyDeviceToHost));
		for(u64 i = 0; i < 2 * N; ++i) out[i] = tmp[i];
	}
	ABI_END
}

int tmfqs_hip_batch_info(tmfqs_hip_batch *b, tmfqs_hip_batch_stats *out) {
	ABI_BEGIN
	if(!b || !out) throw std::invalid_argument("tmfqs_hip_batch_info: null argument");
	out->circuits = b->n.size();
	out->rounds = b->rounds.size();
	out->passes = b->npasses;
	out->items = b->nitems;
	out->tile_qubits = (int)b->T;
	out->graph = b->exec ? 1 : 0;
	double bytes = 0, pass_bytes = 0;
	for(size_t c = 0; c < b->n.size(); ++c) bytes += (double)(1ull << b->n[c]) * (double)b->esz;
	(void)pass_bytes;
	out->state_bytes = bytes;
	ABI_END
}

int tmfqs_hip_batch_adjoint(tmfqs_hip_batch *b, const tmfqs_hip_diag_term *terms, size_t nterms, double *expvals, double *grads,
                            size_t grad_count) {
	char *dblock = nullptr, *dpsi = nullptr;
	ABI_BEGIN
	if(!b) throw std::invalid_argument("null tmfqs_hip_batch handle");
	if(!expvals) throw std::invalid_argument("tmfqs_hip_batch_adjoint: null expectation output");
	size_t total = 0;
	std::vector<std::vector<size_t>> params(b->n.size());
	for(size_t c = 0; c < b->n.size(); ++c) {
		for(size_t i = 0; i < b->ops[c].size(); ++i)
			if(b->ops[c][i].kind == Q_CP || b->ops[c][i].kind == Q_PH) params[c].push_back(i);
		total += params[c].size();
	}
	if(total != grad_count)
		throw std::invalid_argument("tmfqs_hip_batch_adjoint: grad_count must equal the batch's CPHASE and PHASE operations (" +
		                            std::to_string(total) + ")");
	if(grad_count && !grads) throw std::invalid_argument("tmfqs_hip_batch_adjoint: null gradient output");
	HIPC(hipSetDevice(b->cfg.device));
	HIPC(hipStreamSynchronize(b->stream));
	const size_t C = b->n.size();
	std::vector<size_t> off(C + 1, 0);
	for(size_t c = 0; c < C; ++c) off[c + 1] = off[c] + params[c].size();
	std::vector<ObsParams> obs(C);
	for(size_t c = 0; c < C; ++c) obs[c] = make_observable(b->n[c], terms, nterms);
	// Backward plans: per circuit the inverse gates in reverse order, each
	// parameter's gradient reduction (TOP_GRAD) just before its gate is
	// undone, as dual tiled passes over (psi, lambda) one bit narrower than
	// the forward tiles.  Round r runs pass r of every circuit.
	// One more buffer per circuit, when the device has room, takes a copy of
	// the final state for the backward sweep; otherwise psi is stepped back in
	// place and a second batch run restores the final states.
	std::vector<size_t> poff(C, 0);
	size_t psi_bytes = 0;
	for(size_t c = 0; c < C; ++c) {
		poff[c] = psi_bytes;
		psi_bytes += ((((size_t)1 << b->n[c]) * b->esz) + 255) & ~size_t(255);
	}
	{
		size_t freeb = 0, totalb = 0;
		HIPC(hipMemGetInfo(&freeb, &totalb));
		if(psi_bytes + (256u << 20) < freeb && hipMalloc((void **)&dpsi, std::max<size_t>(psi_bytes, 256)) != hipSuccess) {
			(void)hipGetLastError();
			dpsi = nullptr;
		}
	}
	const unsigned Td = std::max(3u, b->T) - 1u;
	const unsigned Wd = tile_stage_width(b->prec, Td, true);
	std::vector<GatePlan> plans(C);
	size_t R = 0;
	for(size_t c = 0; c < C; ++c) {
		if(params[c].empty()) continue;
		std::vector<int64_t> slot(b->ops[c].size(), -1);
		for(size_t p = 0; p < params[c].size(); ++p) slot[params[c][p]] = (int64_t)p;
		std::vector<QOp> rev;
		size_t keep = 0;
		for(size_t i = b->ops[c].size(); i-- > 0;) {
			const QOp &op = b->ops[c][i];
			if(slot[i] >= 0) {
				QOp g;
				g.kind = Q_GRAD;
				g.b0 = op.b0;
				g.b1 = op.b1;
				g.k = op.kind == Q_CP ? 2u : 1u;
				g.state = (u64)slot[i];
				rev.push_back(g);
				keep = rev.size();
			}
			QOp inv = op;
			if(op.kind == Q_CP || op.kind == Q_PH) {
				inv.ci = -op.ci;
				inv.theta = -op.theta;
			}
			rev.push_back(inv);
		}
		rev.resize(keep);  // nothing before the first parameter needs undoing
		PlanCfg pc;
		pc.prec = b->prec;
		pc.dual = true;
		pc.ladders = false;
		pc.T = std::min(Td, b->n[c]);
		pc.C = std::min(pc.T, b->prec == 0 ? 4u : 5u);
		pc.W = Wd;
		plans[c] = plan_gates(pc, b->n[c], rev, 0, rev.size(), std::vector<double>());
		for(const PassLaunch &l : plans[c].launches)
			if(l.kind != PassLaunch::TILE || l.ldsops || l.W != Wd)
				throw std::invalid_argument("batched adjoint needs a pass the batch kernel does not run");
		R = std::max(R, plans[c].launches.size());
	}
	struct DualRound {
		size_t pass0, item0, nitems, grad0, ngrads;
	};
	std::vector<DualRound> rounds;
	std::vector<BatchPass> passes;
	std::vector<BatchItem> items;
	std::vector<BatchGrad> gents;
	std::vector<TileOp> tops;
	std::vector<TileStage> tstages;
	size_t max_items = 1;
	for(size_t r = 0; r < R; ++r) {
		DualRound rd{passes.size(), items.size(), 0, gents.size(), 0};
		for(size_t c = 0; c < C; ++c) {
			if(r >= plans[c].launches.size()) continue;
			const PassLaunch &l = plans[c].launches[r];
			BatchPass bp{};
			char *A = b->dmem + b->offA[c], *B = b->dmem + b->offB[c];
			// psi: the copy, or the final state itself; lambda: the circuit's other buffer
			bp.state = dpsi ? dpsi + poff[c] : (b->in_b[c] ? B : A);
			bp.state2 = b->in_b[c] ? A : B;
			bp.tp = l.tp;
			bp.op0 = (uint32_t)tops.size();
			bp.st0 = (uint32_t)tstages.size();
			bp.nst = (uint32_t)l.stages.size();
			bp.nslots = (uint32_t)l.gidx.size();
			tops.insert(tops.end(), l.ops.begin(), l.ops.end());
			tstages.insert(tstages.end(), l.stages.begin(), l.stages.end());
			const uint32_t local = (uint32_t)(passes.size() - rd.pass0);
			passes.push_back(bp);
			const uint32_t first = (uint32_t)(items.size() - rd.item0);
			const u64 per = 4;  // tiles per work item
			for(u64 t0 = 0; t0 < l.tp.ntiles; t0 += per) {
				BatchItem it{};
				it.pass = local;
				it.tile0 = t0;
				it.ntiles = (uint32_t)std::min<u64>(per, l.tp.ntiles - t0);
				items.push_back(it);
			}
			const uint32_t cnt = (uint32_t)(items.size() - rd.item0) - first;
			for(size_t sl = 0; sl < l.gidx.size(); ++sl) {
				BatchGrad g{};
				g.item0 = first;
				g.nitems = cnt;
				g.slot = (uint32_t)sl;
				g.gindex = off[c] + l.gidx[sl];
				gents.push_back(g);
			}
		}
		rd.nitems = items.size() - rd.item0;
		rd.ngrads = gents.size() - rd.grad0;
		max_items = std::max(max_items, rd.nitems);
		rounds.push_back(rd);
	}
	// one device block: outputs, reduction partials, gradient partials, tables
	auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
	const size_t so = al((total + C + 1) * sizeof(double)), sr = al(2 * kReduceBlocks * sizeof(double)),
	             sg = al(max_items * kMaxGradSlots * sizeof(double)), sp = al(passes.size() * sizeof(BatchPass)),
	             si = al(items.size() * sizeof(BatchItem)), se = al(gents.size() * sizeof(BatchGrad)),
	             sops = al(tops.size() * sizeof(TileOp)), sst = al(tstages.size() * sizeof(TileStage));
	HIPC(hipMalloc((void **)&dblock, so + sr + sg + sp + si + se + sops + sst));
	double *d_out = (double *)dblock, *d_red = (double *)(dblock + so), *d_gpart = (double *)(dblock + so + sr);
	char *tab = dblock + so + sr + sg;
	std::vector<char> host(sp + si + se + sops + sst, 0);
	size_t hoff = 0;
	auto put = [&](const void *src, size_t bytes, size_t span) {
		if(bytes) std::memcpy(host.data() + hoff, src, bytes);
		char *d = tab + hoff;
		hoff += span;
		return d;
	};
	const BatchPass *d_passes = (const BatchPass *)put(passes.data(), passes.size() * sizeof(BatchPass), sp);
	const BatchItem *d_items = (const BatchItem *)put(items.data(), items.size() * sizeof(BatchItem), si);
	const BatchGrad *d_gents = (const BatchGrad *)put(gents.data(), gents.size() * sizeof(BatchGrad), se);
	const TileOp *d_ops = (const TileOp *)put(tops.data(), tops.size() * sizeof(TileOp), sops);
	const TileStage *d_st = (const TileStage *)put(tstages.data(), tstages.size() * sizeof(TileStage), sst);
	if(!host.empty()) HIPC(hipMemcpyAsync(tab, host.data(), host.size(), hipMemcpyHostToDevice, b->stream));
	// forward (the batch run), then lambda = O psi and E per circuit
	if(b->exec)
		HIPC(hipGraphLaunch(b->exec, b->stream));
	else
		batch_enqueue(b);
	for(size_t c = 0; c < C; ++c) {
		char *A = b->dmem + b->offA[c], *B = b->dmem + b->offB[c];
		HIPC(launch_observable(b->L(), b->in_b[c] ? B : A, b->in_b[c] ? A : B, b->n[c], obs[c], d_red, d_out + total + c));
	}
	if(dpsi)  // the backward sweep steps copies of the final states; the states stay in place
		for(size_t c = 0; c < C; ++c) {
			const char *fin = b->dmem + (b->in_b[c] ? b->offB[c] : b->offA[c]);
			HIPC(hipMemcpyAsync(dpsi + poff[c], fin, ((size_t)1 << b->n[c]) * b->esz, hipMemcpyDeviceToDevice, b->stream));
		}
	for(const DualRound &rd : rounds) {
		HIPC(launch_batch_dual(b->L(), Wd, d_passes + rd.pass0, d_items + rd.item0, (unsigned)rd.nitems, d_ops, d_st, d_gpart,
		                       kBatchGrid));
		HIPC(launch_batch_grad_finish(b->L(), d_gents + rd.grad0, (unsigned)rd.ngrads, d_gpart, -2.0, d_out));
	}
	if(total) HIPC(hipMemcpyAsync(grads, d_out, total * sizeof(double), hipMemcpyDeviceToHost, b->stream));
	HIPC(hipMemcpyAsync(expvals, d_out + total, C * sizeof(double), hipMemcpyDeviceToHost, b->stream));
	if(!dpsi) {  // psi was stepped back in place: leave every final state in place again
		if(b->exec)
			HIPC(hipGraphLaunch(b->exec, b->stream));
		else
			batch_enqueue(b);
	}
	HIPC(hipStreamSynchronize(b->stream));
	hipFree(dblock);
	hipFree(dpsi);
	return TMFQS_HIP_OK;
	}
	catch(...) {
		if(b && b->stream) hipStreamSynchronize(b->stream);
		hipFree(dblock);
		hipFree(dpsi);
		return map_exception();
	}
}

int tmfqs_hip_batch_destroy(tmfqs_hip_batch *b) {
	ABI_BEGIN
	destroy_batch(b);
	ABI_END
}

}  // extern "C"
