// Chunked host-resident states: 2^n amplitudes in 2^(n-c) chunks of 2^c in
// pinned host memory (the home storage), streamed through device windows of
// 2^g chunks that run the engine's kernels.
//
// The reference's chunked storage (CompressedStateBackend,
// include/tmfqs/storage/common/compressed_state_backend.h:103-747) keeps every
// chunk compressed, decompresses chunks into an LRU cache
// (chunk_cache.h:83-122) and applies each gate as a pair kernel inside one
// chunk or across two (pair_kernel_executor.h:26-103).  Here the chunks are
// stored uncompressed and the device-resident cache is a window of 2^g whole
// chunks: an operation list is cut into segments that mix at most g chunk
// bits, and one sweep over the windows applies a whole segment to every
// amplitude, so a gate list costs one host round trip per segment rather than
// per gate.
//
// Index bits follow the engine: qubit q is bit n-1-q of the state index, and
// chunk bits are the bits at positions >= c.  A segment's window holds the
// chunk bits S (its gates' mixing chunk bits, filled up with the lowest other
// chunk bits to g of them); window bit c+i is S[i], window bits below c are
// the chunk's own.  The remaining chunk bits are fixed over a window.
#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using namespace tmfqs_hip;

struct tmfqs_hip_chunked {
	tmfqs_hip_config cfg{};
	unsigned n = 0, c = 0, g = 0;  // qubits, chunk qubits, window chunk bits (g = min(window_bits, n - c))
	size_t esz = 16;
	char *home = nullptr;  // pinned: 2^n amplitudes in the windows' storage format
	tmfqs_hip_state *win[2] = {nullptr, nullptr};
	hipEvent_t uploaded[2] = {nullptr, nullptr};  // per window handle: its last upload is complete
	tmfqs_hip_chunked_stats stats{};
};

namespace {

// A status an inner C-ABI call returned (its message is already recorded).
struct InnerStatus {
	int rc;
};

void inner(int rc) {
	if(rc != TMFQS_HIP_OK) throw InnerStatus{rc};
}

void checked(const tmfqs_hip_chunked *c) {
	if(!c) throw std::invalid_argument("chunked state: null handle");
}

bool is_diffusion(int kind) { return kind == TMFQS_HIP_OP_DIFFUSE_LAZY || kind == TMFQS_HIP_OP_DIFFUSE_MATERIALIZED; }

// Index positions of the chunk bits a record mixes (targets; CX's control is
// a condition, phases are diagonal).
void mixing_chunk_bits(const tmfqs_hip_chunked *C, const tmfqs_hip_op &op, const unsigned *qubit_pool, std::vector<unsigned> &out) {
	out.clear();
	auto add = [&](unsigned q) {
		if(q >= C->n) throw std::out_of_range("chunked state: qubit index out of range");
		const unsigned p = C->n - 1u - q;
		if(p >= C->c) out.push_back(p);
	};
	switch(op.kind) {
		case TMFQS_HIP_OP_H:
		case TMFQS_HIP_OP_X: add(op.q0); break;
		case TMFQS_HIP_OP_CX: add(op.q1); break;
		case TMFQS_HIP_OP_SWAP:
			add(op.q0);
			add(op.q1);
			break;
		case TMFQS_HIP_OP_MATRIX:
			if(!qubit_pool) throw std::invalid_argument("chunked state: MATRIX record without a qubit pool");
			for(unsigned i = 0; i < op.k; ++i) add(qubit_pool[op.q0 + i]);
			break;
		default: break;
	}
	std::sort(out.begin(), out.end());
	out.erase(std::unique(out.begin(), out.end()), out.end());
}

// The chunk bits of one segment's windows.
struct WindowMap {
	unsigned n = 0, c = 0, nw = 0;
	std::vector<unsigned> S;     // window chunk bits, ascending (window bit c + i)
	std::vector<unsigned> rest;  // the other chunk bits, ascending (window index bit k)
	std::vector<int> wbit;       // per index bit: its window bit, or -1 when fixed over a window
	std::vector<int> restk;      // per index bit: its position in rest, or -1
	bool contiguous = false;     // S = {c, .., c+g-1}: a window is one run of the home storage

	WindowMap(const tmfqs_hip_chunked *C, const std::vector<unsigned> &need) : n(C->n), c(C->c), nw(C->c + C->g) {
		S = need;
		for(unsigned p = c; p < n && S.size() < C->g; ++p)
			if(std::find(S.begin(), S.end(), p) == S.end()) S.push_back(p);
		std::sort(S.begin(), S.end());
		wbit.assign(n, -1);
		restk.assign(n, -1);
		for(unsigned p = 0; p < c; ++p) wbit[p] = (int)p;
		for(size_t i = 0; i < S.size(); ++i) wbit[S[i]] = (int)(c + i);
		for(unsigned p = c; p < n; ++p)
			if(wbit[p] < 0) {
				restk[p] = (int)rest.size();
				rest.push_back(p);
			}
		contiguous = true;
		for(size_t i = 0; i < S.size(); ++i) contiguous = contiguous && S[i] == c + i;
	}
	u64 windows() const { return 1ull << rest.size(); }
	unsigned wq(unsigned q) const { return nw - 1u - (unsigned)wbit[n - 1u - q]; }  // window qubit of an in-window qubit
	bool inside(unsigned q) const { return wbit[n - 1u - q] >= 0; }
	unsigned fixed(unsigned q, u64 w) const { return (unsigned)(w >> restk[n - 1u - q]) & 1u; }
	// home chunk of window w's local chunk ci
	u64 chunk(u64 w, u64 ci) const {
		u64 gc = 0;
		for(size_t i = 0; i < S.size(); ++i) gc |= ((ci >> i) & 1ull) << (S[i] - c);
		for(size_t k = 0; k < rest.size(); ++k) gc |= ((w >> k) & 1ull) << (rest[k] - c);
		return gc;
	}
	// window index of global state s, or false when s lies in another window
	bool local_state(u64 s, u64 w, u64 &ws) const {
		ws = s & ((1ull << c) - 1ull);
		for(size_t i = 0; i < S.size(); ++i) ws |= ((s >> S[i]) & 1ull) << (c + i);
		for(size_t k = 0; k < rest.size(); ++k)
			if(((s >> rest[k]) & 1ull) != ((w >> k) & 1ull)) return false;
		return true;
	}
};

tmfqs_hip_op make_op(int kind, unsigned q0, unsigned q1 = 0, double theta = 0.0, uint64_t state = 0) {
	tmfqs_hip_op o;
	std::memset(&o, 0, sizeof(o));
	o.kind = kind;
	o.q0 = q0;
	o.q1 = q1;
	o.theta = theta;
	o.state = state;
	return o;
}

// e^{i theta} on every amplitude of the window: PHASE on a qubit, X, PHASE,
// X multiplies each amplitude once (X moves amplitudes without arithmetic).
void global_phase(std::vector<tmfqs_hip_op> &out, unsigned nw, double theta) {
	const unsigned q = nw - 1u;
	out.push_back(make_op(TMFQS_HIP_OP_PHASE, q, 0, theta));
	out.push_back(make_op(TMFQS_HIP_OP_X, q));
	out.push_back(make_op(TMFQS_HIP_OP_PHASE, q, 0, theta));
	out.push_back(make_op(TMFQS_HIP_OP_X, q));
}

// Window w's records of a segment; MATRIX records take their window qubits
// from wpool at the offsets pool_at gives.
void window_ops(const WindowMap &M, u64 w, const tmfqs_hip_op *ops, size_t lo, size_t hi, const std::vector<size_t> &pool_at,
                std::vector<tmfqs_hip_op> &out) {
	out.clear();
	for(size_t i = lo; i < hi; ++i) {
		const tmfqs_hip_op &op = ops[i];
		switch(op.kind) {
			case TMFQS_HIP_OP_H:
			case TMFQS_HIP_OP_X: out.push_back(make_op(op.kind, M.wq(op.q0))); break;
			case TMFQS_HIP_OP_CX:
				if(M.inside(op.q0))
					out.push_back(make_op(TMFQS_HIP_OP_CX, M.wq(op.q0), M.wq(op.q1)));
				else if(M.fixed(op.q0, w))
					out.push_back(make_op(TMFQS_HIP_OP_X, M.wq(op.q1)));
				break;
			case TMFQS_HIP_OP_SWAP: out.push_back(make_op(TMFQS_HIP_OP_SWAP, M.wq(op.q0), M.wq(op.q1))); break;
			case TMFQS_HIP_OP_CPHASE:
			case TMFQS_HIP_OP_PHASE: {
				const unsigned nq = op.kind == TMFQS_HIP_OP_CPHASE ? 2u : 1u;
				const unsigned qs[2] = {op.q0, op.q1};
				unsigned in[2], nin = 0;
				bool off = false;
				for(unsigned j = 0; j < nq; ++j) {
					if(M.inside(qs[j]))
						in[nin++] = M.wq(qs[j]);
					else if(!M.fixed(qs[j], w))
						off = true;
				}
				if(off) break;
				if(nin == 2)
					out.push_back(make_op(TMFQS_HIP_OP_CPHASE, in[0], in[1], op.theta));
				else if(nin == 1)
					out.push_back(make_op(TMFQS_HIP_OP_PHASE, in[0], 0, op.theta));
				else
					global_phase(out, M.nw, op.theta);
				break;
			}
			case TMFQS_HIP_OP_MATRIX: {
				tmfqs_hip_op o = op;
				o.q0 = (uint32_t)pool_at[i - lo];
				out.push_back(o);
				break;
			}
			case TMFQS_HIP_OP_PHASE_FLIP: {
				u64 ws;
				if(M.local_state(op.state, w, ws)) out.push_back(make_op(TMFQS_HIP_OP_PHASE_FLIP, 0, 0, 0.0, ws));
				break;
			}
			default: throw std::invalid_argument("chunked state: unknown operation kind");
		}
	}
}

// Copies window w between the home storage and a window's device buffer.
void move_window(tmfqs_hip_chunked *C, const WindowMap &M, u64 w, void *dev, hipStream_t s, bool to_device) {
	const size_t cb = ((size_t)1 << C->c) * C->esz;
	const u64 nloc = 1ull << C->g;
	auto copy = [&](char *devp, char *hostp, size_t bytes) {
		if(to_device)
			check_device(hipMemcpyAsync(devp, hostp, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync (chunk to device)");
		else
			check_device(hipMemcpyAsync(hostp, devp, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync (chunk to host)");
	};
	if(M.contiguous)
		copy((char *)dev, C->home + (size_t)M.chunk(w, 0) * cb, cb * nloc);
	else
		for(u64 ci = 0; ci < nloc; ++ci) copy((char *)dev + ci * cb, C->home + (size_t)M.chunk(w, ci) * cb, cb);
	(to_device ? C->stats.bytes_h2d : C->stats.bytes_d2h) += (double)cb * (double)nloc;
}

void sync_windows(tmfqs_hip_chunked *C) {
	for(tmfqs_hip_state *h : C->win) {
		hipStream_t s;
		window_flush(h, &s);
		check_device(hipStreamSynchronize(s), "hipStreamSynchronize (chunk window)");
	}
}

// One sweep over the windows of chunk bits `need`; per window: load, body
// (may enqueue work on the window's handle), and write back when `store`.
// Windows alternate between the two handles, and each upload waits for the
// previous window's upload: the streams run half a window apart, so one
// window's download overlaps the next one's upload on the host link.
template <typename Body>
void sweep(tmfqs_hip_chunked *C, const std::vector<unsigned> &need, bool store, Body body) {
	const WindowMap M(C, need);
	for(u64 w = 0; w < M.windows(); ++w) {
		tmfqs_hip_state *h = C->win[w & 1u];
		hipStream_t s;
		void *dev = window_reset(h, &s);
		if(w) check_device(hipStreamWaitEvent(s, C->uploaded[(w - 1) & 1u], 0), "hipStreamWaitEvent (chunk upload)");
		move_window(C, M, w, dev, s, true);
		check_device(hipEventRecord(C->uploaded[w & 1u], s), "hipEventRecord (chunk upload)");
		body(M, w, h);
		if(store) {
			dev = window_flush(h, &s);
			move_window(C, M, w, dev, s, false);
		}
		++C->stats.windows;
	}
	++C->stats.segments;
	sync_windows(C);
}

void run_segment(tmfqs_hip_chunked *C, const tmfqs_hip_op *ops, size_t lo, size_t hi, const std::vector<unsigned> &need,
                 const double *mat_pool, const unsigned *qubit_pool) {
	if(lo == hi) return;
	// window qubits of the segment's MATRIX records (the same in every window)
	const WindowMap M0(C, need);
	std::vector<unsigned> wpool;
	std::vector<size_t> pool_at(hi - lo, 0);
	for(size_t i = lo; i < hi; ++i)
		if(ops[i].kind == TMFQS_HIP_OP_MATRIX) {
			pool_at[i - lo] = wpool.size();
			for(unsigned j = 0; j < ops[i].k; ++j) wpool.push_back(M0.wq(qubit_pool[ops[i].q0 + j]));
		}
	std::vector<tmfqs_hip_op> wops;
	sweep(C, need, true, [&](const WindowMap &M, u64 w, tmfqs_hip_state *h) {
		window_ops(M, w, ops, lo, hi, pool_at, wops);
		if(!wops.empty()) inner(tmfqs_hip_run(h, wops.data(), wops.size(), mat_pool, wpool.empty() ? nullptr : wpool.data()));
	});
}

// 2 mean - a over the whole register, with the mean of every window's sum.
void diffuse(tmfqs_hip_chunked *C) {
	double re = 0.0, im = 0.0;
	sweep(C, {}, false, [&](const WindowMap &, u64, tmfqs_hip_state *h) {
		double wr, wi;
		inner(tmfqs_hip_amplitude_sum(h, &wr, &wi));
		re += wr;
		im += wi;
	});
	const double N = std::ldexp(1.0, (int)C->n);
	const double mr = re / N, mi = im / N;
	sweep(C, {}, true, [&](const WindowMap &, u64, tmfqs_hip_state *h) { inner(tmfqs_hip_inversion_about_mean_value(h, mr, mi)); });
}

bool only_flips(const tmfqs_hip_op *ops, size_t lo, size_t hi) {
	for(size_t i = lo; i < hi; ++i)
		if(ops[i].kind != TMFQS_HIP_OP_PHASE_FLIP) return false;
	return true;
}

// A segment of phase flips only (Grover's oracle between two diffusions)
// negates single amplitudes: done where they are stored, exactly as a
// window would, without a sweep over the register.
void flip_at_home(tmfqs_hip_chunked *C, const tmfqs_hip_op *ops, size_t lo, size_t hi) {
	sync_windows(C);
	for(size_t i = lo; i < hi; ++i) {
		const u64 s = ops[i].state;
		if(C->esz == 16) {
			double *a = (double *)C->home + 2 * s;
			a[0] = -a[0];
			a[1] = -a[1];
		} else {
			float *a = (float *)C->home + 2 * s;
			a[0] = -a[0];
			a[1] = -a[1];
		}
	}
}

void destroy(tmfqs_hip_chunked *C) {
	if(!C) return;
	for(tmfqs_hip_state *&h : C->win)
		if(h) {
			tmfqs_hip_destroy(h);
			h = nullptr;
		}
	for(hipEvent_t &e : C->uploaded)
		if(e) {
			hipEventDestroy(e);
			e = nullptr;
		}
	if(C->home) hipHostFree(C->home);
	delete C;
}

}  // namespace

#define CHUNKED_BEGIN try {
#define CHUNKED_END                            \
	return TMFQS_HIP_OK;                       \
	}                                          \
	catch(const InnerStatus &e) {              \
		return e.rc;                           \
	}                                          \
	catch(...) {                               \
		return status_of_current_exception(); \
	}

extern "C" {

int tmfqs_hip_chunked_create(const tmfqs_hip_config *cfg, unsigned num_qubits, unsigned chunk_qubits, unsigned window_bits,
                             tmfqs_hip_chunked **out) {
	tmfqs_hip_chunked *C = nullptr;
	try {
		if(!out) throw std::invalid_argument("tmfqs_hip_chunked_create: null output pointer");
		*out = nullptr;
		if(num_qubits < 1 || num_qubits > 40) throw std::invalid_argument("tmfqs_hip_chunked_create: num_qubits must be 1..40");
		if(chunk_qubits < 1 || chunk_qubits > num_qubits)
			throw std::invalid_argument("tmfqs_hip_chunked_create: chunk_qubits must be 1..num_qubits");
		if(chunk_qubits < num_qubits && (window_bits < 1 || chunk_qubits + window_bits > num_qubits))
			throw std::invalid_argument("tmfqs_hip_chunked_create: window_bits must be 1..num_qubits - chunk_qubits");
		C = new tmfqs_hip_chunked();
		if(cfg)
			C->cfg = *cfg;
		else
			tmfqs_hip_config_default(&C->cfg);
		C->n = num_qubits;
		C->c = chunk_qubits;
		C->g = chunk_qubits < num_qubits ? window_bits : 0u;
		for(tmfqs_hip_state *&h : C->win) {
			inner(tmfqs_hip_create(&C->cfg, &h));
			inner(tmfqs_hip_init_zero(h, C->c + C->g));
		}
		C->esz = C->cfg.precision == TMFQS_HIP_COMPLEX128 ? 16 : 8;
		const size_t bytes = ((size_t)1 << num_qubits) * C->esz;
		check_device(hipSetDevice(C->cfg.device), "hipSetDevice");
		for(hipEvent_t &e : C->uploaded) check_device(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
		check_device(hipHostMalloc((void **)&C->home, bytes, hipHostMallocDefault), "hipHostMalloc (chunk storage)");
		std::memset(C->home, 0, bytes);
		if(C->esz == 16)
			((double *)C->home)[0] = 1.0;
		else
			((float *)C->home)[0] = 1.0f;
		*out = C;
		return TMFQS_HIP_OK;
	} catch(const InnerStatus &e) {
		destroy(C);
		return e.rc;
	} catch(...) {
		const int rc = status_of_current_exception();
		destroy(C);
		return rc;
	}
}

int tmfqs_hip_chunked_destroy(tmfqs_hip_chunked *c) {
	if(!c) return TMFQS_HIP_OK;
	int rc = TMFQS_HIP_OK;
	try {
		sync_windows(c);
	} catch(const InnerStatus &e) {
		rc = e.rc;
	} catch(...) {
		rc = status_of_current_exception();
	}
	destroy(c);  // released whatever the last copies reported
	return rc;
}

int tmfqs_hip_chunked_init_basis(tmfqs_hip_chunked *c, uint64_t state, double re, double im) {
	CHUNKED_BEGIN
	checked(c);
	if(state >> c->n) throw std::out_of_range("HipChunkedState::initBasis state index out of range");
	sync_windows(c);
	std::memset(c->home, 0, ((size_t)1 << c->n) * c->esz);
	if(c->esz == 16) {
		((double *)c->home)[2 * state] = re;
		((double *)c->home)[2 * state + 1] = im;
	} else {
		((float *)c->home)[2 * state] = (float)re;
		((float *)c->home)[2 * state + 1] = (float)im;
	}
	CHUNKED_END
}

int tmfqs_hip_chunked_load(tmfqs_hip_chunked *c, const double *interleaved, uint64_t elem_count) {
	CHUNKED_BEGIN
	checked(c);
	if(elem_count != 2ull << c->n) throw std::invalid_argument("HipChunkedState: amplitudes size mismatch");
	if(!interleaved) throw std::invalid_argument("HipChunkedState: null amplitude buffer");
	sync_windows(c);
	if(c->esz == 16)
		std::memcpy(c->home, interleaved, (size_t)elem_count * sizeof(double));
	else
		for(u64 i = 0; i < elem_count; ++i) ((float *)c->home)[i] = (float)interleaved[i];
	CHUNKED_END
}

int tmfqs_hip_chunked_export(tmfqs_hip_chunked *c, uint64_t first, uint64_t count, double *out) {
	CHUNKED_BEGIN
	checked(c);
	if(first > (1ull << c->n) || count > (1ull << c->n) - first) throw std::out_of_range("HipChunkedState::export range out of range");
	if(count && !out) throw std::invalid_argument("HipChunkedState: null output buffer");
	sync_windows(c);
	if(c->esz == 16)
		std::memcpy(out, (const double *)c->home + 2 * first, (size_t)count * 2 * sizeof(double));
	else
		for(u64 i = 0; i < 2 * count; ++i) out[i] = (double)((const float *)c->home)[2 * first + i];
	CHUNKED_END
}

int tmfqs_hip_chunked_run(tmfqs_hip_chunked *c, const tmfqs_hip_op *ops, size_t count, const double *mat_pool,
                          const unsigned *qubit_pool) {
	CHUNKED_BEGIN
	checked(c);
	if(count && !ops) throw std::invalid_argument("tmfqs_hip_chunked_run: null operation list");
	// plan every segment before touching the state
	struct Seg {
		size_t lo, hi;
		std::vector<unsigned> need;
		bool diffusion;
	};
	std::vector<Seg> segs;
	std::vector<unsigned> need, bits;
	size_t lo = 0;
	for(size_t i = 0; i < count; ++i) {
		const int k = ops[i].kind;
		if(is_diffusion(k)) {
			segs.push_back({lo, i, need, false});
			segs.push_back({i, i + 1, {}, true});
			need.clear();
			lo = i + 1;
			continue;
		}
		if(k < TMFQS_HIP_OP_H || k > TMFQS_HIP_OP_PHASE) throw std::invalid_argument("tmfqs_hip_chunked_run: unknown operation kind");
		if(k == TMFQS_HIP_OP_PHASE_FLIP && (ops[i].state >> c->n))
			throw std::out_of_range("HipChunkedState::phaseFlipBasisState state index out of range");
		if(k == TMFQS_HIP_OP_CX || k == TMFQS_HIP_OP_CPHASE || k == TMFQS_HIP_OP_SWAP) {
			if(ops[i].q0 >= c->n || ops[i].q1 >= c->n) throw std::out_of_range("chunked state: qubit index out of range");
			if(ops[i].q0 == ops[i].q1) throw std::invalid_argument("chunked state: two-qubit operation requires distinct qubits");
		}
		if((k == TMFQS_HIP_OP_PHASE || k == TMFQS_HIP_OP_H || k == TMFQS_HIP_OP_X) && ops[i].q0 >= c->n)
			throw std::out_of_range("chunked state: qubit index out of range");
		if(k == TMFQS_HIP_OP_MATRIX) {
			if(ops[i].k < 1 || ops[i].k > 5 || !mat_pool || !qubit_pool)
				throw std::invalid_argument("chunked state: MATRIX records take 1..5 qubits and a matrix pool");
			for(unsigned a = 0; a < ops[i].k; ++a)
				for(unsigned b = a + 1; b < ops[i].k; ++b)
					if(qubit_pool[ops[i].q0 + a] == qubit_pool[ops[i].q0 + b])
						throw std::invalid_argument("chunked state: MATRIX record with repeated qubits");
		}
		mixing_chunk_bits(c, ops[i], qubit_pool, bits);
		if(bits.size() > c->g)
			throw std::invalid_argument("chunked state: a record mixes " + std::to_string(bits.size()) + " chunk bits, the window holds " +
			                            std::to_string(c->g));
		std::vector<unsigned> merged = need;
		merged.insert(merged.end(), bits.begin(), bits.end());
		std::sort(merged.begin(), merged.end());
		merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
		if(merged.size() > c->g) {
			segs.push_back({lo, i, need, false});
			lo = i;
			merged = bits;
		}
		need = merged;
	}
	segs.push_back({lo, count, need, false});
	check_device(hipSetDevice(c->cfg.device), "hipSetDevice");
	for(const Seg &s : segs) {
		if(s.diffusion)
			diffuse(c);
		else if(only_flips(ops, s.lo, s.hi))
			flip_at_home(c, ops, s.lo, s.hi);
		else
			run_segment(c, ops, s.lo, s.hi, s.need, mat_pool, qubit_pool);
	}
	CHUNKED_END
}

int tmfqs_hip_chunked_total_probability(tmfqs_hip_chunked *c, double *out) {
	CHUNKED_BEGIN
	checked(c);
	if(!out) throw std::invalid_argument("tmfqs_hip_chunked_total_probability: null output pointer");
	check_device(hipSetDevice(c->cfg.device), "hipSetDevice");
	double total = 0.0;
	sweep(c, {}, false, [&](const WindowMap &, u64, tmfqs_hip_state *h) {
		double p;
		inner(tmfqs_hip_total_probability(h, &p));
		total += p;
	});
	*out = total;
	CHUNKED_END
}

int tmfqs_hip_chunked_clone(tmfqs_hip_chunked *c, tmfqs_hip_chunked **out) {
	tmfqs_hip_chunked *d = nullptr;
	try {
		checked(c);
		if(!out) throw std::invalid_argument("tmfqs_hip_chunked_clone: null output pointer");
		*out = nullptr;
		inner(tmfqs_hip_chunked_create(&c->cfg, c->n, c->c, c->g ? c->g : 1u, &d));
		sync_windows(c);
		std::memcpy(d->home, c->home, ((size_t)1 << c->n) * c->esz);
		*out = d;
		return TMFQS_HIP_OK;
	} catch(const InnerStatus &e) {
		destroy(d);
		return e.rc;
	} catch(...) {
		const int rc = status_of_current_exception();
		destroy(d);
		return rc;
	}
}

int tmfqs_hip_chunked_set_amplitude(tmfqs_hip_chunked *c, uint64_t state, double re, double im) {
	CHUNKED_BEGIN
	checked(c);
	if(state >> c->n) throw std::out_of_range("HipChunkedState::setAmplitude state index out of range");
	sync_windows(c);
	if(c->esz == 16) {
		((double *)c->home)[2 * state] = re;
		((double *)c->home)[2 * state + 1] = im;
	} else {
		((float *)c->home)[2 * state] = (float)re;
		((float *)c->home)[2 * state + 1] = (float)im;
	}
	CHUNKED_END
}

int tmfqs_hip_chunked_affine(tmfqs_hip_chunked *c, double s_re, double s_im, double b_re, double b_im) {
	CHUNKED_BEGIN
	checked(c);
	check_device(hipSetDevice(c->cfg.device), "hipSetDevice");
	sweep(c, {}, true, [&](const WindowMap &, u64, tmfqs_hip_state *h) { inner(tmfqs_hip_affine(h, s_re, s_im, b_re, b_im)); });
	CHUNKED_END
}

int tmfqs_hip_chunked_inversion_about_mean_value(tmfqs_hip_chunked *c, double mean_re, double mean_im) {
	CHUNKED_BEGIN
	checked(c);
	check_device(hipSetDevice(c->cfg.device), "hipSetDevice");
	sweep(c, {}, true, [&](const WindowMap &, u64, tmfqs_hip_state *h) { inner(tmfqs_hip_inversion_about_mean_value(h, mean_re, mean_im)); });
	CHUNKED_END
}

int tmfqs_hip_chunked_amplitude_sum(tmfqs_hip_chunked *c, double *re, double *im) {
	CHUNKED_BEGIN
	checked(c);
	if(!re || !im) throw std::invalid_argument("tmfqs_hip_chunked_amplitude_sum: null output pointer");
	check_device(hipSetDevice(c->cfg.device), "hipSetDevice");
	double sr = 0.0, si = 0.0;
	sweep(c, {}, false, [&](const WindowMap &, u64, tmfqs_hip_state *h) {
		double wr, wi;
		inner(tmfqs_hip_amplitude_sum(h, &wr, &wi));
		sr += wr;
		si += wi;
	});
	*re = sr;
	*im = si;
	CHUNKED_END
}

int tmfqs_hip_chunked_sample(tmfqs_hip_chunked *c, double rnd, uint64_t *out) {
	CHUNKED_BEGIN
	checked(c);
	if(!out) throw std::invalid_argument("tmfqs_hip_chunked_sample: null output pointer");
	if(rnd < 0.0 || rnd >= 1.0) throw std::invalid_argument("HipChunkedState::sampleMeasurement requires rnd in [0,1)");
	check_device(hipSetDevice(c->cfg.device), "hipSetDevice");
	// windows of the lowest chunk bits cover consecutive state ranges, so the
	// cumulative walk runs window by window in state order
	const WindowMap M(c, {});
	const unsigned wbits = c->c + c->g;
	double before = 0.0;
	bool found = false;
	for(u64 w = 0; w < M.windows() && !found; ++w) {
		tmfqs_hip_state *h = c->win[0];
		hipStream_t s;
		void *dev = window_reset(h, &s);
		move_window(c, M, w, dev, s, true);
		++c->stats.windows;
		double total = 0.0;
		inner(tmfqs_hip_total_probability(h, &total));
		if(rnd - before <= total) {
			uint64_t idx = 0;
			if(tmfqs_hip_sample(h, rnd - before, &idx) == TMFQS_HIP_OK) {
				*out = (w << wbits) + idx;
				found = true;
			}
		}
		before += total;
	}
	++c->stats.segments;
	sync_windows(c);
	if(!found) throw std::runtime_error("HipChunkedState::sampleMeasurement cumulative probability did not reach sample");
	CHUNKED_END
}

int tmfqs_hip_chunked_info(tmfqs_hip_chunked *c, tmfqs_hip_chunked_stats *out) {
	CHUNKED_BEGIN
	checked(c);
	if(!out) throw std::invalid_argument("tmfqs_hip_chunked_info: null output pointer");
	*out = c->stats;
	CHUNKED_END
}

}  // extern "C"
