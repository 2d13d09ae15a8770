"""ctypes binding of include/tmfqs_sim.h: QuantumRegister / sim::Simulator
with the Hip storage strategy, as a C++ caller of the reference API would use
them."""

import ctypes
import os

import numpy as np

from . import LIB_DIR, Config, DiagTerm, HipState, HipError, Op, OP_CPHASE, _EXC, default_config, make_terms, lib as hip_lib

SIM_LIB = os.path.join(LIB_DIR, "libtmfqs.so")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SIM_LIB):
            raise OSError(f"{SIM_LIB} is missing: run python -m tmfqsfullstate_amd.build")
        hip_lib()  # libtmfqs.so resolves its dependency from the same directory
        L = ctypes.CDLL(SIM_LIB)
        L.tmfqs_sim_last_error.restype = ctypes.c_char_p
        L.tmfqs_reg_create_basis.argtypes = [ctypes.c_uint, ctypes.c_uint64, ctypes.POINTER(Config), ctypes.POINTER(_vp)]
        L.tmfqs_reg_create_amplitudes.argtypes = [ctypes.c_uint, _dp, ctypes.POINTER(Config), ctypes.POINTER(_vp)]
        L.tmfqs_reg_create_random_phase.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(Config), ctypes.POINTER(_vp)]
        L.tmfqs_reg_create_chunked.argtypes = [ctypes.c_uint, ctypes.c_uint64, ctypes.POINTER(Config), ctypes.c_uint64,
                                               ctypes.c_uint64, ctypes.POINTER(_vp)]
        L.tmfqs_reg_clone.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.tmfqs_reg_free.argtypes = [_vp]
        L.tmfqs_reg_run.argtypes = [_vp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint),
                                    ctypes.c_int, _dp]
        L.tmfqs_sim_run.argtypes = [ctypes.c_uint, ctypes.c_uint64, _dp, ctypes.POINTER(Config), ctypes.POINTER(Op),
                                    ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint), _dp, ctypes.POINTER(_vp)]
        L.tmfqs_reg_qubit_count.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint)]
        L.tmfqs_reg_export.argtypes = [_vp, _dp]
        L.tmfqs_reg_amplitude.argtypes = [_vp, ctypes.c_uint64, _dp, _dp]
        L.tmfqs_reg_total_probability.argtypes = [_vp, _dp]
        L.tmfqs_reg_measure_seed.argtypes = [_vp, ctypes.c_uint32, _u64p]
        L.tmfqs_reg_measure_fixed.argtypes = [_vp, ctypes.c_double, _u64p]
        L.tmfqs_reg_hip_handle.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.tmfqs_circuit_qft.argtypes = [ctypes.c_uint, ctypes.POINTER(Op), ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.tmfqs_circuit_grover.argtypes = [ctypes.c_uint, _u64p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(Op),
                                           ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.tmfqs_circuit_ring.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(Op), ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_size_t)]
        L.tmfqs_reg_adjoint.argtypes = [_vp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint),
                                        ctypes.POINTER(DiagTerm), ctypes.c_size_t, ctypes.c_uint, _dp, _dp, ctypes.c_size_t,
                                        _dp]
        L.tmfqs_reg_expectation.argtypes = [_vp, ctypes.POINTER(DiagTerm), ctypes.c_size_t, _dp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().tmfqs_sim_last_error().decode(errors="replace")
        raise _EXC.get(rc, HipError)(rc, msg)


def _cfg(precision=0, **kw):
    return default_config(precision=precision, **kw)


def qft_ops(n):
    cnt = ctypes.c_size_t()
    check(lib().tmfqs_circuit_qft(n, None, 0, ctypes.byref(cnt)))
    arr = (Op * max(1, cnt.value))()
    check(lib().tmfqs_circuit_qft(n, arr, cnt.value, ctypes.byref(cnt)))
    return arr, cnt.value


def grover_ops(n, marked, materialized):
    m = np.ascontiguousarray(marked, dtype=np.uint64)
    cnt = ctypes.c_size_t()
    check(lib().tmfqs_circuit_grover(n, m.ctypes.data_as(_u64p), len(m), int(materialized), None, 0, ctypes.byref(cnt)))
    arr = (Op * max(1, cnt.value))()
    check(lib().tmfqs_circuit_grover(n, m.ctypes.data_as(_u64p), len(m), int(materialized), arr, cnt.value,
                                     ctypes.byref(cnt)))
    return arr, cnt.value


def ring_ops(n, layers, seed):
    """circuit::makeRingLayers(n, layers, seed) as op records."""
    cnt = ctypes.c_size_t()
    check(lib().tmfqs_circuit_ring(n, layers, seed, None, 0, ctypes.byref(cnt)))
    arr = (Op * max(1, cnt.value))()
    check(lib().tmfqs_circuit_ring(n, layers, seed, arr, cnt.value, ctypes.byref(cnt)))
    return arr, cnt.value


def make_ops(seq):
    arr = (Op * max(1, len(seq)))()
    for i, t in enumerate(seq):
        arr[i].kind = t[0]
        arr[i].q0 = t[1] if len(t) > 1 else 0
        arr[i].q1 = t[2] if len(t) > 2 else 0
        arr[i].theta = t[3] if len(t) > 3 else 0.0
        arr[i].state = t[4] if len(t) > 4 else 0
        arr[i].k = t[5] if len(t) > 5 else 0
    return arr, len(seq)


class Register:
    """tmfqs::QuantumRegister on the Hip strategy."""

    def __init__(self, handle, n):
        self._h = handle
        self.n = n

    @classmethod
    def basis(cls, n, state=0, precision=0, **cfg):
        h = _vp()
        c = _cfg(precision, **cfg)
        check(lib().tmfqs_reg_create_basis(n, state, ctypes.byref(c), ctypes.byref(h)))
        return cls(h, n)

    @classmethod
    def from_amplitudes(cls, n, amps, precision=0, **cfg):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        h = _vp()
        c = _cfg(precision, **cfg)
        check(lib().tmfqs_reg_create_amplitudes(n, a.ctypes.data_as(_dp), ctypes.byref(c), ctypes.byref(h)))
        return cls(h, n)

    @classmethod
    def chunked(cls, n, state=0, chunk_states=0, window_chunks=0, precision=0, **cfg):
        """A register on StorageStrategyKind::HipChunked: 2^n amplitudes in host
        chunks of chunk_states, run through device windows of window_chunks."""
        h = _vp()
        c = _cfg(precision, **cfg)
        check(lib().tmfqs_reg_create_chunked(n, state, ctypes.byref(c), chunk_states, window_chunks, ctypes.byref(h)))
        return cls(h, n)

    @classmethod
    def random_phase(cls, n, seed, precision=0, **cfg):
        h = _vp()
        c = _cfg(precision, **cfg)
        check(lib().tmfqs_reg_create_random_phase(n, seed, ctypes.byref(c), ctypes.byref(h)))
        return cls(h, n)

    def clone(self):
        h = _vp()
        check(lib().tmfqs_reg_clone(self._h, ctypes.byref(h)))
        return Register(h, self.n)

    def close(self):
        if self._h:
            lib().tmfqs_reg_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, ops, count=None, mat_pool=None, mat_qubits=None, trace=False):
        """Simulator::run's timed region on this register; returns executionSeconds."""
        if count is None:
            ops, count = make_ops(ops)
        secs = ctypes.c_double()
        mp = None if mat_pool is None else np.ascontiguousarray(mat_pool, dtype=np.float64).ctypes.data_as(_dp)
        mq = None
        if mat_qubits is not None:
            mqa = np.ascontiguousarray(mat_qubits, dtype=np.uint32)
            mq = mqa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
        check(lib().tmfqs_reg_run(self._h, ops, count, mp, mq, int(trace), ctypes.byref(secs)))
        return secs.value

    def run_traced(self, ops, count=None, mat_pool=None, mat_qubits=None):
        """Simulator::run with the per-operation trace; returns (per-op seconds, executionSeconds)."""
        if count is None:
            ops, count = make_ops(ops)
        per = np.zeros(max(1, count))
        secs = ctypes.c_double()
        mp = None if mat_pool is None else np.ascontiguousarray(mat_pool, dtype=np.float64).ctypes.data_as(_dp)
        mq = None
        if mat_qubits is not None:
            mqa = np.ascontiguousarray(mat_qubits, dtype=np.uint32)
            mq = mqa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
        L = lib()
        L.tmfqs_reg_run_traced.argtypes = [_vp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint), _dp,
                                           _dp]
        check(L.tmfqs_reg_run_traced(self._h, ops, count, mp, mq, per.ctypes.data_as(_dp), ctypes.byref(secs)))
        return per[:count], secs.value

    def export(self):
        out = np.empty(2 << self.n, dtype=np.float64)
        check(lib().tmfqs_reg_export(self._h, out.ctypes.data_as(_dp)))
        return out

    def amplitudes(self):
        e = self.export()
        return e[0::2] + 1j * e[1::2]

    def amplitude(self, s):
        re, im = ctypes.c_double(), ctypes.c_double()
        check(lib().tmfqs_reg_amplitude(self._h, s, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def total_probability(self):
        o = ctypes.c_double()
        check(lib().tmfqs_reg_total_probability(self._h, ctypes.byref(o)))
        return o.value

    def measure_seed(self, seed):
        o = ctypes.c_uint64()
        check(lib().tmfqs_reg_measure_seed(self._h, seed, ctypes.byref(o)))
        return o.value

    def measure_fixed(self, rnd):
        o = ctypes.c_uint64()
        check(lib().tmfqs_reg_measure_fixed(self._h, rnd, ctypes.byref(o)))
        return o.value

    def adjoint(self, ops, count, terms, checkpoint_gates=0, mat_pool=None, mat_qubits=None):
        """sim::adjointGradient: returns (E, gradients per CPHASE op, executionSeconds)."""
        nparams = sum(1 for i in range(count) if ops[i].kind == OP_CPHASE)
        grads = np.zeros(max(1, nparams))
        arr, cnt = make_terms(terms)
        mp = None if mat_pool is None else np.ascontiguousarray(mat_pool, dtype=np.float64)
        mq = None if mat_qubits is None else np.ascontiguousarray(mat_qubits, dtype=np.uint32)
        e, secs = ctypes.c_double(), ctypes.c_double()
        check(lib().tmfqs_reg_adjoint(self._h, ops, count, None if mp is None else mp.ctypes.data_as(_dp),
                                      None if mq is None else mq.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), arr, cnt,
                                      checkpoint_gates, ctypes.byref(e), grads.ctypes.data_as(_dp), nparams,
                                      ctypes.byref(secs)))
        return e.value, grads[:nparams], secs.value

    def expectation(self, terms):
        arr, cnt = make_terms(terms)
        out = ctypes.c_double()
        check(lib().tmfqs_reg_expectation(self._h, arr, cnt, ctypes.byref(out)))
        return out.value

    def hip_state(self):
        """The device handle behind the register, as a non-owning HipState."""
        h = _vp()
        check(lib().tmfqs_reg_hip_handle(self._h, ctypes.byref(h)))
        s = HipState.__new__(HipState)
        s._h = h
        s.n = self.n
        s.close = lambda: None  # owned by the register
        s._owner = self  # keeps the register (and its device state) alive while the view exists
        return s


def simulate(n, ops, count=None, init=None, basis=0, precision=0, mat_pool=None, mat_qubits=None, **cfg):
    """sim::Simulator::run from basis(n, basis) or fromAmplitudes(init);
    returns (Register, executionSeconds)."""
    if count is None:
        ops, count = make_ops(ops)
    c = _cfg(precision, **cfg)
    h = _vp()
    secs = ctypes.c_double()
    ip = None
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float64)
        ip = init.ctypes.data_as(_dp)
    mp = None if mat_pool is None else np.ascontiguousarray(mat_pool, dtype=np.float64).ctypes.data_as(_dp)
    mq = None
    if mat_qubits is not None:
        mqa = np.ascontiguousarray(mat_qubits, dtype=np.uint32)
        mq = mqa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
    check(lib().tmfqs_sim_run(n, basis, ip, ctypes.byref(c), ops, count, mp, mq, ctypes.byref(secs), ctypes.byref(h)))
    return Register(h, n), secs.value


def compare_runs(n, ops, count=None, ref=None, cand=None, init=None, basis=0, mat_pool=None, mat_qubits=None):
    """experiment::compareStates of the circuit simulated with config `ref` and
    config `cand` (dicts of tmfqs_hip_config fields); returns the metrics."""
    from . import Comparison
    if count is None:
        ops, count = make_ops(ops)
    cr, cc = _cfg(**(ref or {})), _cfg(**(cand or {}))
    ip = None
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float64)
        ip = init.ctypes.data_as(_dp)
    mp = None if mat_pool is None else np.ascontiguousarray(mat_pool, dtype=np.float64).ctypes.data_as(_dp)
    mq = None
    if mat_qubits is not None:
        mqa = np.ascontiguousarray(mat_qubits, dtype=np.uint32)
        mq = mqa.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
    out = Comparison()
    L = lib()
    L.tmfqs_sim_compare.argtypes = [ctypes.c_uint, ctypes.c_uint64, _dp, ctypes.POINTER(Config), ctypes.POINTER(Config),
                                    ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint),
                                    ctypes.POINTER(Comparison)]
    check(L.tmfqs_sim_compare(n, basis, ip, ctypes.byref(cr), ctypes.byref(cc), ops, count, mp, mq, ctypes.byref(out)))
    return out.as_dict()


def compare_amplitudes(n, ref, cand):
    """experiment::compareStates of two host amplitude vectors."""
    from . import Comparison
    a = np.ascontiguousarray(ref, dtype=np.float64)
    b = np.ascontiguousarray(cand, dtype=np.float64)
    out = Comparison()
    L = lib()
    L.tmfqs_compare_amplitudes.argtypes = [ctypes.c_uint, _dp, _dp, ctypes.POINTER(Comparison)]
    check(L.tmfqs_compare_amplitudes(n, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), ctypes.byref(out)))
    return out.as_dict()


def random_phase(n, seed):
    """QuantumState::randomPhase(n, seed) amplitudes (host; interleaved)."""
    out = np.empty(2 << n, dtype=np.float64)
    L = lib()
    L.tmfqs_state_random_phase.argtypes = [ctypes.c_uint, ctypes.c_uint, _dp]
    check(L.tmfqs_state_random_phase(n, seed, out.ctypes.data_as(_dp)))
    return out
