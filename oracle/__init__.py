"""CPU oracle for the dense state-vector path (TEST INFRASTRUCTURE ONLY).

Wraps two shared libraries built by ``oracle/Makefile``:

* ``liboracle.so`` -- the clean-room C restatement in ``dense_oracle.c``;
* ``_ref/libtmfqs_ref.so`` -- the reference's own dense path compiled from its
  sources, plus ``ref_driver.cpp`` (present only where it was built).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``tmfqsfullstate_amd``) never imports it.

Pins (what checks each oracle function; the tests in ``tests/test_oracle_*.py``):

* RNG (``mt_unit_draws``, ``random_phase``): the C++ standard's 10000th-draw
  value of ``std::mt19937``, the reference build's ``Mt19937RandomSource`` and
  ``QuantumState::randomPhase`` (bit-exact), committed golden vectors.
* gate kernels (H, X, CX, CPHASE, SWAP, k-qubit matrices): brute-force
  Kronecker-product matrix-vector products in numpy; invariants (H^2 = I,
  CX^2 = I, SWAP = 3 CX, norm preservation); the reference build (bit-exact).
* QFT: the closed-form DFT ``numpy.fft.ifft(a, norm="ortho")``; the reference
  known-answer tests QFT|0> and QFT|5> at n = 3; the reference build.
* Grover: the closed form ``sin^2((2k+1) asin sqrt(M/N))``; the reference
  build (bit-exact amplitudes and measured index); golden vectors.
* register semantics (cached sum, lazy affine overlay, measure): the
  reference build and its 8-amplitude fixed vector test.
* ``expectation`` (diagonal observables) and ``ring_ops`` (the adjoint
  benchmark circuit; no reference counterpart): closed forms (one layer at
  theta = 0 gives |+>^n, two give |0>), a numpy restatement of the
  observable, and the gradient checks below agreeing with each other.
* ``param_shift_grad`` / ``central_fd_grad``: the exact two-term shift rule
  for generators with eigenvalues {0, 1} and FP64 central differences agree
  to O(h^2) on random circuits.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libtmfqs_ref.so")

OK, OUT_OF_RANGE, INVALID_ARGUMENT, LOGIC_ERROR, RUNTIME_ERROR = 0, -1, -2, -3, -4

OP_H, OP_X, OP_CX, OP_CPHASE, OP_SWAP, OP_PHASE_FLIP, OP_DIFFUSE_LAZY, OP_DIFFUSE_MAT, OP_MATRIX, OP_PHASE = range(10)
DIAG_Z, DIAG_PROJECTOR = 0, 1


class DiagTerm(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
        ("coef", ctypes.c_double),
        ("bits", ctypes.c_uint64),
    ]


class Op(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("q0", ctypes.c_uint32),
        ("q1", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("theta", ctypes.c_double),
        ("state", ctypes.c_uint64),
    ]


class OracleError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"oracle error {code} {what}")
        self.code = code


_lib = None
_ref = None

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_uip = ctypes.POINTER(ctypes.c_uint)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_total_probability.restype = ctypes.c_double
        L.orc_mt_next.restype = ctypes.c_uint32
        L.orc_mt_uniform.restype = ctypes.c_double
        L.orc_reg_create.restype = ctypes.c_void_p
        L.orc_reg_free.argtypes = [ctypes.c_void_p]
        L.orc_qft_op_count.restype = ctypes.c_size_t
        L.orc_grover_iterations.restype = ctypes.c_uint
        L.orc_grover_iterations.argtypes = [ctypes.c_uint64, ctypes.c_size_t]
        L.orc_grover_ops.restype = ctypes.c_longlong
        L.orc_controlled_phase.argtypes = [_dp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_double]
        L.orc_inversion_about_mean_value.argtypes = [_dp, ctypes.c_uint, ctypes.c_double, ctypes.c_double]
        L.orc_affine.argtypes = [_dp, ctypes.c_uint] + [ctypes.c_double] * 4
        L.orc_sample.argtypes = [_dp, ctypes.c_uint, ctypes.c_double, _u64p]
        L.orc_phase_flip.argtypes = [_dp, ctypes.c_uint, ctypes.c_uint64]
        L.orc_reg_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(Op), ctypes.c_size_t, _dp, _uip]
        L.orc_reg_load.argtypes = [ctypes.c_void_p, _dp]
        L.orc_reg_export.argtypes = [ctypes.c_void_p, _dp]
        L.orc_reg_measure.argtypes = [ctypes.c_void_p, ctypes.c_double, _u64p]
        L.orc_reg_total_probability.argtypes = [ctypes.c_void_p, _dp]
        L.orc_reg_inversion_lazy_value.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_reg_inversion_materialized_value.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_reg_inversion_lazy.argtypes = [ctypes.c_void_p]
        L.orc_reg_inversion_materialized.argtypes = [ctypes.c_void_p]
        L.orc_reg_amplitude.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _dp, _dp]
        L.orc_reg_create.argtypes = [ctypes.c_uint]
        L.orc_expectation.argtypes = [_dp, ctypes.c_uint, ctypes.POINTER(DiagTerm), ctypes.c_size_t, _dp]
        L.orc_ring_ops.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint32, ctypes.POINTER(Op)]
        L.orc_ring_ops.restype = ctypes.c_size_t
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise OSError(f"{REF_PATH} is not built")
        R = ctypes.CDLL(REF_PATH)
        R.ref_run.argtypes = [ctypes.c_uint, _dp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, _uip, _dp, _dp,
                              ctypes.c_longlong, _u64p, _dp]
        R.ref_register_inversion.argtypes = [ctypes.c_uint, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double, _dp, _u64p, _dp]
        R.ref_time_ops.argtypes = [ctypes.c_uint, ctypes.POINTER(Op), ctypes.c_size_t, _dp]
        R.ref_time_ops_each.argtypes = [ctypes.c_uint, ctypes.POINTER(Op), ctypes.c_size_t, _dp, _dp]
        R.ref_unit_draws.argtypes = [ctypes.c_uint32, ctypes.c_size_t, _dp]
        R.ref_random_phase.argtypes = [ctypes.c_uint, ctypes.c_uint, _dp]
        _ref = R
    return _ref


def _check(rc, what=""):
    if rc != 0:
        raise OracleError(rc, what)


# ---------------------------------------------------------------- helpers
def make_ops(seq):
    """seq: list of tuples (kind, q0, q1, theta, state[, k]) -> ctypes array."""
    arr = (Op * max(1, len(seq)))()
    for i, t in enumerate(seq):
        kind = t[0]
        arr[i].kind = kind
        arr[i].q0 = t[1] if len(t) > 1 else 0
        arr[i].q1 = t[2] if len(t) > 2 else 0
        arr[i].theta = t[3] if len(t) > 3 else 0.0
        arr[i].state = t[4] if len(t) > 4 else 0
        arr[i].k = t[5] if len(t) > 5 else 0
    return arr, len(seq)


def qft_ops(n):
    L = lib()
    cnt = L.orc_qft_op_count(n)
    arr = (Op * max(1, cnt))()
    L.orc_qft_ops(ctypes.c_uint(n), arr)
    return arr, cnt


def grover_ops(n, marked, materialized):
    L = lib()
    m = np.ascontiguousarray(marked, dtype=np.uint64)
    mp = m.ctypes.data_as(_u64p)
    cnt = L.orc_grover_ops(ctypes.c_uint(n), mp, ctypes.c_size_t(len(m)), ctypes.c_int(int(materialized)), None,
                           ctypes.c_size_t(0))
    if cnt < 0:
        raise OracleError(int(cnt), "grover_ops")
    arr = (Op * max(1, cnt))()
    rc = L.orc_grover_ops(ctypes.c_uint(n), mp, ctypes.c_size_t(len(m)), ctypes.c_int(int(materialized)), arr,
                          ctypes.c_size_t(cnt))
    if rc < 0:
        raise OracleError(int(rc), "grover_ops")
    return arr, cnt


def grover_iterations(n, m):
    return lib().orc_grover_iterations(ctypes.c_uint64(1 << n), ctypes.c_size_t(m))


def mt_unit_draws(seed, count):
    out = np.empty(count, dtype=np.float64)
    lib().orc_mt_unit_draws(ctypes.c_uint32(seed), ctypes.c_size_t(count), _dptr(out))
    return out


def random_phase(n, seed):
    out = np.empty(2 << n, dtype=np.float64)
    lib().orc_random_phase(ctypes.c_uint(n), ctypes.c_uint32(seed), _dptr(out))
    return out


def to_complex(inter):
    return inter[0::2] + 1j * inter[1::2]


def to_interleaved(c):
    out = np.empty(2 * len(c), dtype=np.float64)
    out[0::2] = c.real
    out[1::2] = c.imag
    return out


class Register:
    """Dense register with the reference's QuantumRegister semantics."""

    def __init__(self, n, amps=None):
        self.n = n
        self.h = lib().orc_reg_create(n)
        if not self.h:
            raise MemoryError("orc_reg_create")
        if amps is not None:
            a = np.ascontiguousarray(amps, dtype=np.float64)
            _check(lib().orc_reg_load(self.h, _dptr(a)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_reg_free(self.h)
            self.h = None

    def run(self, ops, count, mat_pool=None, mat_qubits=None):
        mp = None if mat_pool is None else _dptr(mat_pool)
        mq = None if mat_qubits is None else mat_qubits.ctypes.data_as(_uip)
        _check(lib().orc_reg_run(self.h, ops, ctypes.c_size_t(count), mp, mq), "run")

    def amplitudes(self):
        out = np.empty(2 << self.n, dtype=np.float64)
        lib().orc_reg_export(self.h, _dptr(out))
        return out

    def amplitude(self, s):
        re, im = ctypes.c_double(), ctypes.c_double()
        _check(lib().orc_reg_amplitude(self.h, ctypes.c_uint64(s), ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def measure(self, rnd):
        out = ctypes.c_uint64()
        _check(lib().orc_reg_measure(self.h, ctypes.c_double(rnd), ctypes.byref(out)), "measure")
        return out.value

    def total_probability(self):
        out = ctypes.c_double()
        _check(lib().orc_reg_total_probability(self.h, ctypes.byref(out)))
        return out.value

    def inversion(self, materialized, mean=None):
        L = lib()
        if mean is None:
            fn = L.orc_reg_inversion_materialized if materialized else L.orc_reg_inversion_lazy
            _check(fn(self.h))
        else:
            fn = L.orc_reg_inversion_materialized_value if materialized else L.orc_reg_inversion_lazy_value
            _check(fn(self.h, ctypes.c_double(mean.real), ctypes.c_double(mean.imag)))


def run_circuit(n, seq_or_ops, init=None, mat_pool=None, mat_qubits=None):
    """Runs ops on the oracle register; returns interleaved amplitudes."""
    if isinstance(seq_or_ops, tuple):
        ops, cnt = seq_or_ops
    else:
        ops, cnt = make_ops(seq_or_ops)
    if init is None:
        reg = Register(n)
    else:
        reg = Register(n, init)
    reg.run(ops, cnt, mat_pool, mat_qubits)
    return reg


def apply_gate(amps, n, mat, qubits):
    a = np.ascontiguousarray(amps, dtype=np.float64).copy()
    m = np.ascontiguousarray(mat, dtype=np.float64)
    q = np.ascontiguousarray(qubits, dtype=np.uint32)
    _check(lib().orc_apply_gate(_dptr(a), ctypes.c_uint(n), _dptr(m), ctypes.c_uint(len(q)),
                                q.ctypes.data_as(_uip)), "apply_gate")
    return a


def sample(amps, n, rnd):
    out = ctypes.c_uint64()
    a = np.ascontiguousarray(amps, dtype=np.float64)
    _check(lib().orc_sample(_dptr(a), ctypes.c_uint(n), ctypes.c_double(rnd), ctypes.byref(out)), "sample")
    return out.value


def total_probability(amps, n):
    a = np.ascontiguousarray(amps, dtype=np.float64)
    return lib().orc_total_probability(_dptr(a), ctypes.c_uint(n))


def make_terms(terms):
    """[(kind, coef, bits), ...] -> ctypes array of DiagTerm."""
    arr = (DiagTerm * max(1, len(terms)))()
    for i, (kind, coef, bits) in enumerate(terms):
        arr[i].kind = kind
        arr[i].coef = coef
        arr[i].bits = bits
    return arr, len(terms)


def expectation(amps, n, terms):
    arr, cnt = make_terms(terms)
    a = np.ascontiguousarray(amps, dtype=np.float64)
    out = ctypes.c_double()
    _check(lib().orc_expectation(_dptr(a), ctypes.c_uint(n), arr, ctypes.c_size_t(cnt), ctypes.byref(out)), "expectation")
    return out.value


def ring_ops(n, layers, seed):
    cnt = 2 * n * layers
    arr = (Op * max(1, cnt))()
    got = lib().orc_ring_ops(ctypes.c_uint(n), ctypes.c_uint(layers), ctypes.c_uint32(seed), arr)
    assert got == cnt
    return arr, cnt


def param_indices(ops, cnt):
    return [i for i in range(cnt) if ops[i].kind in (OP_CPHASE, OP_PHASE)]


def energy(n, ops, cnt, terms, init=None, mat_pool=None, mat_qubits=None):
    reg = run_circuit(n, (ops, cnt), init=init, mat_pool=mat_pool, mat_qubits=mat_qubits)
    return expectation(reg.amplitudes(), n, terms)


def _shifted(ops, cnt, i, delta):
    cp = (Op * max(1, cnt))()
    ctypes.memmove(cp, ops, ctypes.sizeof(Op) * cnt)
    cp[i].theta = ops[i].theta + delta
    return cp


def param_shift_grad(n, ops, cnt, terms, init=None, mat_pool=None, mat_qubits=None):
    """dE/dtheta for every CPHASE/PHASE op: (E(theta + pi/2) - E(theta - pi/2)) / 2."""
    out = []
    for i in param_indices(ops, cnt):
        ep = energy(n, _shifted(ops, cnt, i, np.pi / 2), cnt, terms, init, mat_pool, mat_qubits)
        em = energy(n, _shifted(ops, cnt, i, -np.pi / 2), cnt, terms, init, mat_pool, mat_qubits)
        out.append(0.5 * (ep - em))
    return np.array(out)


def central_fd_grad(n, ops, cnt, terms, h=1e-6, init=None, mat_pool=None, mat_qubits=None):
    out = []
    for i in param_indices(ops, cnt):
        ep = energy(n, _shifted(ops, cnt, i, h), cnt, terms, init, mat_pool, mat_qubits)
        em = energy(n, _shifted(ops, cnt, i, -h), cnt, terms, init, mat_pool, mat_qubits)
        out.append((ep - em) / (2 * h))
    return np.array(out)


# ------------------------------------------------------- reference build
def ref_run(n, seq_or_ops, init=None, mat_pool=None, mat_qubits=None, measure_seed=-1):
    """Runs ops through the reference's Simulator (dense).  Returns
    (interleaved amplitudes, total probability, measured or None, exec seconds)."""
    if isinstance(seq_or_ops, tuple):
        ops, cnt = seq_or_ops
    else:
        ops, cnt = make_ops(seq_or_ops)
    out = np.empty(2 << n, dtype=np.float64)
    tp = ctypes.c_double()
    meas = ctypes.c_uint64()
    secs = ctypes.c_double()
    ip = None
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float64)
        ip = _dptr(init)
    mp = None if mat_pool is None else _dptr(mat_pool)
    mq = None if mat_qubits is None else mat_qubits.ctypes.data_as(_uip)
    rc = ref().ref_run(ctypes.c_uint(n), ip, ops, ctypes.c_size_t(cnt), mp, mq, _dptr(out), ctypes.byref(tp),
                       ctypes.c_longlong(measure_seed), ctypes.byref(meas), ctypes.byref(secs))
    _check(rc, "ref_run")
    return out, tp.value, (meas.value if measure_seed >= 0 else None), secs.value


def ref_unit_draws(seed, count):
    out = np.empty(count, dtype=np.float64)
    _check(ref().ref_unit_draws(ctypes.c_uint32(seed), ctypes.c_size_t(count), _dptr(out)))
    return out


def ref_random_phase(n, seed):
    out = np.empty(2 << n, dtype=np.float64)
    _check(ref().ref_random_phase(ctypes.c_uint(n), ctypes.c_uint(seed), _dptr(out)))
    return out


def ref_time_ops(n, seq_or_ops):
    return time_ops_each(n, seq_or_ops, reference=True)[0]


def time_ops_each(n, seq_or_ops, reference):
    """Single-thread CPU time of ops on a uniform 2^n-state register: the
    reference build (reference=True) or the restatement.  Returns (total
    seconds, per-operation seconds)."""
    if isinstance(seq_or_ops, tuple):
        ops, cnt = seq_or_ops
    else:
        ops, cnt = make_ops(seq_or_ops)
    secs = ctypes.c_double()
    per = np.zeros(max(1, cnt), dtype=np.float64)
    if reference:
        fn = ref().ref_time_ops_each
    else:
        fn = lib().orc_time_ops_each
        fn.argtypes = [ctypes.c_uint, ctypes.POINTER(Op), ctypes.c_size_t, _dp, _dp]
    _check(fn(ctypes.c_uint(n), ops, ctypes.c_size_t(cnt), ctypes.byref(secs), _dptr(per)), "time_ops")
    return secs.value, per[:cnt]
