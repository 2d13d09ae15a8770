"""MI355X (gfx950) dense state-vector engine: Python bindings.

``HipState`` binds the C-ABI of ``lib/libtmfqs_hip.so`` (declared in
``include/tmfqs_hip.h``) with ctypes; it is the handle a test or a Python
caller uses in place of the reference's ``IStateBackend``.  ``Simulator``
(module ``sim``) binds the C++ API's simulator entry points.

Nothing here computes on the CPU: every operation runs on the GPU through the
native library, and importing a binding raises if the library is missing.
"""

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libtmfqs_hip.so")

C128, C64 = 0, 1

OK = 0
ERR_OUT_OF_RANGE = -1
ERR_INVALID_ARGUMENT = -2
ERR_LOGIC = -3
ERR_RUNTIME = -4
ERR_OUT_OF_MEMORY = -5
ERR_DEVICE = -6

OP_H, OP_X, OP_CX, OP_CPHASE, OP_SWAP, OP_PHASE_FLIP, OP_DIFFUSE_LAZY, OP_DIFFUSE_MAT, OP_MATRIX, OP_PHASE = range(10)
DIAG_Z, DIAG_PROJECTOR = 0, 1


class Config(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("device", ctypes.c_int),
        ("fusion", ctypes.c_int),
        ("tile_qubits", ctypes.c_int),
        ("coalesce_qubits", ctypes.c_int),
        ("grover_block", ctypes.c_int),
        ("exact_phases", ctypes.c_int),
        ("fuse_qubits", ctypes.c_int),
        ("fuse_min_gates", ctypes.c_int),
        ("qft_plan", ctypes.c_int),
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


class DiagTerm(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
        ("coef", ctypes.c_double),
        ("bits", ctypes.c_uint64),
    ]


def make_terms(terms):
    """[(kind, coef, bits), ...] -> (ctypes DiagTerm array, count)."""
    arr = (DiagTerm * max(1, len(terms)))()
    for i, (kind, coef, bits) in enumerate(terms):
        arr[i].kind = kind
        arr[i].coef = coef
        arr[i].bits = bits
    return arr, len(terms)


def circuit_tune_key(n, precision, ops, count):
    """Key of a tuned single-circuit configuration: precision, qubit count and
    the circuit's shape (each gate's kind and qubits; angles and matrix
    entries do not change what is fastest)."""
    import hashlib

    h = hashlib.sha1(f"circuit;precision={int(precision)};n={int(n)}".encode())
    for i in range(count):
        o = ops[i]
        h.update(f";{o.kind},{o.q0},{o.q1},{o.k}".encode())
    return h.hexdigest()


def _pools(mat_pool, qubit_pool):
    mp = None if mat_pool is None else np.ascontiguousarray(mat_pool, dtype=np.float64)
    qp = None if qubit_pool is None else np.ascontiguousarray(qubit_pool, dtype=np.uint32)
    return (mp, qp, None if mp is None else mp.ctypes.data_as(_dp),
            None if qp is None else qp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))


class Comparison(ctypes.Structure):
    _fields_ = [
        ("bitwise_equal", ctypes.c_int),
        ("max_abs_amplitude_error", ctypes.c_double),
        ("max_abs_component_error", ctypes.c_double),
        ("rmse_amplitude", ctypes.c_double),
        ("rel_l2", ctypes.c_double),
        ("max_abs_probability_error", ctypes.c_double),
        ("total_probability_diff", ctypes.c_double),
        ("worst_state", ctypes.c_uint64),
        ("overlap_re", ctypes.c_double),
        ("overlap_im", ctypes.c_double),
        ("fidelity", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HipError(RuntimeError):
    """Raised for a non-zero status; ``code`` is the tmfqs_hip_status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class OutOfRange(HipError, IndexError):
    pass


class InvalidArgument(HipError, ValueError):
    pass


_EXC = {ERR_OUT_OF_RANGE: OutOfRange, ERR_INVALID_ARGUMENT: InvalidArgument}

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def lib():
    """Loads libtmfqs_hip.so; raises OSError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise OSError(f"{HIP_LIB} is missing: run python -m tmfqsfullstate_amd.build")
        L = ctypes.CDLL(HIP_LIB)
        L.tmfqs_hip_last_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        L.tmfqs_hip_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        L.tmfqs_hip_clone.argtypes = [vp, ctypes.POINTER(vp)]
        L.tmfqs_hip_destroy.argtypes = [vp]
        L.tmfqs_hip_init_zero.argtypes = [vp, ctypes.c_uint]
        L.tmfqs_hip_init_basis.argtypes = [vp, ctypes.c_uint, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.tmfqs_hip_init_uniform.argtypes = [vp, ctypes.c_uint, _u64p, ctypes.c_size_t]
        L.tmfqs_hip_load.argtypes = [vp, ctypes.c_uint, _dp, ctypes.c_uint64]
        L.tmfqs_hip_init_random_phase.argtypes = [vp, ctypes.c_uint, ctypes.c_uint]
        L.tmfqs_hip_mt_jump_outputs.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_uint32)]
        L.tmfqs_hip_phase_eval.argtypes = [ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_int)]
        L.tmfqs_hip_export.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, _dp]
        L.tmfqs_hip_amplitude.argtypes = [vp, ctypes.c_uint64, _dp, _dp]
        L.tmfqs_hip_set_amplitude.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.tmfqs_hip_probability.argtypes = [vp, ctypes.c_uint64, _dp]
        L.tmfqs_hip_total_probability.argtypes = [vp, _dp]
        L.tmfqs_hip_sample.argtypes = [vp, ctypes.c_double, _u64p]
        L.tmfqs_hip_amplitude_sum.argtypes = [vp, _dp, _dp]
        L.tmfqs_hip_total_probability_affine.argtypes = [vp] + [ctypes.c_double] * 4 + [_dp]
        L.tmfqs_hip_sample_affine.argtypes = [vp, ctypes.c_double] + [ctypes.c_double] * 4 + [_u64p]
        L.tmfqs_hip_phase_flip.argtypes = [vp, ctypes.c_uint64]
        L.tmfqs_hip_inversion_about_mean.argtypes = [vp]
        L.tmfqs_hip_inversion_about_mean_value.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.tmfqs_hip_affine.argtypes = [vp] + [ctypes.c_double] * 4
        L.tmfqs_hip_hadamard.argtypes = [vp, ctypes.c_uint]
        L.tmfqs_hip_pauli_x.argtypes = [vp, ctypes.c_uint]
        L.tmfqs_hip_controlled_phase.argtypes = [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_double]
        L.tmfqs_hip_controlled_not.argtypes = [vp, ctypes.c_uint, ctypes.c_uint]
        L.tmfqs_hip_swap.argtypes = [vp, ctypes.c_uint, ctypes.c_uint]
        L.tmfqs_hip_apply_gate.argtypes = [vp, _dp, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
        L.tmfqs_hip_begin_batch.argtypes = [vp]
        L.tmfqs_hip_end_batch.argtypes = [vp]
        L.tmfqs_hip_synchronize.argtypes = [vp]
        L.tmfqs_hip_profile_enable.argtypes = [vp, ctypes.c_int]
        L.tmfqs_hip_profile_reset.argtypes = [vp]
        L.tmfqs_hip_profile_read.argtypes = [vp, ctypes.c_void_p, _u64p, _dp, _dp, ctypes.POINTER(ctypes.c_size_t)]
        L.tmfqs_hip_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.tmfqs_hip_copy_probe.argtypes = [vp, ctypes.c_int, _dp]
        L.tmfqs_hip_compare.argtypes = [vp, vp, ctypes.POINTER(Comparison)]
        L.tmfqs_hip_qubit_count.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
        L.tmfqs_hip_run.argtypes = [vp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint)]
        L.tmfqs_hip_expectation.argtypes = [vp, ctypes.POINTER(DiagTerm), ctypes.c_size_t, _dp]
        L.tmfqs_hip_tune.argtypes = [vp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _dp]
        L.tmfqs_hip_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tmfqs_hip_get_config.argtypes = [vp, ctypes.POINTER(Config)]
        L.tmfqs_hip_expectation_device.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_uint,
                                                   ctypes.POINTER(DiagTerm), ctypes.c_size_t, vp]
        L.tmfqs_hip_adjoint.argtypes = [vp, ctypes.POINTER(Op), ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_uint),
                                        ctypes.POINTER(DiagTerm), ctypes.c_size_t, ctypes.c_uint, _dp, _dp,
                                        ctypes.c_size_t]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().tmfqs_hip_last_error().decode(errors="replace")
        raise _EXC.get(rc, HipError)(rc, msg)


def device_count():
    c = ctypes.c_int(0)
    check(lib().tmfqs_hip_device_count(ctypes.byref(c)))
    return c.value


def default_config(**kw):
    cfg = Config()
    lib().tmfqs_hip_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, int(v))
    return cfg


def mt_jump_outputs(seed, jump_words, count):
    """The count std::mt19937(seed) outputs after its first jump_words, through
    the jump-ahead polynomials (host computation)."""
    out = np.empty(count, dtype=np.uint32)
    check(lib().tmfqs_hip_mt_jump_outputs(seed, jump_words, count,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return out


def phase_eval(x):
    """The device's double-double (cos hi, cos lo, sin hi, sin lo) of x and its
    near-midpoint flags (cos, sin), evaluated on the host."""
    out = np.empty(4, dtype=np.float64)
    fl = (ctypes.c_int * 2)()
    check(lib().tmfqs_hip_phase_eval(x, out.ctypes.data_as(_dp), fl))
    return out, (bool(fl[0]), bool(fl[1]))


def expectation_device(amps_ptr, n, terms, out_ptr, precision=C128, device=0, stream=None):
    """tmfqs_hip_expectation_device: <psi|O|psi> of the 2^n amplitudes at the
    device address amps_ptr into the device double at out_ptr, enqueued on
    `stream` (a hipStream_t as an integer; None = the default stream) without
    allocating or waiting."""
    arr, cnt = make_terms(terms)
    check(lib().tmfqs_hip_expectation_device(device, stream, precision, amps_ptr, n, arr, cnt, out_ptr))


class HipState:
    """One device state vector behind the C-ABI."""

    def __init__(self, precision=C128, device=0, _handle=None, **cfg):
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            c = default_config(precision=precision, device=device, **cfg)
            check(lib().tmfqs_hip_create(ctypes.byref(c), ctypes.byref(self._h)))
        self.n = None

    def close(self):
        if self._h:
            lib().tmfqs_hip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def clone(self):
        out = ctypes.c_void_p()
        check(lib().tmfqs_hip_clone(self._h, ctypes.byref(out)))
        c = HipState(_handle=out)
        c.n = self.n
        return c

    # ---- init / IO
    def init_zero(self, n):
        check(lib().tmfqs_hip_init_zero(self._h, n))
        self.n = n
        return self

    def init_basis(self, n, state=0, amp=1.0 + 0j):
        check(lib().tmfqs_hip_init_basis(self._h, n, state, amp.real, amp.imag))
        self.n = n
        return self

    def init_uniform(self, n, states):
        a = np.ascontiguousarray(states, dtype=np.uint64)
        check(lib().tmfqs_hip_init_uniform(self._h, n, a.ctypes.data_as(_u64p), len(a)))
        self.n = n
        return self

    def init_random_phase(self, n, seed):
        """QuantumState::randomPhase(n, seed) generated in device memory."""
        check(lib().tmfqs_hip_init_random_phase(self._h, n, seed))
        self.n = n
        return self

    def load(self, n, interleaved):
        a = np.ascontiguousarray(interleaved, dtype=np.float64)
        check(lib().tmfqs_hip_load(self._h, n, a.ctypes.data_as(_dp), a.size))
        self.n = n
        return self

    def export(self, first=0, count=None):
        if count is None:
            count = (1 << self.n) - first
        out = np.empty(2 * count, dtype=np.float64)
        check(lib().tmfqs_hip_export(self._h, first, count, out.ctypes.data_as(_dp)))
        return out

    def amplitudes(self):
        e = self.export()
        return e[0::2] + 1j * e[1::2]

    def amplitude(self, s):
        re, im = ctypes.c_double(), ctypes.c_double()
        check(lib().tmfqs_hip_amplitude(self._h, s, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def set_amplitude(self, s, v):
        check(lib().tmfqs_hip_set_amplitude(self._h, s, v.real, v.imag))

    def probability(self, s):
        o = ctypes.c_double()
        check(lib().tmfqs_hip_probability(self._h, s, ctypes.byref(o)))
        return o.value

    def total_probability(self):
        o = ctypes.c_double()
        check(lib().tmfqs_hip_total_probability(self._h, ctypes.byref(o)))
        return o.value

    def sample(self, rnd):
        o = ctypes.c_uint64()
        check(lib().tmfqs_hip_sample(self._h, rnd, ctypes.byref(o)))
        return o.value

    def amplitude_sum(self):
        re, im = ctypes.c_double(), ctypes.c_double()
        check(lib().tmfqs_hip_amplitude_sum(self._h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def total_probability_affine(self, scale, bias):
        o = ctypes.c_double()
        check(lib().tmfqs_hip_total_probability_affine(self._h, scale.real, scale.imag, bias.real, bias.imag,
                                                       ctypes.byref(o)))
        return o.value

    def sample_affine(self, rnd, scale, bias):
        o = ctypes.c_uint64()
        check(lib().tmfqs_hip_sample_affine(self._h, rnd, scale.real, scale.imag, bias.real, bias.imag,
                                            ctypes.byref(o)))
        return o.value

    # ---- operations
    def h(self, q):
        check(lib().tmfqs_hip_hadamard(self._h, q))

    def x(self, q):
        check(lib().tmfqs_hip_pauli_x(self._h, q))

    def cx(self, c, t):
        check(lib().tmfqs_hip_controlled_not(self._h, c, t))

    def cphase(self, c, t, theta):
        check(lib().tmfqs_hip_controlled_phase(self._h, c, t, theta))

    def swap(self, a, b):
        check(lib().tmfqs_hip_swap(self._h, a, b))

    def gate(self, mat, qubits):
        m = np.ascontiguousarray(mat, dtype=np.complex128)
        inter = np.empty(2 * m.size)
        inter[0::2] = m.real.ravel()
        inter[1::2] = m.imag.ravel()
        q = (ctypes.c_uint * max(1, len(qubits)))(*qubits)
        check(lib().tmfqs_hip_apply_gate(self._h, inter.ctypes.data_as(_dp), len(qubits), q))

    def phase_flip(self, s):
        check(lib().tmfqs_hip_phase_flip(self._h, s))

    def inversion_about_mean(self, mean=None):
        if mean is None:
            check(lib().tmfqs_hip_inversion_about_mean(self._h))
        else:
            check(lib().tmfqs_hip_inversion_about_mean_value(self._h, mean.real, mean.imag))

    def affine(self, scale, bias):
        check(lib().tmfqs_hip_affine(self._h, scale.real, scale.imag, bias.real, bias.imag))

    def begin_batch(self):
        check(lib().tmfqs_hip_begin_batch(self._h))

    def end_batch(self):
        check(lib().tmfqs_hip_end_batch(self._h))

    def synchronize(self):
        check(lib().tmfqs_hip_synchronize(self._h))

    # ---- observability
    def profile(self, on=True):
        check(lib().tmfqs_hip_profile_enable(self._h, int(on)))

    def profile_reset(self):
        check(lib().tmfqs_hip_profile_reset(self._h))

    def profile_read(self):
        cap = 64
        names = (ctypes.c_char * 64 * cap)()
        launches = (ctypes.c_uint64 * cap)()
        ms = (ctypes.c_double * cap)()
        byts = (ctypes.c_double * cap)()
        cnt = ctypes.c_size_t(cap)
        check(lib().tmfqs_hip_profile_read(self._h, ctypes.cast(names, ctypes.c_void_p), launches, ms, byts,
                                           ctypes.byref(cnt)))
        out = {}
        for i in range(cnt.value):
            nm = bytes(names[i]).split(b"\0", 1)[0].decode()
            out[nm] = {"launches": launches[i], "ms": ms[i], "bytes": byts[i]}
        return out

    def copy_probe(self, reps=3):
        ms = ctypes.c_double()
        check(lib().tmfqs_hip_copy_probe(self._h, reps, ctypes.byref(ms)))
        return ms.value

    def compare(self, other):
        """Device comparison with self as the reference state."""
        c = Comparison()
        check(lib().tmfqs_hip_compare(self._h, other._h, ctypes.byref(c)))
        return c.as_dict()

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().tmfqs_hip_stream(self._h, ctypes.byref(s)))
        return s.value

    # ---- operation lists, observables, adjoint gradients
    def run(self, ops, count, mat_pool=None, qubit_pool=None):
        """tmfqs_hip_run: enqueue op records (see include/tmfqs_hip.h)."""
        mp, qp, mpp, qpp = _pools(mat_pool, qubit_pool)
        check(lib().tmfqs_hip_run(self._h, ops, count, mpp, qpp))

    def tune(self, ops, count, mat_pool=None, qubit_pool=None, cache=None):
        """tmfqs_hip_tune: times the gate list for every tile width and gate-run
        composition width on a scratch copy of the state and keeps the fastest;
        returns (tile_qubits, fuse_qubits, device ms).  With cache (a JSON
        file path) a result stored for the same precision, qubit count and
        circuit shape (see circuit_tune_key) is applied without timing, and a
        new result is stored there."""
        from .batch import tune_cache_lookup_entry, tune_cache_store

        key = circuit_tune_key(self.n, self.config().precision, ops, count)
        if cache:
            entry = tune_cache_lookup_entry(cache, key)
            if entry is not None:
                check(lib().tmfqs_hip_set_tuning(self._h, int(entry["tile_qubits"]), int(entry["fuse_qubits"])))
                return int(entry["tile_qubits"]), int(entry["fuse_qubits"]), float(entry.get("ms", 0.0))
        mp, qp, mpp, qpp = _pools(mat_pool, qubit_pool)
        t, f, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().tmfqs_hip_tune(self._h, ops, count, mpp, qpp, ctypes.byref(t), ctypes.byref(f), ctypes.byref(ms)))
        if cache:
            tune_cache_store(cache, key, t.value, fuse_qubits=f.value, ms=ms.value)
        return t.value, f.value, ms.value

    def config(self):
        c = Config()
        check(lib().tmfqs_hip_get_config(self._h, ctypes.byref(c)))
        return c

    def expectation(self, terms):
        arr, cnt = make_terms(terms)
        out = ctypes.c_double()
        check(lib().tmfqs_hip_expectation(self._h, arr, cnt, ctypes.byref(out)))
        return out.value

    def adjoint(self, ops, count, terms, checkpoint_gates=0, mat_pool=None, qubit_pool=None):
        """Runs the ops from the current state; returns (E, dE/dtheta per CPHASE/PHASE op)."""
        nparams = sum(1 for i in range(count) if ops[i].kind in (OP_CPHASE, OP_PHASE))
        grads = np.zeros(max(1, nparams))
        arr, cnt = make_terms(terms)
        mp, qp, mpp, qpp = _pools(mat_pool, qubit_pool)
        e = ctypes.c_double()
        check(lib().tmfqs_hip_adjoint(self._h, ops, count, mpp, qpp, arr, cnt, checkpoint_gates, ctypes.byref(e),
                                      grads.ctypes.data_as(_dp), nparams))
        return e.value, grads[:nparams]

    # ---- convenience: run a list of (kind, q0, q1, theta, state) ops with the
    # backend semantics (swap as one permutation; diffusion needs the mean)
    def run_gates(self, seq):
        for t in seq:
            k = t[0]
            if k == OP_H:
                self.h(t[1])
            elif k == OP_X:
                self.x(t[1])
            elif k == OP_CX:
                self.cx(t[1], t[2])
            elif k == OP_CPHASE:
                self.cphase(t[1], t[2], t[3])
            elif k == OP_SWAP:
                self.swap(t[1], t[2])
            elif k == OP_PHASE_FLIP:
                self.phase_flip(t[4])
            else:
                raise ValueError(f"op kind {k} needs register semantics")
