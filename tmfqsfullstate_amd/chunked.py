"""Chunked host-resident states: ctypes binding of the tmfqs_hip_chunked_*
entry points of include/tmfqs_hip.h (a register of 2^n amplitudes held as
2^(n-c) chunks in pinned host memory and streamed through device windows of
2^g chunks)."""

import ctypes

import numpy as np

from . import C128, Config, Op, _pools, check, default_config, lib


class Stats(ctypes.Structure):
    _fields_ = [
        ("segments", ctypes.c_uint64),
        ("windows", ctypes.c_uint64),
        ("bytes_h2d", ctypes.c_double),
        ("bytes_d2h", ctypes.c_double),
    ]


def _bind():
    L = lib()
    if not getattr(L, "_chunked_bound", False):
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.tmfqs_hip_chunked_create.argtypes = [ctypes.POINTER(Config), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                               ctypes.POINTER(vp)]
        L.tmfqs_hip_chunked_destroy.argtypes = [vp]
        L.tmfqs_hip_chunked_init_basis.argtypes = [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.tmfqs_hip_chunked_load.argtypes = [vp, dp, ctypes.c_uint64]
        L.tmfqs_hip_chunked_export.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, dp]
        L.tmfqs_hip_chunked_run.argtypes = [vp, ctypes.POINTER(Op), ctypes.c_size_t, dp, ctypes.POINTER(ctypes.c_uint)]
        L.tmfqs_hip_chunked_total_probability.argtypes = [vp, dp]
        L.tmfqs_hip_chunked_info.argtypes = [vp, ctypes.POINTER(Stats)]
        L._chunked_bound = True
    return L


class ChunkedState:
    """A register of n qubits in 2^(n - chunk_qubits) host chunks, run through
    device windows of 2^window_bits chunks (tmfqs_hip_chunked)."""

    def __init__(self, n, chunk_qubits, window_bits, precision=C128, **cfg):
        L = _bind()
        self.n = n
        c = default_config(precision=precision, **cfg)
        self._h = ctypes.c_void_p()
        check(L.tmfqs_hip_chunked_create(ctypes.byref(c), n, chunk_qubits, window_bits, ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            _bind().tmfqs_hip_chunked_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_basis(self, state=0, amp=1.0 + 0j):
        check(_bind().tmfqs_hip_chunked_init_basis(self._h, state, amp.real, amp.imag))
        return self

    def load(self, amps):
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128)).view(np.float64)
        check(_bind().tmfqs_hip_chunked_load(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size))
        return self

    def export(self, first=0, count=None):
        if count is None:
            count = (1 << self.n) - first
        out = np.empty(2 * count, dtype=np.float64)
        check(_bind().tmfqs_hip_chunked_export(self._h, first, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out.view(np.complex128)

    def run(self, ops, count, mat_pool=None, qubit_pool=None):
        mp, qp, mpp, qpp = _pools(mat_pool, qubit_pool)
        check(_bind().tmfqs_hip_chunked_run(self._h, ops, count, mpp, qpp))
        return self

    def total_probability(self):
        out = ctypes.c_double()
        check(_bind().tmfqs_hip_chunked_total_probability(self._h, ctypes.byref(out)))
        return out.value

    def info(self):
        s = Stats()
        check(_bind().tmfqs_hip_chunked_info(self._h, ctypes.byref(s)))
        return {"segments": s.segments, "windows": s.windows, "bytes_h2d": s.bytes_h2d, "bytes_d2h": s.bytes_d2h}


def main(argv=None):
    """python -m tmfqsfullstate_amd.chunked --qubits 30 --chunk-qubits 24
    --window-bits 4: QFT-n from |0..0> on a chunked register; prints the time,
    the window sweeps and the host-link traffic, and checks the output against
    the closed form (every amplitude 2^(-n/2))."""
    import argparse
    import json
    import time

    from . import sim

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--chunk-qubits", type=int, default=24)
    ap.add_argument("--window-bits", type=int, default=4)
    ap.add_argument("--precision", choices=["complex128", "complex64"], default="complex128")
    a = ap.parse_args(argv)
    ops, count = sim.qft_ops(a.qubits)
    prec = C128 if a.precision == "complex128" else 1
    with ChunkedState(a.qubits, a.chunk_qubits, a.window_bits, precision=prec) as reg:
        t0 = time.perf_counter()
        reg.run(ops, count)
        secs = time.perf_counter() - t0
        info = reg.info()
        head = reg.export(0, min(1 << a.qubits, 1 << 20))
    err = float(np.max(np.abs(head - 2.0 ** (-a.qubits / 2))))
    moved = info["bytes_h2d"] + info["bytes_d2h"]
    print(json.dumps({"qubits": a.qubits, "precision": a.precision, "seconds": secs, **info,
                      "host_link_GBps": moved / secs / 1e9, "max_abs_error_first_2^20": err}))


if __name__ == "__main__":
    main()
