"""C-ABI surface checks that need no GPU: the library builds, loads, exports
every entry point include/tmfqs_hip.h declares, and the ctypes record layouts
match the C ones."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import tmfqsfullstate_amd as T
import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const char \*|int |void )\s*(tmfqs_\w+)\s*\(", text, re.M)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_hip_library_exports_every_declared_symbol():
    names = declared("tmfqs_hip.h")
    assert len(names) >= 39
    syms = exported(T.HIP_LIB)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    lib = T.lib()
    for n in names:
        getattr(lib, n)


def test_record_layouts():
    assert ctypes.sizeof(T.Op) == 32
    assert ctypes.sizeof(T.Op) == ctypes.sizeof(O.Op)
    for (a, *_), (b, *_) in zip(T.Op._fields_, O.Op._fields_):
        assert getattr(T.Op, a).offset == getattr(O.Op, b).offset
    assert ctypes.sizeof(T.Config) == 10 * 4
    assert [f for f, _ in T.Config._fields_][-1] == "qft_plan"


def test_abi_version_and_errors_without_device():
    lib = T.lib()
    assert lib.tmfqs_hip_abi_version() == 3
    # Creating a handle on a machine without a GPU fails cleanly (device error),
    # never by crashing; on a GPU machine it succeeds.
    h = ctypes.c_void_p()
    cfg = T.default_config()
    rc = lib.tmfqs_hip_create(ctypes.byref(cfg), ctypes.byref(h))
    if rc == 0:
        assert lib.tmfqs_hip_destroy(h) == 0
    else:
        assert rc == T.ERR_DEVICE
        assert b"device" in lib.tmfqs_hip_last_error()
    # invalid handle -> invalid argument, message set
    assert lib.tmfqs_hip_hadamard(None, 0) == T.ERR_INVALID_ARGUMENT
    assert b"null" in lib.tmfqs_hip_last_error()
    bad = T.default_config(precision=7)
    rc = lib.tmfqs_hip_create(ctypes.byref(bad), ctypes.byref(h))
    assert rc in (T.ERR_INVALID_ARGUMENT, T.ERR_DEVICE)


def test_sim_library_exports_and_host_entry_points():
    from tmfqsfullstate_amd import sim as S
    names = declared("tmfqs_sim.h")
    syms = exported(S.SIM_LIB)
    assert not [n for n in names if n not in syms]
    # host-only entry points of the C++ API, pinned against the oracle
    for n, seed in ((0, 1), (5, 7), (12, 7), (16, 123456)):
        assert np.array_equal(S.random_phase(n, seed), O.random_phase(n, seed))
    GOLDEN = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))
    assert np.array_equal(S.random_phase(12, 7), GOLDEN["random_phase_12_7"])
    for n in (1, 4, 30):
        a, na = S.qft_ops(n)
        b, nb = O.qft_ops(n)
        assert na == nb and all((a[i].kind, a[i].q0, a[i].q1, a[i].theta) == (b[i].kind, b[i].q0, b[i].q1, b[i].theta)
                                for i in range(na))
    a, na = S.grover_ops(12, [5, 4095, 5], False)
    b, nb = O.grover_ops(12, [5, 4095, 5], False)
    assert na == nb and all((a[i].kind, a[i].state) == (b[i].kind, b[i].state) for i in range(na))
    with pytest.raises(T.InvalidArgument):
        S.grover_ops(3, [8], True)
