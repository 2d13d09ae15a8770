"""QuantumState::randomPhase generated on the device (tmfqs_hip_init_random_phase,
reference src/tmfqs/state/quantum_state.cpp:264-277) against the oracle's
host construction, bit for bit, and the large-state runs it makes cheap:
QFT-32 complex64 on randomPhase(32, 11) against the complex128 run, and the
33/34-qubit capacity checks against the QFT's closed form."""

import ctypes
import math
import time

import numpy as np
import pytest

import oracle
import tmfqsfullstate_amd as T
from tmfqsfullstate_amd import sim as S

pytestmark = pytest.mark.gpu

_libm = ctypes.CDLL("libm.so.6")
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_libm.sincos.restype = None


def cplx(inter):
    return inter[0::2] + 1j * inter[1::2]


def device_random_phase(n, seed, precision=T.C128):
    s = T.HipState(precision=precision)
    s.init_random_phase(n, seed)
    return s


def expected_amplitude(n, seed, state):
    """randomPhase(n, seed)'s amplitude of one state from the seed's words
    2 state and 2 state + 1 (host jump-ahead), generate_canonical<double, 53>
    and glibc's sincos."""
    w0, w1 = (float(v) for v in T.mt_jump_outputs(seed, 2 * state, 2))
    r = (w0 + w1 * 4294967296.0) * 2.0**-64
    r = min(r, math.nextafter(1.0, 0.0))
    x = r * (2.0 * math.acos(-1.0))
    sn, cs = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(x, ctypes.byref(sn), ctypes.byref(cs))
    norm = 1.0 / math.sqrt(float(1 << n))
    return norm * cs.value, norm * sn.value


@pytest.mark.parametrize("n", list(range(1, 21)))
def test_bit_exact_against_the_oracle_small(n):
    for seed in (7, 11, 123456, 0, 2**32 - 1):
        want = oracle.random_phase(n, seed)
        with device_random_phase(n, seed) as s:
            assert np.array_equal(s.export(), want), (n, seed)
        with device_random_phase(n, seed, T.C64) as s:
            assert np.array_equal(s.export(), want.astype(np.float32).astype(np.float64)), (n, seed)


@pytest.mark.parametrize("n,seed", [(24, 7), (26, 11)])
def test_bit_exact_against_the_oracle_full_state(n, seed):
    print(f"randomPhase({n}, {seed}) against the oracle", flush=True)
    want = oracle.random_phase(n, seed)
    with device_random_phase(n, seed) as s:
        got = s.export()
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad[:8], got[bad[:8]], want[bad[:8]])


def test_register_api_materializes_on_the_device():
    # tmfqs_reg_create_random_phase: what QuantumState::randomPhase(n, seed)
    # materializes to on a Hip register.
    n = 14
    want = oracle.random_phase(n, 5)
    r = S.Register.random_phase(n, 5)
    assert np.array_equal(r.export(), want)
    assert r.total_probability() == oracle.total_probability(want, n)
    r.close()


@pytest.mark.parametrize("n,seed", [(30, 7), (32, 11)])
def test_large_states_spot_values_and_time(n, seed):
    N = 1 << n
    print(f"randomPhase({n}, {seed}) on the device", flush=True)
    t0 = time.perf_counter()
    s = device_random_phase(n, seed)
    s.synchronize()
    secs = time.perf_counter() - t0
    print(f"  resident in {secs:.3f} s", flush=True)
    if n == 30:
        assert secs < 2.0, secs
    # States at and next to the stream boundaries of the generator (streams
    # of L = 2^(n - 10) states for n >= 11) and scattered ones.
    per = (2 * N) // 2048 // 2
    picks = [0, 1, per - 1, per, per + 1, 7 * per - 1, 7 * per, N // 2, N - 2, N - 1]
    rng = np.random.default_rng(n)
    picks += [int(v) for v in rng.integers(0, N, 40)]
    for st in picks:
        got = s.export(st, 1)
        want = expected_amplitude(n, seed, st)
        assert got[0] == want[0] and got[1] == want[1], (st, got, want)
    assert abs(s.total_probability() - 1.0) < 1e-12
    s.close()


def test_qft32_complex64_on_random_phase_against_complex128():
    # QFT-32 on randomPhase(32, 11) in complex64 against the same run in
    # complex128: slices of both exported states agree within the float32
    # precision of amplitudes of size 2^-16.
    n = 32
    ops, cnt = S.qft_ops(n)
    slices = [(0, 1 << 16), ((1 << 31) - 4096, 8192), ((1 << 32) - (1 << 16), 1 << 16)]
    print("QFT-32 complex128 on randomPhase(32, 11)", flush=True)
    with device_random_phase(n, 11, T.C128) as s:
        s.run(ops, cnt)
        s.synchronize()
        ref = [cplx(s.export(a, c)) for a, c in slices]
        p128 = s.total_probability()
    print("QFT-32 complex64 on randomPhase(32, 11)", flush=True)
    with device_random_phase(n, 11, T.C64) as s:
        s.run(ops, cnt)
        s.synchronize()
        got = [cplx(s.export(a, c)) for a, c in slices]
        p64 = s.total_probability()
    scale = math.sqrt(1 << n)
    for (a, c), g, r in zip(slices, got, ref):
        err = np.abs(g - r).max() * scale
        assert err < 2e-5, (a, err)
        fid = abs(np.vdot(r, g)) ** 2 / (np.vdot(r, r).real * np.vdot(g, g).real)
        assert fid > 1 - 1e-9, (a, fid)
    assert abs(p128 - 1.0) < 1e-12 and abs(p64 - 1.0) < 1e-5


@pytest.mark.parametrize(
    "precision,n,plan", [(T.C128, 33, 0), (T.C64, 34, 0), (T.C64, 34, 1)], ids=["c128-n33", "c64-n34", "c64-n34-in-place"]
)
def test_capacity_qft_basis_closed_form(precision, n, plan):
    # 128 GiB states: QFT|s> = sum_y e^{2 pi i s y / N} |y> / sqrt(N) on three
    # slices (the reference's dense backend stops at 31 qubits).  The default
    # plan's out-of-place tail takes a second 128 GiB buffer; the in-place
    # plan needs none.
    N = 1 << n
    src = 0x13579BDF1 % N
    print(f"QFT-{n} of |{src}> ({'c128' if precision == T.C128 else 'c64'}, plan {plan})", flush=True)
    with T.HipState(precision=precision, qft_plan=plan) as s:
        s.init_basis(n, src)
        ops, cnt = S.qft_ops(n)
        s.run(ops, cnt)
        s.synchronize()
        for first in (0, N // 2 - 1000, N - 4096):
            got = cplx(s.export(first, 4096))
            y = np.arange(first, first + 4096, dtype=np.uint64)
            ph = (np.uint64(src) * y) % np.uint64(N)  # exact mod N: uint64 products wrap mod 2^64
            want = np.exp(2j * np.pi * (ph.astype(np.float64) / N)) / np.sqrt(N)
            err = np.abs(got - want).max() * np.sqrt(N)
            assert err <= (1e-10 if precision == T.C128 else 1e-5), (first, err)
