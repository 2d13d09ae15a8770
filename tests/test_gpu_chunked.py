"""Chunked host-resident states (tmfqs_hip_chunked_*) against the CPU oracle
and the device register.

A chunked register streams its chunks through device windows that run the
engine's own kernels; a window rewrites the records that reach outside it (a
controlled phase with a fixed control becomes a phase gate, one with both
qubits fixed a phase of the whole window, each amplitude multiplied once as
the reference does).  complex128: bit for bit with the oracle in the unfused
and exact modes for every record kind but the diffusions, whose mean is
summed window by window; max-abs 1e-12 in the fused mode (fused phase
ladders) and with diffusions.  complex64: fidelity error <= 1e-6.
"""

import ctypes

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tests import numpy_ref as R
from tmfqsfullstate_amd import sim as S
from tmfqsfullstate_amd.chunked import ChunkedState

pytestmark = pytest.mark.gpu

TOL = 1e-12


def cplx(inter):
    return inter[0::2] + 1j * inter[1::2]


def random_circuit(n, c, g, seed, count=300, matrices=True):
    """Records over every kind a chunked register takes; MATRIX records mix at
    most g chunk bits (qubit q is a chunk bit when n - 1 - q >= c)."""
    rng = np.random.default_rng(seed)
    seq, mats, mq = [], [], []
    chunk = lambda q: n - 1 - q >= c
    while len(seq) < count:
        kind = int(rng.integers(0, 8))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        th = float(rng.uniform(0, 2 * np.pi))
        if kind == 0:
            seq.append((T.OP_H, a))
        elif kind == 1:
            seq.append((T.OP_X, a))
        elif kind == 2:
            seq.append((T.OP_CX, a, b))
        elif kind == 3:
            seq.append((T.OP_CPHASE, a, b, th))
        elif kind == 4:
            if chunk(a) + chunk(b) <= g:
                seq.append((T.OP_SWAP, a, b))
        elif kind == 5:
            seq.append((T.OP_PHASE, a, 0, th))
        elif kind == 6:
            seq.append((T.OP_PHASE_FLIP, 0, 0, 0.0, int(rng.integers(1 << n))))
        elif matrices:
            k = int(rng.integers(1, 4))
            qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
            if sum(chunk(q) for q in qs) > g:
                continue
            u = R.random_unitary(rng, 1 << k)
            seq.append((T.OP_MATRIX, len(mq), 0, 0.0, len(mats), k))
            mq.extend(qs)
            mats.extend(R.interleave(u))
    ops, cnt = S.make_ops(seq)
    return ops, cnt, np.array(mats, dtype=np.float64), np.array(mq, dtype=np.uint32)


def oracle_run(n, ops, cnt, init, mats=None, mq=None):
    arr = (O.Op * max(1, cnt))()
    ctypes.memmove(arr, ops, ctypes.sizeof(O.Op) * cnt)
    mp = mats if mats is not None and len(mats) else None
    qp = mq if mq is not None and len(mq) else None
    return O.run_circuit(n, (arr, cnt), init=init, mat_pool=mp, mat_qubits=qp).amplitudes()


SHAPES = [(6, 3), (8, 1), (4, 5), (10, 2), (13, 0)]


@pytest.mark.parametrize("c,g", SHAPES, ids=[f"c{c}-g{g}" for c, g in SHAPES])
@pytest.mark.parametrize("fusion", [0, 1], ids=["unfused", "fused"])
def test_random_circuits_vs_oracle(c, g, fusion):
    n = 13
    print(f"chunked random circuit n={n} c={c} g={g} fusion={fusion}", flush=True)
    ops, cnt, mats, mq = random_circuit(n, c, g, seed=c * 10 + g)
    init = O.random_phase(n, 3)
    want = oracle_run(n, ops, cnt, init, mats, mq)
    with ChunkedState(n, c, g, fusion=fusion) as s:
        s.load(cplx(init))
        s.run(ops, cnt, mats if len(mats) else None, mq if len(mq) else None)
        got = s.export()
        info = s.info()
    assert np.max(np.abs(got - cplx(want))) <= TOL, (c, g)
    if c < n:
        assert info["windows"] >= info["segments"] * (1 << (n - c - g))


@pytest.mark.parametrize("c,g", SHAPES[:4], ids=[f"c{c}-g{g}" for c, g in SHAPES[:4]])
@pytest.mark.parametrize("mode", [dict(fusion=0), dict(fusion=1, exact_phases=1)], ids=["unfused", "exact"])
def test_random_circuits_bitexact_in_exact_modes(c, g, mode):
    n = 13
    print(f"chunked exact-mode circuit n={n} c={c} g={g} {mode}", flush=True)
    ops, cnt, mats, mq = random_circuit(n, c, g, seed=c * 10 + g + 100)
    init = O.random_phase(n, 3)
    want = oracle_run(n, ops, cnt, init, mats, mq)
    with ChunkedState(n, c, g, **mode) as s:
        s.load(cplx(init))
        s.run(ops, cnt, mats, mq)
        assert np.array_equal(s.export().view(np.float64), want)


def test_single_window_is_the_device_register_bit_for_bit():
    """With the whole register in one window a chunked run is the engine's run."""
    n = 12
    ops, cnt, mats, mq = random_circuit(n, n, 0, seed=5)
    init = O.random_phase(n, 9)
    with ChunkedState(n, n, 0) as s, T.HipState() as d:
        s.load(cplx(init))
        s.run(ops, cnt, mats, mq)
        d.load(n, init)
        d.run(ops, cnt, mats, mq)
        assert np.array_equal(s.export().view(np.float64), d.export())


def test_qft_vs_oracle():
    n = 16
    print("chunked QFT-16, c=8, g=4", flush=True)
    init = O.random_phase(n, 7)
    ops, cnt = S.qft_ops(n)
    want = cplx(oracle_run(n, ops, cnt, init))
    with ChunkedState(n, 8, 4) as s:
        s.load(cplx(init))
        s.run(ops, cnt)
        got = s.export()
        segs = s.info()["segments"]
    assert np.max(np.abs(got - want)) <= TOL
    assert segs >= 2  # the targets on chunk bits do not fit one window


def test_qft_exact_mode_bitexact():
    """QFT-14 through windows of 2^9 amplitudes in the exact mode: every
    controlled phase its own product, as in the reference, bit for bit."""
    n = 14
    init = O.random_phase(n, 8)
    ops, cnt = S.qft_ops(n)
    want = oracle_run(n, ops, cnt, init)
    with ChunkedState(n, 7, 2, exact_phases=1) as s:
        s.load(cplx(init))
        s.run(ops, cnt)
        assert np.array_equal(s.export().view(np.float64), want)


def test_grover_complex64():
    n = 12
    marked = [7, 2900, 4000]
    ops, cnt = S.grover_ops(n, marked, True)
    want = cplx(oracle_run(n, ops, cnt, None))
    with ChunkedState(n, 6, 3, precision=T.C64) as s:
        s.run(ops, cnt)
        got = s.export()
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real)
    assert 1.0 - fid <= 1e-6
    p_marked = float(np.sum(np.abs(got[marked]) ** 2))
    assert abs(p_marked - float(np.sum(np.abs(want[marked]) ** 2))) <= 1e-5


@pytest.mark.parametrize("materialized", [True, False], ids=["materialized", "lazy"])
def test_grover_vs_oracle(materialized):
    n = 10
    marked = [3, 700]
    ops, cnt = S.grover_ops(n, marked, materialized)
    want = cplx(oracle_run(n, ops, cnt, None))
    with ChunkedState(n, 5, 2) as s:
        s.run(ops, cnt)
        sweeps = s.info()["segments"]
        got = s.export()
        p = s.total_probability()
    assert np.max(np.abs(got - want)) <= TOL
    # the Hadamard layer mixes the 5 chunk bits two at a time (3 sweeps, the
    # first oracle riding along), then two sweeps per diffusion; the later
    # oracles' flips negate amplitudes in host memory without a sweep
    assert sweeps == 3 + 2 * O.grover_iterations(n, len(marked))
    assert abs(p - 1.0) <= 1e-12
    assert int(np.argmax(np.abs(got))) in marked


def test_complex64_fidelity():
    n = 14
    ops, cnt, mats, mq = random_circuit(n, 7, 3, seed=21)
    init = O.random_phase(n, 4)
    want = cplx(oracle_run(n, ops, cnt, init, mats, mq))
    with ChunkedState(n, 7, 3, precision=T.C64) as s:
        s.load(cplx(init))
        s.run(ops, cnt, mats, mq)
        got = s.export()
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real)
    assert 1.0 - fid <= 1e-6


def test_init_export_and_total_probability():
    n = 12
    with ChunkedState(n, 6, 2) as s:
        s.init_basis(1234, 0.6 + 0.8j)
        full = s.export()
        assert full[1234] == 0.6 + 0.8j and np.count_nonzero(full) == 1
        assert np.array_equal(s.export(1230, 10), full[1230:1240])
        assert abs(s.total_probability() - 1.0) <= 1e-15
        with pytest.raises(IndexError):
            s.export(4090, 10)
        with pytest.raises(IndexError):
            s.init_basis(1 << n)


def test_larger_register_qft_matches_device_register():
    """QFT-24 through 2^3 windows of 2^21 amplitudes against the device register."""
    n = 24
    print("chunked QFT-24 vs device register", flush=True)
    ops, cnt = S.qft_ops(n)
    with T.HipState() as d:
        d.init_random_phase(n, 11)
        init = d.export()
        d.run(ops, cnt)
        want = cplx(d.export())
    with ChunkedState(n, 18, 3) as s:
        s.load(cplx(init))
        s.run(ops, cnt)
        got = s.export()
        info = s.info()
    assert np.max(np.abs(got - want)) <= TOL
    assert info["bytes_h2d"] == info["bytes_d2h"] and info["bytes_h2d"] % (16 << n) == 0


def test_errors():
    with pytest.raises(ValueError):
        ChunkedState(10, 0, 1)
    with pytest.raises(ValueError):
        ChunkedState(10, 6, 5)  # c + g > n
    with pytest.raises(ValueError):
        ChunkedState(10, 6, 0)  # no window bit for the chunk bits
    n = 10
    rng = np.random.default_rng(1)
    u = R.random_unitary(rng, 8)
    with ChunkedState(n, 6, 2) as s:
        s.init_basis(5)
        before = s.export()
        # MATRIX on qubits 0, 1, 2 (index bits 9, 8, 7: three chunk bits) with a two-bit window
        ops, cnt = S.make_ops([(T.OP_H, 9), (T.OP_MATRIX, 0, 0, 0.0, 0, 3)])
        with pytest.raises(ValueError):
            s.run(ops, cnt, np.array(R.interleave(u)), np.array([0, 1, 2], dtype=np.uint32))
        assert np.array_equal(s.export(), before)  # refused before anything ran
        ops, cnt = S.make_ops([(T.OP_H, 10)])
        with pytest.raises(IndexError):
            s.run(ops, cnt)


# ---- the C++ API on StorageStrategyKind::HipChunked (QuantumRegister, Simulator)

@pytest.mark.parametrize("materialized", [True, False], ids=["materialized", "lazy"])
def test_register_grover_through_simulator(materialized):
    """Grover-12 on a chunked register (2^6-state chunks, windows of 4 chunks)
    through Simulator::run: amplitudes, total probability and sampled indices
    against the oracle register."""
    n = 12
    marked = [5, 3000]
    ops, cnt = S.grover_ops(n, marked, materialized)
    reg = S.Register.chunked(n, 0, chunk_states=1 << 6, window_chunks=4)
    reg.run(ops, cnt)
    arr = (O.Op * max(1, cnt))()
    ctypes.memmove(arr, ops, ctypes.sizeof(O.Op) * cnt)
    ref = O.run_circuit(n, (arr, cnt))
    want = cplx(ref.amplitudes())
    assert np.max(np.abs(reg.amplitudes() - want)) <= TOL
    assert abs(reg.total_probability() - ref.total_probability()) <= 1e-12
    for rnd in (0.0, 0.1, 0.49, 0.5, 0.93):
        assert reg.measure_fixed(rnd) == ref.measure(rnd), rnd
    reg.close()


def test_register_qft_and_clone():
    """QFT-14 of |10...0> through Simulator::run on a chunked register (2^8-state
    chunks, windows of 8), and a deep copy taken before it that keeps its own
    chunks and then gives the same result."""
    n = 14
    ops, cnt = S.qft_ops(n)
    state = 1 << (n - 1)
    init = np.zeros(2 << n)
    init[2 * state] = 1.0
    want = cplx(oracle_run(n, ops, cnt, init))
    reg = S.Register.chunked(n, state, chunk_states=1 << 8, window_chunks=8)
    twin = reg.clone()
    reg.run(ops, cnt)
    got = reg.amplitudes()
    assert np.max(np.abs(got - want)) <= TOL
    assert twin.amplitude(state) == 1.0
    twin.run(ops, cnt)
    assert np.array_equal(twin.amplitudes(), got)
    twin.close()
    reg.close()


def test_register_chunked_config_errors():
    with pytest.raises(ValueError):
        S.Register.chunked(10, 0, chunk_states=100)  # not a power of two
    with pytest.raises(ValueError):
        S.Register.chunked(10, 0, chunk_states=1 << 4, window_chunks=3)
