"""Pins the CPU oracle (CPU only; no GPU).

Each oracle function is checked against something that does not share its
code: the C++ standard's mt19937 value, brute-force numpy models, closed forms
(DFT, Grover amplitude), the reference's own known-answer tests, the golden
fixtures produced by the reference build, and -- where it is built -- the
reference build itself (bit-exact).
"""

import os

import numpy as np
import pytest

import oracle as O
from tests import numpy_ref as R

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


def c(inter):
    return O.to_complex(inter)


# ------------------------------------------------------------------ RNG
def test_mt19937_standard_value():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937
    # (seed 5489) produces 4123659995.
    import ctypes

    class MT(ctypes.Structure):
        _fields_ = [("mt", ctypes.c_uint32 * 624), ("idx", ctypes.c_int)]

    g = MT()
    O.lib().orc_mt_seed(ctypes.byref(g), ctypes.c_uint32(5489))
    for _ in range(9999):
        O.lib().orc_mt_next(ctypes.byref(g))
    assert O.lib().orc_mt_next(ctypes.byref(g)) == 4123659995


@pytest.mark.parametrize("seed", [123456, 100, 7, 3, 42])
def test_unit_draws_golden(seed):
    got = O.mt_unit_draws(seed, 16)
    assert np.array_equal(got, GOLDEN[f"draws_{seed}"])
    assert np.all((got >= 0) & (got < 1))


def test_random_phase_golden():
    got = O.random_phase(12, 7)
    assert np.array_equal(got, GOLDEN["random_phase_12_7"])
    amps = c(got)
    assert np.allclose(np.abs(amps), 2.0 ** -6, atol=1e-15)


# ------------------------------------------------------- gate kernels
@pytest.mark.parametrize("n", [3, 6])
def test_single_gates_vs_kron(n):
    rng = np.random.default_rng(n)
    st = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    st /= np.linalg.norm(st)
    for q in range(n):
        got = c(O.run_circuit(n, [(O.OP_H, q)], init=R.interleave(st)).amplitudes())
        assert np.allclose(got, R.apply(st, n, R.H, [q]), atol=1e-14)
        got = c(O.run_circuit(n, [(O.OP_X, q)], init=R.interleave(st)).amplitudes())
        assert np.array_equal(got, R.apply(st, n, R.X, [q]))
    for cq in range(n):
        for t in range(n):
            if cq == t:
                continue
            got = c(O.run_circuit(n, [(O.OP_CX, cq, t)], init=R.interleave(st)).amplitudes())
            assert np.array_equal(got, R.apply(st, n, R.CX, [cq, t]))
            got = c(O.run_circuit(n, [(O.OP_CPHASE, cq, t, 0.3 + cq - t)], init=R.interleave(st)).amplitudes())
            assert np.allclose(got, R.apply(st, n, R.cphase(0.3 + cq - t), [cq, t]), atol=1e-14)
            got = c(O.run_circuit(n, [(O.OP_SWAP, cq, t)], init=R.interleave(st)).amplitudes())
            assert np.array_equal(got, R.apply(st, n, R.SWAP, [cq, t]))


def test_gate_invariants():
    n = 5
    st = O.random_phase(n, 9)
    for q in range(n):
        twice = O.run_circuit(n, [(O.OP_H, q), (O.OP_H, q)], init=st).amplitudes()
        assert np.allclose(twice, st, atol=1e-15)
        cx2 = O.run_circuit(n, [(O.OP_CX, q, (q + 1) % n), (O.OP_CX, q, (q + 1) % n)], init=st).amplitudes()
        assert np.array_equal(cx2, st)
    three = O.run_circuit(n, [(O.OP_CX, 1, 3), (O.OP_CX, 3, 1), (O.OP_CX, 1, 3)], init=st).amplitudes()
    sw = O.run_circuit(n, [(O.OP_SWAP, 1, 3)], init=st).amplitudes()
    assert np.array_equal(three, sw)
    mix = O.run_circuit(n, [(O.OP_H, 0), (O.OP_CPHASE, 0, 4, 0.2), (O.OP_H, 4)], init=st)
    assert abs(mix.total_probability() - 1.0) < 1e-14


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_apply_gate_golden_and_kron(k):
    n = 8
    st = GOLDEN[f"gate_k{k}_in"]
    mat = GOLDEN[f"gate_k{k}_mat"]
    qubits = GOLDEN[f"gate_k{k}_qubits"]
    got = O.apply_gate(st, n, mat, qubits)
    assert np.array_equal(got, GOLDEN[f"gate_k{k}_out"])
    u = (mat[0::2] + 1j * mat[1::2]).reshape(1 << k, 1 << k)
    want = R.apply(c(st), n, u, list(qubits))
    assert np.allclose(c(got), want, atol=1e-13)


def test_apply_gate_errors():
    st = np.zeros(2 << 2)
    st[0] = 1
    cx = R.interleave(R.CX)
    with pytest.raises(O.OracleError) as e:
        O.apply_gate(st, 2, cx, [1, 1])
    assert e.value.code == O.INVALID_ARGUMENT
    with pytest.raises(O.OracleError) as e:
        O.apply_gate(st, 2, R.interleave(R.H), [2])
    assert e.value.code == O.OUT_OF_RANGE


def test_mixed_circuit_golden():
    ops = [tuple([int(r[0]), int(r[1]), int(r[2]), float(r[3])]) for r in GOLDEN["cx_cp_swap_ops"]]
    got = O.run_circuit(6, ops, init=O.random_phase(6, 11)).amplitudes()
    assert np.array_equal(got, GOLDEN["cx_cp_swap_amps"])


# ------------------------------------------------------------------ QFT
def test_qft_known_answers():
    # test_circuit.cpp:76-106: QFT|0> uniform, QFT|5> = e^{2 pi i 5 y / 8}/sqrt(8).
    n = 3
    reg = O.run_circuit(n, O.qft_ops(n))
    assert np.allclose(c(reg.amplitudes()), np.full(8, 1 / np.sqrt(8)), atol=1e-6)
    init = np.zeros(16)
    init[10] = 1.0
    reg = O.run_circuit(n, O.qft_ops(n), init=init)
    y = np.arange(8)
    assert np.allclose(c(reg.amplitudes()), np.exp(2j * np.pi * 5 * y / 8) / np.sqrt(8), atol=1e-6)


@pytest.mark.parametrize("n", [1, 2, 4, 9, 12])
def test_qft_vs_dft(n):
    init = O.random_phase(n, 7)
    got = c(O.run_circuit(n, O.qft_ops(n), init=init).amplitudes())
    want = np.fft.ifft(c(init), norm="ortho")
    assert np.abs(got - want).max() < 1e-13


def test_qft_golden():
    got = O.run_circuit(12, O.qft_ops(12), init=GOLDEN["random_phase_12_7"]).amplitudes()
    assert np.array_equal(got, GOLDEN["qft_12_7"])


def test_qft_op_count():
    for n in (26, 30, 32):
        assert O.qft_ops(n)[1] == n + n * (n - 1) // 2 + n // 2


# --------------------------------------------------------------- Grover
@pytest.mark.parametrize("n", [10, 12, 14])
def test_grover_golden_and_closed_form(n):
    meta = GOLDEN[f"grover_{n}_meta"]
    marked = int(meta[0])
    reg = O.run_circuit(n, O.grover_ops(n, [marked], True))
    amps = reg.amplitudes()
    assert np.array_equal(amps, GOLDEN[f"grover_{n}_amps"])
    k = O.grover_iterations(n, 1)
    p = abs(c(amps)[marked]) ** 2
    assert abs(p - R.grover_p_marked(n, 1, k)) < 1e-12
    rnd = O.mt_unit_draws(123456, 1)[0]
    assert reg.measure(rnd) == int(meta[1])


def test_grover_iteration_counts():
    assert O.grover_iterations(20, 1) == 804
    assert O.grover_iterations(30, 1) == 25735
    assert O.grover_iterations(30, 3) == 14858


def test_grover_lazy_matches_materialized():
    for n, marked in ((5, [3]), (6, [1, 40]), (7, [100, 3, 3])):
        a = O.run_circuit(n, O.grover_ops(n, marked, True)).amplitudes()
        b = O.run_circuit(n, O.grover_ops(n, marked, False)).amplitudes()
        assert np.abs(a - b).max() < 1e-12


def test_grover_reference_kats():
    # test_circuit.cpp:66-74 (n=2 marked 2, seeds 200..231) and :191-217.
    for t in range(32):
        rnd = O.mt_unit_draws(200 + t, 1)[0]
        reg = O.run_circuit(2, O.grover_ops(2, [2], False))
        assert reg.measure(rnd) == 2
    for t in range(32):
        rnd = O.mt_unit_draws(300 + t, 1)[0]
        reg = O.run_circuit(3, O.grover_ops(3, [1, 6], False))
        assert reg.measure(rnd) in (1, 6)
    with pytest.raises(O.OracleError):
        O.grover_ops(3, [8], False)
    ops, cnt = O.grover_ops(2, [0, 1, 2, 3], False)
    assert cnt == 2  # zero iterations: only the H layer


# ------------------------------------------------------ register semantics
AMP8 = np.array([0.20, 0.10, -0.15, 0.05, 0.30, -0.20, 0.00, 0.10,
                 -0.25, -0.15, 0.18, 0.12, -0.08, 0.04, 0.22, -0.06])


def test_mean_aware_inversion_matches_default():
    # test_circuit.cpp:125-176 at 1e-12.
    mean = complex(AMP8[0::2].sum() / 8, AMP8[1::2].sum() / 8)
    a = O.Register(3, AMP8)
    b = O.Register(3, AMP8)
    a.inversion(False)
    b.inversion(False, mean)
    assert np.abs(a.amplitudes() - b.amplitudes()).max() < 1e-12
    assert abs(a.total_probability() - b.total_probability()) < 1e-12


def test_overlay_measurement_fixed_rnd():
    # test_storage_parity.cpp:151-214: rnd = total/2 measured through the overlay.
    mean = complex(AMP8[0::2].sum() / 8, AMP8[1::2].sum() / 8)
    lazy = O.Register(3, AMP8)
    lazy.inversion(False, mean)
    mat = O.Register(3, AMP8)
    mat.inversion(True, mean)
    tp = mat.total_probability()
    assert abs(lazy.total_probability() - tp) < 1e-12
    assert lazy.measure(tp / 2) == mat.measure(tp / 2)


def test_measure_errors():
    reg = O.Register(2)
    with pytest.raises(O.OracleError) as e:
        reg.measure(1.0)
    assert e.value.code == O.INVALID_ARGUMENT
    with pytest.raises(O.OracleError) as e:
        reg.measure(-0.1)
    assert e.value.code == O.INVALID_ARGUMENT
    st = np.zeros(8)
    st[0] = 0.5  # unnormalised: cumulative never reaches 0.9
    with pytest.raises(O.OracleError) as e:
        O.sample(st, 2, 0.9)
    assert e.value.code == O.RUNTIME_ERROR


def test_sample_brute_force():
    rng = np.random.default_rng(5)
    n = 6
    st = O.random_phase(n, 4)
    p = c(st)
    probs = np.abs(p) ** 2
    for rnd in rng.random(50):
        cum = 0.0
        want = None
        for i, pi in enumerate(probs):
            cum += pi
            if rnd <= cum:
                want = i
                break
        assert O.sample(st, n, rnd) == want


# --------------------------------------------- reference build (here only)
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build not present")


@needs_ref
def test_oracle_bitexact_vs_reference_build():
    for seed in (123456, 1, 2 ** 31 + 5):
        assert np.array_equal(O.mt_unit_draws(seed, 64), O.ref_unit_draws(seed, 64))
    for n in (1, 5, 10):
        assert np.array_equal(O.random_phase(n, 3), O.ref_random_phase(n, 3))
        init = O.random_phase(n, 3)
        got = O.run_circuit(n, O.qft_ops(n), init=init).amplitudes()
        assert np.array_equal(got, O.ref_run(n, O.qft_ops(n), init=init)[0])
    for n, marked, mat in ((8, [200], True), (8, [3, 77], False), (9, [0], True)):
        ops = O.grover_ops(n, marked, mat)
        reg = O.run_circuit(n, ops)
        ref_amps, ref_tp, ref_meas, _ = O.ref_run(n, ops, measure_seed=123456)
        assert np.array_equal(reg.amplitudes(), ref_amps)
        assert reg.measure(O.mt_unit_draws(123456, 1)[0]) == ref_meas
        assert reg.total_probability() == ref_tp
    rng = np.random.default_rng(77)
    n = 7
    for trial in range(20):
        seq = []
        for _ in range(40):
            kind = int(rng.integers(0, 5))
            q0, q1 = (int(x) for x in rng.choice(n, 2, replace=False))
            seq.append((kind, q0, q1, float(rng.uniform(0, 2 * np.pi))))
        init = O.random_phase(n, trial)
        got = O.run_circuit(n, seq, init=init).amplitudes()
        assert np.array_equal(got, O.ref_run(n, seq, init=init)[0])


@needs_ref
def test_threaded_oracle_loops_bitexact_vs_reference_build():
    # states of 2^16 amplitudes and more run the oracle's pair, block and
    # elementwise loops split over threads: the bits must not change
    n = 17
    rng = np.random.default_rng(1717)
    mats, quals, seq = [], [], []
    for trial in range(60):
        kind = int(rng.integers(0, 7))
        q0, q1, q2 = (int(x) for x in rng.choice(n, 3, replace=False))
        if kind == 5:
            seq.append((O.OP_DIFFUSE_MAT,))
        elif kind == 6:
            k = int(rng.integers(1, 4))
            qs = [int(x) for x in rng.choice(n, k, replace=False)]
            d = 1 << k
            m = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
            u, _ = np.linalg.qr(m)
            seq.append((O.OP_MATRIX, len(quals), 0, 0.0, sum(x.size for x in mats), k))
            mats.append(O.to_interleaved(u.reshape(-1)))
            quals.extend(qs)
        else:
            seq.append((kind, q0, q1, float(rng.uniform(0, 2 * np.pi))))
    pool = np.ascontiguousarray(np.concatenate(mats), dtype=np.float64)
    qpool = np.ascontiguousarray(quals, dtype=np.uint32)
    init = O.random_phase(n, 17)
    got = O.run_circuit(n, seq, init=init, mat_pool=pool, mat_qubits=qpool).amplitudes()
    want = O.ref_run(n, seq, init=init, mat_pool=pool, mat_qubits=qpool)[0]
    assert sum(t[0] == O.OP_MATRIX for t in seq) >= 5 and sum(t[0] == O.OP_DIFFUSE_MAT for t in seq) >= 5
    assert np.array_equal(got, want)
