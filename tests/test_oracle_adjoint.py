"""Pins the oracle's diagonal-observable expectation, the adjoint benchmark
circuit builder and its CPU gradients (CPU only; no GPU)."""

import ctypes
import math
import os

import numpy as np

import oracle as O
from tests import numpy_ref as R

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
Z0_PLUS_P0 = [(O.DIAG_Z, 1.0, 1 << 0), (O.DIAG_PROJECTOR, 1.0, 0)]


def numpy_expectation(amps, n, terms):
    a = O.to_complex(amps)
    x = np.arange(1 << n, dtype=np.uint64)
    o = np.zeros(1 << n)
    for kind, coef, bits in terms:
        if kind == O.DIAG_Z:
            par = np.zeros(1 << n, dtype=np.uint64)
            for q in range(n):
                if (bits >> q) & 1:
                    par ^= (x >> np.uint64(n - 1 - q)) & np.uint64(1)
            o += np.where(par == 1, -coef, coef)
        else:
            o[bits] += coef
    return float(np.sum(o * np.abs(a) ** 2))


def test_expectation_vs_numpy():
    rng = np.random.default_rng(5)
    n = 6
    a = rng.normal(size=2 << n)
    terms = [(O.DIAG_Z, 0.7, 0b1), (O.DIAG_Z, -1.3, 0b101010), (O.DIAG_PROJECTOR, 2.0, 0), (O.DIAG_PROJECTOR, 0.5, 45)]
    assert abs(O.expectation(a, n, terms) - numpy_expectation(a, n, terms)) < 1e-12


def test_expectation_errors():
    a = np.zeros(2 << 3)
    for terms in ([(O.DIAG_Z, 1.0, 1 << 3)], [(O.DIAG_PROJECTOR, 1.0, 8)]):
        try:
            O.expectation(a, 3, terms)
        except O.OracleError as e:
            assert e.code == O.OUT_OF_RANGE
        else:
            raise AssertionError("expected out of range")


def _zero_angles(ops, cnt):
    for i in range(cnt):
        if ops[i].kind == O.OP_CPHASE:
            ops[i].theta = 0.0


def test_ring_closed_forms():
    # theta = 0: one layer is H on every qubit (|+>^n: <Z_0> = 0, |<0|+^n>|^2 = 2^-n);
    # two layers return to |0> (E = 1 + 1).
    n = 6
    for layers, want in ((1, 2.0 ** -n), (2, 2.0)):
        ops, cnt = O.ring_ops(n, layers, 3)
        _zero_angles(ops, cnt)
        assert abs(O.energy(n, ops, cnt, Z0_PLUS_P0) - want) < 1e-14


def test_ring_angles_follow_mt19937():
    # theta_{l,i} = 2 acos(-1) * (the (n l + i)-th uniform [0, 1) draw of mt19937(3))
    ops, cnt = O.ring_ops(4, 4, 3)
    thetas = [ops[i].theta for i in range(cnt) if ops[i].kind == O.OP_CPHASE]
    assert len(thetas) == 16
    assert np.array_equal(np.array(thetas), 2.0 * math.acos(-1.0) * GOLDEN["draws_3"])
    structure = [(ops[i].kind, ops[i].q0, ops[i].q1) for i in range(8)]
    assert structure == [(O.OP_H, 0, 0), (O.OP_H, 1, 0), (O.OP_H, 2, 0), (O.OP_H, 3, 0),
                         (O.OP_CPHASE, 0, 1), (O.OP_CPHASE, 1, 2), (O.OP_CPHASE, 2, 3), (O.OP_CPHASE, 3, 0)]


def test_phase_op_matches_kron():
    # ORC_OP_PHASE = QuantumGate::PhaseShift(theta) through applyGate
    n, q, th = 4, 2, 0.734
    rng = np.random.default_rng(9)
    a = rng.normal(size=2 << n)
    ops, cnt = O.make_ops([(O.OP_PHASE, q, 0, th)])
    got = O.to_complex(O.run_circuit(n, (ops, cnt), init=a).amplitudes())
    want = R.apply(O.to_complex(a), n, np.diag([1.0, np.exp(1j * th)]), [q])
    assert np.max(np.abs(got - want)) < 1e-15


def mixed_circuit(n, seed):
    """A circuit using every differentiable record kind, with a 1-qubit matrix."""
    rng = np.random.default_rng(seed)
    seq = []
    for _ in range(6):
        for q in range(n):
            seq.append((O.OP_H, q))
        a, b = rng.choice(n, 2, replace=False)
        seq.append((O.OP_CX, int(a), int(b)))
        a, b = rng.choice(n, 2, replace=False)
        seq.append((O.OP_CPHASE, int(a), int(b), float(rng.uniform(0, 2 * np.pi))))
        seq.append((O.OP_PHASE, int(rng.integers(n)), 0, float(rng.uniform(0, 2 * np.pi))))
        a, b = rng.choice(n, 2, replace=False)
        seq.append((O.OP_SWAP, int(a), int(b)))
        seq.append((O.OP_X, int(rng.integers(n))))
        seq.append((O.OP_MATRIX, 0, 0, 0.0, 0, 1))
    ops, cnt = O.make_ops(seq)
    mats, qubits = [], []
    for i in range(cnt):
        if ops[i].kind == O.OP_MATRIX:
            u = R.random_unitary(rng, 2)
            ops[i].q0 = len(qubits)
            ops[i].state = len(mats)
            qubits.append(int(rng.integers(n)))
            mats.extend(R.interleave(u.reshape(-1)))
    return ops, cnt, np.array(mats, dtype=np.float64), np.array(qubits, dtype=np.uint32)


def test_param_shift_agrees_with_central_fd():
    n = 5
    ops, cnt, mats, qubits = mixed_circuit(n, 11)
    terms = [(O.DIAG_Z, 1.0, 0b1), (O.DIAG_Z, 0.5, 0b10110), (O.DIAG_PROJECTOR, 1.0, 0)]
    ps = O.param_shift_grad(n, ops, cnt, terms, mat_pool=mats, mat_qubits=qubits)
    fd = O.central_fd_grad(n, ops, cnt, terms, h=1e-6, mat_pool=mats, mat_qubits=qubits)
    assert len(ps) == 12
    assert np.max(np.abs(ps - fd)) < 1e-8
    assert np.max(np.abs(ps)) > 1e-3
