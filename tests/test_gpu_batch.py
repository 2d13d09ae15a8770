"""Batched circuits (tmfqs_hip_batch_*) against the oracle and against the
engine's own one-circuit-at-a-time path.  The seed-42 workload itself is
checked against the oracle in tests/test_gpu_configs.py."""

import ctypes

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tmfqsfullstate_amd.batch import Batch

pytestmark = pytest.mark.gpu

def oracle_ops(ops, cnt):
    arr = (O.Op * max(1, cnt))()
    ctypes.memmove(arr, ops, ctypes.sizeof(O.Op) * cnt)
    return arr


def random_circuits(seed, count, nmin=4, nmax=12):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        seq = []
        for _ in range(int(rng.integers(20, 80))):
            k = int(rng.integers(0, 6))
            if k == 0:
                seq.append((T.OP_H, int(rng.integers(n))))
            elif k == 1:
                seq.append((T.OP_X, int(rng.integers(n))))
            elif k == 5:
                seq.append((T.OP_PHASE, int(rng.integers(n)), 0, float(rng.uniform(0, 6.3))))
            else:
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                kind = {2: T.OP_CX, 3: T.OP_CPHASE, 4: T.OP_SWAP}[k]
                seq.append((kind, a, b, float(rng.uniform(0, 6.3))))
        arr = (T.Op * len(seq))()
        for i, g in enumerate(seq):
            arr[i].kind, arr[i].q0 = g[0], g[1]
            arr[i].q1 = g[2] if len(g) > 2 else 0
            arr[i].theta = g[3] if len(g) > 3 else 0.0
        out.append((n, arr, len(seq)))
    return out


def sequential(n, ops, cnt, init=0, precision=T.C128):
    with T.HipState(precision=precision) as s:
        s.init_basis(n, init)
        s.run(ops, cnt)
        return s.export()


def test_small_batch_vs_oracle_bitexact():
    circuits = random_circuits(1, 24)
    inits = [(7 * i) % (1 << c[0]) for i, c in enumerate(circuits)]
    print("small batch: 24 circuits", flush=True)
    with Batch(circuits, init_states=inits) as b:
        b.run()
        assert b.info()["graph"] == 1
        for i, (n, ops, cnt) in enumerate(circuits):
            init = np.zeros(2 << n)
            init[2 * inits[i]] = 1.0
            ref = O.run_circuit(n, (oracle_ops(ops, cnt), cnt), init=init).amplitudes()
            assert np.array_equal(b.export(i), ref), f"circuit {i}"


def test_graph_replays_are_stable():
    circuits = random_circuits(2, 16, 6, 14)
    print("graph replays", flush=True)
    with Batch(circuits) as b:
        b.run()
        first = [b.export(i) for i in range(len(circuits))]
        for _ in range(3):
            with T.HipState() as s:  # unrelated eager work between replays
                s.init_basis(12, 3)
                s.h(0)
                s.total_probability()
            b.run()
            for i in range(len(circuits)):
                assert np.array_equal(b.export(i), first[i])


def test_tuned_batch_stays_bitexact():
    circuits = random_circuits(3, 12, 10, 16)
    print("tune", flush=True)
    with Batch(circuits) as b:
        t = b.tune()
        assert 9 <= t <= 12
        b.run()
        for i, (n, ops, cnt) in enumerate(circuits):
            assert np.array_equal(b.export(i), sequential(n, ops, cnt))


def test_tune_cache_round_trip(tmp_path):
    circuits = random_circuits(4, 10, 10, 16)
    path = str(tmp_path / "tune.json")
    print("tune and store", flush=True)
    with Batch(circuits, tune_cache=path) as b:
        assert b.cached_tile is None
        t = b.tune()
    print("reuse", t, flush=True)
    with Batch(circuits, tune_cache=path) as b:
        assert b.cached_tile == t and b.info()["tile_qubits"] == t
        assert b.tune() == t
        b.run()
        for i, (n, ops, cnt) in enumerate(circuits):
            assert np.array_equal(b.export(i), sequential(n, ops, cnt))


def test_complex64_batch_fidelity():
    circuits = random_circuits(4, 12, 8, 16)
    print("complex64 batch", flush=True)
    with Batch(circuits, precision=T.C64) as b:
        b.run()
        for i, (n, ops, cnt) in enumerate(circuits):
            ref = O.to_complex(sequential(n, ops, cnt))
            got = O.to_complex(b.export(i))
            fid = abs(np.vdot(ref, got)) ** 2
            assert 1.0 - fid <= 1e-6
            assert np.array_equal(b.export(i), sequential(n, ops, cnt, precision=T.C64))


def test_batch_errors():
    good = random_circuits(5, 2)
    bad = (T.Op * 1)()
    bad[0].kind = T.OP_MATRIX
    with pytest.raises(T.InvalidArgument):
        Batch(good + [(6, bad, 1)])
    with pytest.raises(T.InvalidArgument):
        Batch([(3, good[0][1], 0)])
    with pytest.raises(T.OutOfRange):
        Batch(good, init_states=[1 << 20, 0])
    bad[0].kind = T.OP_CX
    bad[0].q0 = bad[0].q1 = 1
    with pytest.raises(T.InvalidArgument):
        Batch([(6, bad, 1)])


def test_batched_adjoint_vs_single_and_oracle():
    circuits = random_circuits(6, 6, 5, 10)
    terms = [(T.DIAG_Z, 1.0, 1 << 0), (T.DIAG_PROJECTOR, 1.0, 0)]
    print("batched adjoint", flush=True)
    with Batch(circuits) as b:
        e, g = b.adjoint(terms)
    for i, (n, ops, cnt) in enumerate(circuits):
        with T.HipState() as s:
            s.init_basis(n, 0)
            e1, g1 = s.adjoint(ops, cnt, terms)
        # the batch steps every circuit back from its final state in one dual
        # sweep (the single path restarts from checkpoints): same values up
        # to rounding
        assert abs(e[i] - e1) <= 1e-14 and np.max(np.abs(g[i] - g1)) <= 1e-12
        ps = O.param_shift_grad(n, oracle_ops(ops, cnt), cnt, terms)
        assert np.max(np.abs(g[i] - ps)) <= 1e-10


def test_batched_adjoint_many_parameters_and_states_kept():
    # Circuits with more gradient reductions per pass than one launch holds
    # (64), uneven sizes up to 14 qubits, and the forward states still in
    # place after the adjoint.  complex64 follows its own rounding path (the
    # batch steps back from the final FP32 state through every gate).
    rng = np.random.default_rng(17)
    circuits = []
    for n in (14, 7, 12):
        seq = []
        for _ in range(240):
            k = int(rng.integers(0, 4))
            if k == 0:
                seq.append((T.OP_H, int(rng.integers(n))))
            elif k == 1:
                seq.append((T.OP_PHASE, int(rng.integers(n)), 0, float(rng.uniform(0, 6.3))))
            else:
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                seq.append((T.OP_CPHASE if k == 2 else T.OP_CX, a, b, float(rng.uniform(0, 6.3))))
        arr = (T.Op * len(seq))()
        for i, g in enumerate(seq):
            arr[i].kind, arr[i].q0 = g[0], g[1]
            arr[i].q1 = g[2] if len(g) > 2 else 0
            arr[i].theta = g[3] if len(g) > 3 else 0.0
        circuits.append((n, arr, len(seq)))
    terms = [(T.DIAG_Z, 0.5, (1 << 0) | (1 << 3)), (T.DIAG_PROJECTOR, -1.0, 5)]
    print("batched adjoint, many parameters", flush=True)
    for prec in (T.C128, T.C64):
        with Batch(circuits, precision=prec) as b:
            b.run()
            before = [b.export(i) for i in range(len(circuits))]
            e, g = b.adjoint(terms)
            for i in range(len(circuits)):
                assert np.array_equal(b.export(i), before[i])
        for i, (n, ops, cnt) in enumerate(circuits):
            with T.HipState(precision=prec) as s:
                s.init_basis(n, 0)
                e1, g1 = s.adjoint(ops, cnt, terms)
            tol = 1e-12 if prec == T.C128 else 1e-4
            assert len(g[i]) == len(g1) > 64
            assert abs(e[i] - e1) <= tol and np.max(np.abs(g[i] - g1)) <= tol, (prec, i)
            if prec == T.C128:
                ps = O.param_shift_grad(n, oracle_ops(ops, cnt), cnt, terms)
                assert np.max(np.abs(g[i] - ps)) <= 1e-10
