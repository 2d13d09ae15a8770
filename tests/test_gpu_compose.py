"""Gate-run composition (fuse_qubits): runs of consecutive gates on at most
k qubits multiplied into one dense block on the host and applied per tile
(five-qubit blocks on MFMA), against the CPU oracle applying every gate;
and the circuit autotuner (tmfqs_hip_tune) with its JSON cache.

Tolerances: complex128 max-abs amplitude error <= 1e-12 against the oracle
(composition changes the rounding, not the result); complex64 fidelity error
<= 1e-6."""

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tests import numpy_ref as R

pytestmark = pytest.mark.gpu


def deep_run_circuit(n, seed, depth=60, width=5):
    """Gates on `width` qubits chosen at random (a deep few-qubit run: random
    1-qubit unitaries, CX, H, CP, phase), between gates on the other qubits;
    returns (T op records, count, matrix pool, qubit pool)."""
    rng = np.random.default_rng(seed)
    sub = [int(q) for q in rng.choice(n, width, replace=False)]
    seq, mats, quals = [], [], []

    def matrix(qs, u):
        seq.append((T.OP_MATRIX, len(quals), 0, 0.0, sum(m.size for m in mats), len(qs)))
        quals.extend(qs)
        mats.append(R.interleave(u.reshape(-1)))

    for layer in range(2):
        for q in range(n):  # gates outside the run break it
            seq.append((T.OP_H, q))
        for _ in range(depth):
            k = int(rng.integers(0, 5))
            a, b = (int(x) for x in rng.choice(sub, 2, replace=False))
            if k == 0:
                matrix([a], R.random_unitary(rng, 2))
            elif k == 1:
                seq.append((T.OP_CX, a, b))
            elif k == 2:
                seq.append((T.OP_H, a))
            elif k == 3:
                seq.append((T.OP_CPHASE, a, b, float(rng.uniform(0, 6.3))))
            else:
                seq.append((T.OP_PHASE, a, 0, float(rng.uniform(0, 6.3))))
    arr = (T.Op * len(seq))()
    for i, g in enumerate(seq):
        arr[i].kind, arr[i].q0 = g[0], g[1]
        arr[i].q1 = g[2] if len(g) > 2 else 0
        arr[i].theta = g[3] if len(g) > 3 else 0.0
        arr[i].state = g[4] if len(g) > 4 else 0
        arr[i].k = g[5] if len(g) > 5 else 0
    pool = np.concatenate(mats) if mats else np.zeros(1)
    return arr, len(seq), np.ascontiguousarray(pool, dtype=np.float64), np.ascontiguousarray(quals, dtype=np.uint32)


def oracle_state(n, arr, cnt, pool, quals, init):
    import ctypes
    oa = (O.Op * cnt)()
    ctypes.memmove(oa, arr, ctypes.sizeof(O.Op) * cnt)
    return O.run_circuit(n, (oa, cnt), init=init, mat_pool=pool, mat_qubits=quals).amplitudes()


def composed_runs(s):
    prof = s.profile_read()
    return prof.get("gate_runs_composed", {}).get("launches", 0)


@pytest.mark.parametrize("fuse", [2, 3, 4, 5])
@pytest.mark.parametrize("n", [9, 14, 18])
def test_composed_runs_vs_oracle(n, fuse):
    arr, cnt, pool, quals = deep_run_circuit(n, 10 * n + fuse)
    init = O.random_phase(n, n)
    want = oracle_state(n, arr, cnt, pool, quals, init)
    print(f"composition n={n} fuse_qubits={fuse}", flush=True)
    with T.HipState(fuse_qubits=fuse, fuse_min_gates=8) as s:
        s.load(n, init)
        s.profile(True)
        s.run(arr, cnt, pool, quals)
        got = s.export()
        runs = composed_runs(s)
    assert np.abs(O.to_complex(got) - O.to_complex(want)).max() <= 1e-12
    if fuse == 5:
        assert runs >= 2, runs  # each deep five-qubit run became one block
    with T.HipState(precision=T.C64, fuse_qubits=fuse, fuse_min_gates=8) as s:
        s.load(n, init)
        s.run(arr, cnt, pool, quals)
        got64 = O.to_complex(s.export())
    w = O.to_complex(want)
    assert 1.0 - abs(np.vdot(w, got64)) ** 2 / (np.vdot(w, w).real * np.vdot(got64, got64).real) <= 1e-6


def test_exact_mode_never_composes():
    n = 12
    arr, cnt, pool, quals = deep_run_circuit(n, 3)
    init = O.random_phase(n, 3)
    with T.HipState(fuse_qubits=5, fuse_min_gates=8, exact_phases=1) as s:
        s.load(n, init)
        s.profile(True)
        s.run(arr, cnt, pool, quals)
        got = s.export()
        assert composed_runs(s) == 0
    assert np.array_equal(got, oracle_state(n, arr, cnt, pool, quals, init))


def test_tuner_keeps_state_and_caches(tmp_path):
    n = 20
    arr, cnt, pool, quals = deep_run_circuit(n, 77, depth=120)
    init = O.random_phase(n, 5)
    want = oracle_state(n, arr, cnt, pool, quals, init)
    cache = str(tmp_path / "tune.json")
    print("tune n=20", flush=True)
    with T.HipState() as s:
        s.load(n, init)
        t, f, ms = s.tune(arr, cnt, pool, quals, cache=cache)
        assert t in (0, 11, 12) and f in (0, 3, 4, 5) and ms > 0
        assert np.array_equal(s.export(), init)  # tuning runs on a scratch copy
        c = s.config()
        assert (c.tile_qubits, c.fuse_qubits) == (t, f)
        s.run(arr, cnt, pool, quals)
        got = s.export()
    assert np.abs(O.to_complex(got) - O.to_complex(want)).max() <= 1e-12
    with T.HipState() as s:
        s.load(n, init)
        assert s.tune(arr, cnt, pool, quals, cache=cache)[:2] == (t, f)  # from the cache, nothing timed
        c = s.config()
        assert (c.tile_qubits, c.fuse_qubits) == (t, f)
