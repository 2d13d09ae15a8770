"""The committed batched-circuit workload (CPU only)."""

import os

import numpy as np

from tests.golden import make_batch_circuits as G
from tmfqsfullstate_amd import OP_CPHASE, OP_CX, OP_H, OP_SWAP
from tmfqsfullstate_amd.batch import load_circuit_list

PATH = os.path.join(os.path.dirname(__file__), "golden", "batch_circuits_seed42.txt")


def test_fixture_matches_generator_and_spec():
    circuits = load_circuit_list(PATH)
    assert len(circuits) == 256
    regen = G.generate()
    kinds = {OP_H: 0, OP_CX: 0, OP_CPHASE: 0, OP_SWAP: 0}
    for (n, ops, cnt), (rn, rg) in zip(circuits, regen):
        assert 10 <= n <= 22 and 50 <= cnt <= 200
        assert n == rn and cnt == len(rg)
        for i in range(cnt):
            o, g = ops[i], rg[i]
            kinds[o.kind] += 1
            assert o.q0 < n and (o.kind == OP_H or (o.q1 < n and o.q1 != o.q0))
            if o.kind == OP_CPHASE:
                assert o.theta == g[3] and 0.0 <= o.theta < 2 * np.pi
    total = sum(kinds.values())
    for k, v in kinds.items():  # each gate kind about a quarter of the gates
        assert abs(v / total - 0.25) < 0.02


def test_tune_cache_key_and_file(tmp_path):
    from tmfqsfullstate_amd import C64, C128
    from tmfqsfullstate_amd.batch import tune_cache_lookup, tune_cache_store, tune_key

    circuits = load_circuit_list(PATH)
    k128, k64 = tune_key(circuits, C128), tune_key(circuits, C64)
    assert k128 != k64 and k128 == tune_key(load_circuit_list(PATH), C128)
    # angles do not enter the key; qubit counts and gate kinds do
    n, ops, cnt = circuits[0]
    j = next(i for i in range(cnt) if ops[i].kind == OP_CPHASE)
    ops[j].theta += 1.0
    assert tune_key(circuits, C128) == k128
    ops[j].kind = OP_CX
    assert tune_key(circuits, C128) != k128
    assert tune_key(circuits[1:], C128) != k128

    path = str(tmp_path / "tune.json")
    assert tune_cache_lookup(path, k128) is None
    tune_cache_store(path, k128, 12, circuits=256)
    tune_cache_store(path, k64, 11)
    assert tune_cache_lookup(path, k128) == 12 and tune_cache_lookup(path, k64) == 11
    tune_cache_store(path, k128, 10)
    assert tune_cache_lookup(path, k128) == 10
    assert sorted(os.listdir(tmp_path)) == ["tune.json"]
