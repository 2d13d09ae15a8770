"""The benchmark workloads at their full sizes on the GPU, each checked
against the CPU oracle (the reference dense backend restated, bit-exact with
the reference build) or against what the mathematics fixes where the oracle
cannot run at that size:

* Grover-30 complex64, marked {1, 2^29 + 5, 2^30 - 1}, 14858 iterations:
  marked-state probability against the closed form sin^2((2k+1) asin
  sqrt(M/N)), fidelity against the complex128 run, measured index.
* adjoint n = 24 complex128 (8 ring layers, 192 angles): the expectation
  against the oracle and four gradients against the oracle's parameter-shift
  rule (eight 24-qubit oracle circuits); complex64 against the complex128
  adjoint.
* the seed-42 batch of 256 circuits: complex128 bit-exact against the
  oracle for every circuit; complex64 fidelity against the oracle; batched
  adjoint gradients against the single-circuit adjoint and, on a subset,
  against the oracle's parameter-shift rule.
* QFT-26 on randomPhase(26, 7): against the oracle (bit-exact with exact
  phases, <= 1e-12 fused) as well as the DFT.

Tolerances: as stated in each test.
"""

import ctypes
import functools
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tmfqsfullstate_amd import sim as S
from tmfqsfullstate_amd.batch import Batch, load_circuit_list

pytestmark = pytest.mark.gpu

SEED42 = os.path.join(os.path.dirname(__file__), "golden", "batch_circuits_seed42.txt")
Z0_PLUS_P0 = [(T.DIAG_Z, 1.0, 1 << 0), (T.DIAG_PROJECTOR, 1.0, 0)]


def oracle_ops(ops, cnt):
    arr = (O.Op * max(1, cnt))()
    ctypes.memmove(arr, ops, ctypes.sizeof(O.Op) * cnt)
    return arr


@functools.lru_cache(maxsize=1)
def seed42():
    return load_circuit_list(SEED42)


@functools.lru_cache(maxsize=1)
def seed42_oracle_states():
    """Final oracle states of every seed-42 circuit from |0...0>."""
    def one(c):
        n, ops, cnt = c
        init = np.zeros(2 << n)
        init[0] = 1.0
        return O.run_circuit(n, (oracle_ops(ops, cnt), cnt), init=init).amplitudes()

    with ThreadPoolExecutor(4) as ex:
        return list(ex.map(one, seed42()))


def test_grover30_complex64_three_marked():
    n = 30
    marked = [1, (1 << 29) + 5, (1 << 30) - 1]
    iters = O.grover_iterations(n, len(marked))
    assert iters == 14858
    theta = np.arcsin(np.sqrt(len(marked) / 2.0 ** n))
    closed = np.sin((2 * iters + 1) * theta) ** 2  # summed over the marked states
    assert abs(closed - 0.999999999767017) < 1e-14
    ops, cnt = S.grover_ops(n, marked, True)
    print("Grover-30 complex64 (3 marked)", flush=True)
    r64, _ = S.simulate(n, ops, cnt, precision=T.C64)
    p64 = sum(abs(r64.amplitude(m)) ** 2 for m in marked)
    assert abs(p64 - closed) <= 1e-6, p64
    m64 = r64.measure_seed(123456)
    assert m64 in marked
    print("Grover-30 complex128 (3 marked)", flush=True)
    r128, _ = S.simulate(n, ops, cnt)
    p128 = sum(abs(r128.amplitude(m)) ** 2 for m in marked)
    assert abs(p128 - closed) <= 1e-12, p128
    cmp = r128.hip_state().compare(r64.hip_state())
    assert 1.0 - cmp["fidelity"] <= 1e-6, cmp
    assert cmp["max_abs_amplitude_error"] <= 1e-6
    assert r128.measure_seed(123456) == m64


def test_adjoint_n24_complex128_vs_oracle_parameter_shift():
    # E(theta + pi/2) - E(theta - pi/2) over 2 for four of the 192 angles (the
    # first and last layer's ends and two in between), each energy a full
    # 24-qubit oracle circuit (about 3 s per circuit on the host).
    n = 24
    ops, cnt = S.ring_ops(n, 8, 3)
    with T.HipState() as s:
        s.init_basis(n, 0)
        e, g = s.adjoint(ops, cnt, Z0_PLUS_P0)
    oops, ocnt = O.ring_ops(n, 8, 3)
    assert ocnt == cnt
    for i in range(cnt):
        assert oops[i].kind == ops[i].kind and oops[i].theta == ops[i].theta
    terms = [(O.DIAG_Z, 1.0, 1 << 0), (O.DIAG_PROJECTOR, 1.0, 0)]
    print("adjoint n=24: oracle energy", flush=True)
    assert abs(e - O.energy(n, oops, ocnt, terms)) <= 1e-12
    params = O.param_indices(oops, ocnt)
    assert len(params) == g.size == 192
    for j in (0, 61, 130, 191):
        print(f"adjoint n=24: oracle parameter shift of angle {j}", flush=True)
        i = params[j]
        ep = O.energy(n, O._shifted(oops, ocnt, i, np.pi / 2), ocnt, terms)
        em = O.energy(n, O._shifted(oops, ocnt, i, -np.pi / 2), ocnt, terms)
        assert abs(g[j] - 0.5 * (ep - em)) <= 1e-10, j


def test_adjoint_n24_complex64_vs_complex128():
    n = 24
    ops, cnt = S.ring_ops(n, 8, 3)
    assert cnt == 384
    res = {}
    for prec in (T.C128, T.C64):
        print(f"adjoint n=24 precision {prec}", flush=True)
        with T.HipState(precision=prec) as s:
            s.init_basis(n, 0)
            res[prec] = s.adjoint(ops, cnt, Z0_PLUS_P0)
    e128, g128 = res[T.C128]
    e64, g64 = res[T.C64]
    assert g64.shape == g128.shape == (192,)
    assert abs(e64 - e128) <= 1e-5
    assert np.max(np.abs(g64 - g128)) <= 1e-5


def test_seed42_complex128_bitexact_vs_oracle():
    circuits = seed42()
    print("seed-42 workload: oracle states", flush=True)
    want = seed42_oracle_states()
    print("seed-42 workload: batch", flush=True)
    with Batch(circuits) as b:
        secs = b.run()
        info = b.info()
        assert info["circuits"] == 256 and info["graph"] == 1 and secs > 0
        for i, (n, ops, cnt) in enumerate(circuits):
            assert np.array_equal(b.export(i), want[i]), f"circuit {i} (n={n})"


def test_seed42_complex64_fidelity_vs_oracle():
    circuits = seed42()
    want = seed42_oracle_states()
    print("seed-42 workload complex64: batch", flush=True)
    worst = 0.0
    with Batch(circuits, precision=T.C64) as b:
        b.run()
        for i, (n, ops, cnt) in enumerate(circuits):
            got = O.to_complex(b.export(i))
            ref = O.to_complex(want[i])
            fid = abs(np.vdot(ref, got)) ** 2 / (np.vdot(got, got).real * np.vdot(ref, ref).real)
            worst = max(worst, 1.0 - fid)
            assert 1.0 - fid <= 1e-6, f"circuit {i} (n={n}): fidelity error {1.0 - fid}"
            assert np.abs(got - ref).max() <= 1e-5, f"circuit {i}"
    print(f"worst fidelity error {worst:.3e}", flush=True)


def test_seed42_batched_adjoint():
    circuits = seed42()
    print("seed-42 workload: batched adjoint", flush=True)
    with Batch(circuits) as b:
        e, g = b.adjoint(Z0_PLUS_P0)
    assert len(e) == len(g) == 256
    print("seed-42 workload: single-circuit adjoints", flush=True)
    for i, (n, ops, cnt) in enumerate(circuits):
        with T.HipState() as s:
            s.init_basis(n, 0)
            e1, g1 = s.adjoint(ops, cnt, Z0_PLUS_P0)
        assert len(g[i]) == len(g1) == sum(1 for j in range(cnt) if ops[j].kind == T.OP_CPHASE)
        assert abs(e[i] - e1) <= 1e-12 and np.max(np.abs(g[i] - g1), initial=0.0) <= 1e-12, f"circuit {i}"
    # the oracle's parameter-shift rule on the smallest circuits and one mid-sized one
    order = sorted(range(256), key=lambda i: (circuits[i][0], circuits[i][2]))
    subset = order[:6] + [next(i for i in order if circuits[i][0] == 16)]
    print(f"seed-42 workload: oracle parameter shift on circuits {subset}", flush=True)
    for i in subset:
        n, ops, cnt = circuits[i]
        ps = O.param_shift_grad(n, oracle_ops(ops, cnt), cnt, Z0_PLUS_P0)
        assert len(ps) == len(g[i]) and np.max(np.abs(g[i] - ps), initial=0.0) <= 1e-10, f"circuit {i}"


@pytest.mark.parametrize("exact", [1, 0], ids=["exact-phases", "fused"])
def test_qft26_random_phase_vs_oracle(exact):
    n = 26
    print(f"QFT-26 randomPhase(26, 7), exact_phases={exact}", flush=True)
    reg = S.Register.random_phase(n, 7, exact_phases=exact)
    ops, cnt = S.qft_ops(n)
    reg.run(ops, cnt)
    got = reg.export()
    init = O.random_phase(n, 7)
    print("QFT-26 oracle", flush=True)
    want = O.run_circuit(n, O.qft_ops(n), init=init).amplitudes()
    if exact:
        assert np.array_equal(got, want)
    else:
        assert np.abs(O.to_complex(got) - O.to_complex(want)).max() <= 1e-12
    dft = np.fft.ifft(O.to_complex(init), norm="ortho")
    assert np.abs(O.to_complex(want) - dft).max() <= 1e-12
