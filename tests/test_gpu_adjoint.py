"""Adjoint gradients on the GPU (tmfqs_hip_adjoint) against the CPU oracle's
parameter-shift and finite-difference gradients, and against the GPU's own
forward runs at the benchmark size (n = 24, 8 ring layers, 192 angles)."""

import ctypes

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tests.test_oracle_adjoint import Z0_PLUS_P0, mixed_circuit

pytestmark = pytest.mark.gpu


def to_hip_ops(ops, cnt):
    arr = (T.Op * max(1, cnt))()
    ctypes.memmove(arr, ops, ctypes.sizeof(T.Op) * cnt)
    return arr


def gpu_adjoint(n, ops, cnt, terms, precision=T.C128, ck=0, mats=None, qubits=None, init=0):
    with T.HipState(precision=precision) as s:
        s.init_basis(n, init)
        e, g = s.adjoint(to_hip_ops(ops, cnt), cnt, terms, ck, mats, qubits)
        final = s.export()
    return e, g, final


@pytest.mark.parametrize("n,layers", [(4, 2), (6, 3), (10, 3)])
def test_ring_vs_oracle_parameter_shift(n, layers):
    print(f"ring n={n} layers={layers}", flush=True)
    ops, cnt = O.ring_ops(n, layers, 3)
    e, g, final = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0)
    ref_state = O.run_circuit(n, (ops, cnt)).amplitudes()
    assert np.max(np.abs(final - ref_state)) <= 1e-12
    assert abs(e - O.expectation(ref_state, n, Z0_PLUS_P0)) <= 1e-12
    ps = O.param_shift_grad(n, ops, cnt, Z0_PLUS_P0)
    assert g.shape == ps.shape
    assert np.max(np.abs(g - ps)) <= 1e-10


def test_long_phase_runs_vs_oracle():
    # 150 controlled phases and phase gates between two Hadamard layers: a
    # complex128 dual pass takes such a run as its gradient reductions
    # followed by summed-angle ops of at most 64 angles (here three, and the
    # reductions span three launches of at most 64 slots).  Gradients against
    # the oracle's parameter shift, the final state against the oracle.
    n = 12
    rng = np.random.default_rng(5)
    seq = [(O.OP_H, q) for q in range(n)]
    for i in range(150):
        if i % 5 == 4:
            seq.append((O.OP_PHASE, int(rng.integers(n)), 0, float(rng.uniform(0, 2 * np.pi))))
        else:
            a, b = rng.choice(n, 2, replace=False)
            seq.append((O.OP_CPHASE, int(a), int(b), float(rng.uniform(0, 2 * np.pi))))
    seq += [(O.OP_H, q) for q in range(n)]
    ops, cnt = O.make_ops(seq)
    print("phase runs n=12", flush=True)
    e, g, final = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0)
    ref_state = O.run_circuit(n, (ops, cnt)).amplitudes()
    assert np.max(np.abs(final - ref_state)) <= 1e-12
    assert abs(e - O.expectation(ref_state, n, Z0_PLUS_P0)) <= 1e-12
    ps = O.param_shift_grad(n, ops, cnt, Z0_PLUS_P0)
    assert g.shape == ps.shape == (150,)
    assert np.max(np.abs(ps)) > 1e-3
    assert np.max(np.abs(g - ps)) <= 1e-10


def test_mixed_circuit_vs_oracle():
    n = 7
    ops, cnt, mats, qubits = mixed_circuit(n, 21)
    terms = [(O.DIAG_Z, 1.0, 0b1), (O.DIAG_Z, -0.5, 0b1011001), (O.DIAG_PROJECTOR, 1.0, 0), (O.DIAG_PROJECTOR, 0.25, 77)]
    print("mixed n=7", flush=True)
    e, g, final = gpu_adjoint(n, ops, cnt, terms, mats=mats, qubits=qubits, init=5)
    init = np.zeros(2 << n)
    init[10] = 1.0
    reg = O.run_circuit(n, (ops, cnt), init=init, mat_pool=mats, mat_qubits=qubits)
    assert np.max(np.abs(final - reg.amplitudes())) <= 1e-12
    assert abs(e - O.expectation(reg.amplitudes(), n, terms)) <= 1e-12
    ps = O.param_shift_grad(n, ops, cnt, terms, init=init, mat_pool=mats, mat_qubits=qubits)
    assert np.max(np.abs(g - ps)) <= 1e-10


def test_five_qubit_blocks_in_the_backward_sweep():
    # 5-qubit matrices between the ring layers at n = 12 (11-bit dual tiles:
    # FP64 MFMA on both states of the backward sweep) against the oracle's
    # parameter shift
    from tests import numpy_ref as R
    n = 12
    rng = np.random.default_rng(55)
    ring, rcnt = O.ring_ops(n, 3, 9)
    seq = []
    for i in range(rcnt):
        o = ring[i]
        seq.append((o.kind, o.q0, o.q1, o.theta, o.state, o.k))
        if i % 20 == 19:
            seq.append((O.OP_MATRIX, 0, 0, 0.0, 0, 5))
    ops, cnt = O.make_ops(seq)
    mats, qubits = [], []
    for i in range(cnt):
        if ops[i].kind == O.OP_MATRIX:
            ops[i].q0 = len(qubits)
            ops[i].state = len(mats)
            qubits.extend(int(q) for q in rng.choice(n, 5, replace=False))
            mats.extend(R.interleave(R.random_unitary(rng, 32).reshape(-1)))
    mats = np.array(mats, dtype=np.float64)
    qubits = np.array(qubits, dtype=np.uint32)
    assert sum(1 for i in range(cnt) if ops[i].kind == O.OP_MATRIX) >= 2
    print("ring n=12 with 5-qubit blocks", flush=True)
    e, g, final = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0, mats=mats, qubits=qubits)
    reg = O.run_circuit(n, (ops, cnt), mat_pool=mats, mat_qubits=qubits)
    assert np.max(np.abs(final - reg.amplitudes())) <= 1e-12
    assert abs(e - O.expectation(reg.amplitudes(), n, Z0_PLUS_P0)) <= 1e-12
    ps = O.param_shift_grad(n, ops, cnt, Z0_PLUS_P0, mat_pool=mats, mat_qubits=qubits)
    assert np.max(np.abs(g - ps)) <= 1e-10


def test_checkpoint_spacing_does_not_change_gradients():
    n = 9
    ops, cnt = O.ring_ops(n, 4, 7)
    print("checkpoint variants n=9", flush=True)
    _, g0, _ = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0, ck=0)
    for ck in (1, 5, 10000):
        _, g, _ = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0, ck=ck)
        assert np.max(np.abs(g - g0)) <= 1e-12


def test_complex64_gradients():
    n = 10
    ops, cnt = O.ring_ops(n, 3, 3)
    print("complex64 n=10", flush=True)
    e, g, _ = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0, precision=T.C64)
    ps = O.param_shift_grad(n, ops, cnt, Z0_PLUS_P0)
    assert abs(e - O.energy(n, ops, cnt, Z0_PLUS_P0)) <= 1e-5
    assert np.max(np.abs(g - ps)) <= 1e-5


def test_ring_n14_vs_oracle_all_angles():
    # 8 layers, 112 angles: parameter shift and central differences on the CPU oracle
    n = 14
    ops, cnt = O.ring_ops(n, 8, 3)
    print("ring n=14 layers=8", flush=True)
    e, g, _ = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0)
    ps = O.param_shift_grad(n, ops, cnt, Z0_PLUS_P0)
    assert np.max(np.abs(g - ps)) <= 1e-10
    sel = list(range(0, len(ps), 16))
    params = O.param_indices(ops, cnt)
    for j in sel:
        i = params[j]
        h = 1e-6
        ep = O.energy(n, O._shifted(ops, cnt, i, h), cnt, Z0_PLUS_P0)
        em = O.energy(n, O._shifted(ops, cnt, i, -h), cnt, Z0_PLUS_P0)
        assert abs(g[j] - (ep - em) / (2 * h)) <= 1e-8


def gpu_energy(s, n, ops, cnt, terms):
    s.init_basis(n, 0)
    s.run(ops, cnt)
    return s.expectation(terms)


def test_benchmark_size_vs_gpu_shift_and_fd():
    # n = 24, 8 layers, 192 angles from mt19937(3): adjoint against the
    # parameter-shift rule and FP64 central differences, both from forward
    # runs of the same engine.
    import tmfqsfullstate_amd.sim as S
    n = 24
    ops, cnt = S.ring_ops(n, 8, 3)
    assert cnt == 384
    print("ring n=24 layers=8: adjoint", flush=True)
    e, g, _ = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0)
    assert g.shape == (192,)
    params = [i for i in range(cnt) if ops[i].kind == T.OP_CPHASE]
    print("ring n=24: parameter shift on sampled angles", flush=True)
    with T.HipState() as s:
        assert abs(gpu_energy(s, n, ops, cnt, Z0_PLUS_P0) - e) <= 1e-12
        for j in range(0, 192, 8):
            i = params[j]
            t0 = ops[i].theta
            vals = {}
            for d in (np.pi / 2, -np.pi / 2, 1e-6, -1e-6):
                ops[i].theta = t0 + d
                vals[d] = gpu_energy(s, n, ops, cnt, Z0_PLUS_P0)
            ops[i].theta = t0
            shift = 0.5 * (vals[np.pi / 2] - vals[-np.pi / 2])
            fd = (vals[1e-6] - vals[-1e-6]) / 2e-6
            assert abs(g[j] - shift) <= 1e-9
            assert abs(g[j] - fd) <= 1e-8


def test_cpp_api_matches_c_abi():
    import tmfqsfullstate_amd.sim as S
    n = 10
    ops, cnt = S.ring_ops(n, 3, 5)
    reg = S.Register.basis(n, 0)
    e1, g1, secs = reg.adjoint(ops, cnt, Z0_PLUS_P0)
    e2, g2, final = gpu_adjoint(n, ops, cnt, Z0_PLUS_P0)
    assert e1 == e2 and np.array_equal(g1, g2) and secs > 0
    assert np.array_equal(reg.export(), final)
    assert reg.expectation(Z0_PLUS_P0) == e1


def test_torch_autograd_wrapper():
    import torch
    from tmfqsfullstate_amd.autograd import AdjointCircuit, CircuitExpectation
    n = 8
    ops, cnt = O.ring_ops(n, 2, 3)
    circ = AdjointCircuit(n, to_hip_ops(ops, cnt), cnt, Z0_PLUS_P0)
    th0 = np.array([ops[i].theta for i in O.param_indices(ops, cnt)])
    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    e = CircuitExpectation.apply(th, circ)
    (2.0 * e).backward()
    ps = O.param_shift_grad(n, ops, cnt, Z0_PLUS_P0)
    assert abs(e.item() - O.energy(n, ops, cnt, Z0_PLUS_P0)) <= 1e-12
    assert np.max(np.abs(th.grad.numpy() - 2.0 * ps)) <= 1e-10


def test_adjoint_errors():
    n = 4
    ops, cnt = O.ring_ops(n, 1, 3)
    hops = to_hip_ops(ops, cnt)
    with T.HipState() as s:
        s.init_basis(n, 0)
        g = np.zeros(8)
        arr, nt = T.make_terms(Z0_PLUS_P0)
        e = ctypes.c_double()
        rc = T.lib().tmfqs_hip_adjoint(s.handle, hops, cnt, None, None, arr, nt, 0, ctypes.byref(e),
                                       g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 3)
        assert rc == T.ERR_INVALID_ARGUMENT
        with pytest.raises(T.OutOfRange):
            s.adjoint(hops, cnt, [(T.DIAG_Z, 1.0, 1 << 4)])
        with pytest.raises(T.OutOfRange):
            s.adjoint(hops, cnt, [(T.DIAG_PROJECTOR, 1.0, 16)])
        bad = (T.Op * 1)()
        bad[0].kind = T.OP_DIFFUSE_MAT
        with pytest.raises(T.InvalidArgument):
            s.adjoint(bad, 1, Z0_PLUS_P0)
