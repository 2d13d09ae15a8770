"""GPU tests through the public C++ API (QuantumRegister / sim::Simulator with
the Hip strategy), via include/tmfqs_sim.h and the C++ test drivers.

Tolerances: complex128 amplitudes within 1e-12 of the reference dense
backend's (the CPU oracle, itself bit-exact with the reference build);
measured indices equal.
"""

import os
import subprocess

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tmfqsfullstate_amd import sim as S
from tests import numpy_ref as R

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(T.__file__), "bin")
TOL = 1e-12


def cplx(x):
    return x[0::2] + 1j * x[1::2]


@pytest.mark.parametrize("exe", ["test_hip_circuit", "test_hip_register"])
def test_cpp_drivers(exe):
    p = subprocess.run([os.path.join(BIN, exe)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert " 0 failures" in p.stdout and "FAIL" not in p.stdout


def test_circuit_builders_match_oracle():
    for n in (1, 5, 12, 30):
        a, na = S.qft_ops(n)
        b, nb = O.qft_ops(n)
        assert na == nb
        for i in range(na):
            assert (a[i].kind, a[i].q0, a[i].q1, a[i].theta) == (b[i].kind, b[i].q0, b[i].q1, b[i].theta)
    a, na = S.grover_ops(10, [7, 3, 7], True)
    b, nb = O.grover_ops(10, [7, 3, 7], True)
    assert na == nb and all(a[i].kind == b[i].kind and a[i].state == b[i].state for i in range(na))


@pytest.mark.parametrize("materialized", [True, False])
def test_simulator_grover(materialized):
    for n, marked in ((10, [511]), (13, [1, 4096, 8191])):
        ops, cnt = S.grover_ops(n, marked, materialized)
        reg, secs = S.simulate(n, ops, cnt)
        assert secs > 0
        want = O.run_circuit(n, O.grover_ops(n, marked, materialized))
        # sums, means, diffusion steps and the overlay reproduce the reference bit for bit
        assert np.array_equal(reg.export(), want.amplitudes())
        for seed in (123456, 7, 8):
            assert reg.measure_seed(seed) == want.measure(O.mt_unit_draws(seed, 1)[0])
        assert reg.total_probability() == want.total_probability()


def test_simulator_qft_random_phase_and_exact_mode():
    n = 14
    init = O.random_phase(n, 7)
    ops, cnt = S.qft_ops(n)
    want = O.run_circuit(n, O.qft_ops(n), init=init).amplitudes()
    reg, _ = S.simulate(n, ops, cnt, init=init)
    assert np.abs(reg.amplitudes() - cplx(want)).max() <= TOL
    exact, _ = S.simulate(n, ops, cnt, init=init, exact_phases=1)
    assert np.array_equal(exact.export(), want)
    rp = S.Register.random_phase(12, 7)
    assert np.array_equal(rp.export(), O.random_phase(12, 7))


def test_register_matrix_gates_and_fixed_measure():
    n = 8
    rng = np.random.default_rng(3)
    init = O.random_phase(n, 5)
    pool, mq, seq = [], [], []
    for k in (1, 2, 3, 5):
        u = R.random_unitary(rng, 1 << k)
        qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
        seq.append((T.OP_MATRIX, len(mq), 0, 0.0, sum(len(p) for p in pool), k))
        pool.append(R.interleave(u))
        mq += qs
    pool = np.concatenate(pool)
    mq = np.array(mq, dtype=np.uint32)
    reg = S.Register.from_amplitudes(n, init)
    ops, cnt = S.make_ops(seq)
    reg.run(ops, cnt, mat_pool=pool, mat_qubits=mq)
    oops, ocnt = O.make_ops(seq)
    want = O.run_circuit(n, (oops, ocnt), init=init, mat_pool=pool, mat_qubits=mq)
    assert np.array_equal(reg.export(), want.amplitudes())
    tp = want.total_probability()
    assert reg.measure_fixed(tp / 2) == want.measure(tp / 2)


def test_overlay_paths_match_oracle():
    # lazy Grover leaves the affine overlay active: amplitudes, total
    # probability and measurement go through the overlay-aware device reads
    n = 11
    ops, cnt = S.grover_ops(n, [1000], False)
    reg, _ = S.simulate(n, ops, cnt)
    want = O.run_circuit(n, O.grover_ops(n, [1000], False))
    assert np.abs(reg.amplitudes() - cplx(want.amplitudes())).max() <= TOL
    assert reg.total_probability() == want.total_probability()
    for rnd in (0.0001, 0.3, 0.99999):
        assert reg.measure_fixed(rnd) == want.measure(rnd)


def test_device_compare_and_complex64():
    n = 16
    init = O.random_phase(n, 1)
    ops, cnt = S.qft_ops(n)
    a, _ = S.simulate(n, ops, cnt, init=init)
    b, _ = S.simulate(n, ops, cnt, init=init, precision=T.C64)
    c = a.hip_state().compare(b.hip_state())
    host_a, host_b = a.amplitudes(), b.amplitudes()
    assert abs(c["max_abs_amplitude_error"] - np.abs(host_a - host_b).max()) < 1e-15
    fid = abs(np.vdot(host_a, host_b)) ** 2 / (np.vdot(host_a, host_a).real * np.vdot(host_b, host_b).real)
    assert abs(c["fidelity"] - fid) < 1e-12
    assert 1 - c["fidelity"] <= 1e-6
    same = a.hip_state().compare(a.clone().hip_state())
    assert same["bitwise_equal"] == 1 and same["max_abs_amplitude_error"] == 0.0


def test_errors_map_to_exceptions():
    with pytest.raises(T.OutOfRange):
        S.simulate(3, [(T.OP_H, 3)])
    with pytest.raises(T.InvalidArgument):
        S.simulate(3, [(T.OP_CX, 1, 1)])
    with pytest.raises(T.InvalidArgument):
        S.grover_ops(3, [8], True)
    with pytest.raises(T.InvalidArgument):
        S.Register.basis(64)


def test_config_qft26_random_phase_vs_fft():
    # QFT-26 on QuantumState::randomPhase(26, 7) (the QFT parity workload) against the DFT
    n = 26
    print("QFT-26 randomPhase(26, 7)", flush=True)
    reg = S.Register.random_phase(n, 7)
    ops, cnt = S.qft_ops(n)
    reg.run(ops, cnt)
    got = reg.amplitudes()
    init = S.random_phase(n, 7)
    want = np.fft.ifft(init[0::2] + 1j * init[1::2], norm="ortho")
    assert np.abs(got - want).max() <= 1e-12


def test_config_grover20_bitexact():
    # Grover n=20, marked 524287, 804 iterations, materialized diffusion; measure(Mt19937(123456))
    n, marked = 20, 524287
    print("Grover-20 vs oracle", flush=True)
    ops, cnt = S.grover_ops(n, [marked], True)
    reg, _ = S.simulate(n, ops, cnt)
    oops, ocnt = O.grover_ops(n, [marked], True)
    ref = O.run_circuit(n, (oops, ocnt))
    assert np.array_equal(reg.export(), ref.amplitudes())
    p = abs(reg.amplitude(marked)) ** 2
    assert abs(p - 0.999999756965361) < 1e-12
    draw = O.mt_unit_draws(123456, 1)[0]
    assert reg.measure_seed(123456) == ref.measure(draw) == marked


def test_trace_report(tmp_path, monkeypatch):
    # TMFQS_HIP_TRACE=<file>: each handle appends one JSON line of its kernels
    # when it is destroyed.
    import json

    path = tmp_path / "trace.jsonl"
    monkeypatch.setenv("TMFQS_HIP_TRACE", str(path))
    n = 16
    ops, cnt = S.qft_ops(n)
    print("trace QFT-16", flush=True)
    s = T.HipState()
    s.init_basis(n, 5)
    s.run(ops, cnt)
    got = cplx(s.export())
    s.close()
    monkeypatch.delenv("TMFQS_HIP_TRACE")
    want = np.fft.ifft(np.eye(1 << n)[5]) * np.sqrt(1 << n)
    assert np.max(np.abs(got - want)) < TOL
    lines = path.read_text().splitlines()
    assert len(lines) == 1
    rec = json.loads(lines[0])
    assert rec["qubits"] == n and rec["precision"] == "complex128"
    names = {k["name"]: k for k in rec["kernels"]}
    # the default QFT plan at n = 16: one six-target block pass and the
    # ten-target tail fused with the bit reversal
    assert set(names) == {"k_qft_wide_block", "k_qft_wide_tail"}
    assert names["k_qft_wide_block"]["launches"] == 1
    amp_bytes = 16 << n
    for k in rec["kernels"]:
        assert k["launches"] >= 1 and k["ms"] > 0 and k["bytes"] > 0
        assert k["fraction_of_hbm_peak"] == pytest.approx(k["gb_per_s"] / 8000, rel=1e-3, abs=1e-5)
    assert names["k_qft_wide_tail"]["bytes"] == 2 * amp_bytes
    assert rec["total_ms"] == pytest.approx(sum(k["ms"] for k in rec["kernels"]), rel=1e-5)


def sequential_metrics(ref, cand):
    """experiment::compareStates (reference comparison.cpp:12-63) in state
    order, restated in Python on interleaved vectors."""
    r = ref[0::2] + 1j * ref[1::2]
    c = cand[0::2] + 1j * cand[1::2]
    m = dict(bitwise_equal=int(np.array_equal(ref, cand)), max_abs_amplitude_error=0.0, max_abs_component_error=0.0,
             max_abs_probability_error=0.0, worst_state=0)
    d2 = r2 = rt = ct = 0.0
    for s in range(r.size):
        dr, di = c[s].real - r[s].real, c[s].imag - r[s].imag
        sd = dr * dr + di * di
        ae = sd ** 0.5
        if ae > m["max_abs_amplitude_error"]:
            m["max_abs_amplitude_error"], m["worst_state"] = ae, s
        m["max_abs_component_error"] = max(m["max_abs_component_error"], abs(dr), abs(di))
        d2 += sd
        rp = r[s].real * r[s].real + r[s].imag * r[s].imag
        cp = c[s].real * c[s].real + c[s].imag * c[s].imag
        r2 += rp
        rt += rp
        ct += cp
        m["max_abs_probability_error"] = max(m["max_abs_probability_error"], abs(cp - rp))
    m["rmse_amplitude"] = (d2 / r.size) ** 0.5
    m["rel_l2"] = 0.0 if r2 == 0 else (d2 / r2) ** 0.5
    m["total_probability_diff"] = abs(ct - rt)
    return m


FIELDS = ["bitwise_equal", "max_abs_amplitude_error", "max_abs_component_error", "rmse_amplitude", "rel_l2",
          "max_abs_probability_error", "total_probability_diff", "worst_state"]


def test_compare_states_all_metrics():
    # every StateComparison field and worstState: the device path (two Hip
    # simulation results) against a sequential restatement (maxima and the
    # worst state exactly, sums to rounding); the host path (vectors)
    # bit for bit
    n = 12
    init = O.random_phase(n, 21)
    ops, cnt = S.qft_ops(n)
    a, _ = S.simulate(n, ops, cnt, init=init)
    b, _ = S.simulate(n, ops, cnt, init=init, precision=T.C64)
    ra, rb = a.export(), b.export()
    want = sequential_metrics(ra, rb)
    assert want["bitwise_equal"] == 0 and want["max_abs_amplitude_error"] > 0
    dev = S.compare_runs(n, ops, cnt, ref={}, cand={"precision": T.C64}, init=init)
    host = S.compare_amplitudes(n, ra, rb)
    for f in FIELDS:
        assert host[f] == want[f], f
        if f in ("bitwise_equal", "worst_state", "max_abs_amplitude_error", "max_abs_component_error",
                 "max_abs_probability_error"):
            assert dev[f] == want[f], f
        else:
            assert abs(dev[f] - want[f]) <= 1e-13 * max(1.0, abs(want[f])), (f, dev[f], want[f])
    same = S.compare_runs(n, ops, cnt, ref={}, cand={}, init=init)
    assert same["bitwise_equal"] == 1 and all(same[f] == 0 for f in FIELDS[1:])
    # a single perturbed amplitude: it is the worst state
    rc = ra.copy()
    rc[2 * 777] += 3e-9
    host = S.compare_amplitudes(n, ra, rc)
    assert host["worst_state"] == 777 and host["bitwise_equal"] == 0
    assert host == {**host, **sequential_metrics(ra, rc)}


def test_traced_run_times_each_operation_on_the_device():
    # per-operation RunReport seconds from stream events: one entry per
    # operation, all positive, their sum within the run's execution seconds
    n = 22
    ops, cnt = S.qft_ops(n)
    reg = S.Register.basis(n, 3)
    per, secs = reg.run_traced(ops, cnt)
    assert per.shape == (cnt,) and np.all(per > 0)
    assert per.sum() <= secs
    want = O.run_circuit(n, O.qft_ops(n), init=_basis(n, 3)).amplitudes()
    assert np.abs(reg.export() - want).max() <= 1e-12  # traced runs compute the same state


def _basis(n, s):
    v = np.zeros(2 << n)
    v[2 * s] = 1.0
    return v
