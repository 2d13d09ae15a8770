"""The command-line programs on the device (reference examples/grover.cpp,
groverLazy.cpp, qftG.cpp, experiment_compare.cpp): their output lines, the
measured state against the oracle's register measured with the same
Mt19937 draw, the compare rows against the oracle-pinned tolerances, the
JSON report, and one small campaign through all of them."""

import csv
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from tmfqsfullstate_amd import campaign

pytestmark = pytest.mark.gpu

BIN = campaign.DEFAULT_BIN


def run(prog, *args, expect=0):
    p = subprocess.run([str(BIN / prog), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert p.returncode == expect, (prog, args, p.returncode, p.stdout, p.stderr)
    return p.stdout


def field(out, label):
    for line in out.splitlines():
        if line.startswith(label + ":"):
            return line.split(":", 1)[1].strip()
    raise AssertionError(f"no '{label}' line in:\n{out}")


def oracle_measure(n, ops, seed, init=None):
    reg = oracle.run_circuit(n, ops, init=init)
    return reg.measure(float(oracle.mt_unit_draws(seed, 1)[0]))


@pytest.mark.parametrize("prog,materialized", [("grover", True), ("groverLazy", False)])
@pytest.mark.parametrize("n,marked", [(9, [64]), (10, [128, 512, 896])])
def test_grover_programs_measure_the_oracle_state(prog, materialized, n, marked):
    for seed in (1, 7, 12345):
        out = run(prog, n, ",".join(map(str, marked)), "--seed", seed, "--verbose")
        assert field(out, "Grover mode") == ("normal" if materialized else "lazy")
        assert field(out, "Resolved strategy") == "hip"
        got = int(field(out, "Grover search result"))
        assert got == oracle_measure(n, oracle.grover_ops(n, marked, materialized), seed), (prog, n, seed)
        assert got in marked
        ops, cnt = oracle.grover_ops(n, marked, materialized)
        assert int(field(out, "Circuit operations")) == cnt
        assert float(field(out, "Marked-state probability")) > 0.99
        assert float(field(out, "Execution seconds")) > 0


def test_programs_on_the_chunked_strategy():
    """grover and qftG with --strategy hip-chunked (2^6-state host chunks,
    windows of 4 chunks): the measured state equals the oracle's."""
    n, marked = 10, [128, 512, 896]
    for seed in (1, 7):
        out = run("grover", n, ",".join(map(str, marked)), "--strategy", "hip-chunked", "--chunk-states", 64,
                  "--cache-slots", 4, "--seed", seed, "--verbose")
        assert field(out, "Resolved strategy") == "hip-chunked"
        got = int(field(out, "Grover search result"))
        assert got == oracle_measure(n, oracle.grover_ops(n, marked, True), seed)
    out = run("qftG", 11, "--input-family", "random-phase", "--strategy", "hip-chunked", "--chunk-states", 128,
              "--cache-slots", 2, "--seed", 3)
    k = int(field(out, "Measured state (k)"))
    assert k == oracle_measure(11, oracle.qft_ops(11), 3, init=oracle.random_phase(11, 123456))
    run("grover", n, "128", "--strategy", "hip-chunked", "--chunk-states", 100, expect=2)


def pattern_input(n):
    states = [8 * x + (x % 2) for x in range(1 << n) if 8 * x + (x % 2) < (1 << n)]
    a = np.zeros(2 << n)
    a[2 * np.array(states)] = 1.0 / np.sqrt(len(states))
    return a


@pytest.mark.parametrize("family", ["pattern", "random-phase"])
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_qftg_measures_the_oracle_state(family, precision):
    n = 11
    init = pattern_input(n) if family == "pattern" else oracle.random_phase(n, 123456)
    for seed in (3, 99):
        out = run("qftG", n, "--input-family", family, "--precision", precision, "--seed", seed)
        assert field(out, "Input family") == family
        if family == "random-phase":
            assert field(out, "Random seed") == "123456"
        k = int(field(out, "Measured state (k)"))
        want = oracle_measure(n, oracle.qft_ops(n), seed, init=init)
        if precision == "complex128":
            assert k == want, (family, seed)
        else:
            # complex64 cumulative sums can cross a bucket edge the complex128
            # ones do not: the outcome must still be a state the oracle gives
            # weight to.
            amps = oracle.run_circuit(n, oracle.qft_ops(n), init=init).amplitudes()
            assert amps[2 * k] ** 2 + amps[2 * k + 1] ** 2 > 1e-9
        r = field(out, "Calculated r (N/k)")
        if k == 0:
            assert r == "undefined (division by zero)"
        else:
            assert float(r) == pytest.approx((1 << n) / k, rel=1e-5)


def compare_row(out):
    return next(csv.DictReader(out.strip().splitlines()[-2:]))


@pytest.mark.parametrize("workload", campaign.WORKLOADS)
def test_experiment_compare_rows(workload):
    n = 9
    exact = compare_row(run("experiment_compare", "--qubits", n, "--workload", workload, "--exact"))
    assert exact["bitwise_equal"] == "true" and float(exact["max_abs_amplitude_error"]) == 0.0
    assert exact["accuracy_kind"] == "reference" and exact["workload"] == workload and exact["qubits"] == str(n)
    fused = compare_row(run("experiment_compare", "--qubits", n, "--workload", workload))
    assert fused["strategy"] == "hip" and fused["accuracy_kind"] == "fused"
    assert float(fused["max_abs_amplitude_error"]) <= 1e-12
    c64 = compare_row(run("experiment_compare", "--qubits", n, "--workload", workload, "--precision", "complex64"))
    assert c64["strategy"] == "hip-complex64" and c64["accuracy_kind"] == "lossy"
    assert 0.0 < float(c64["max_abs_amplitude_error"]) <= 1e-5
    assert float(c64["rel_l2"]) <= 1e-5 and 0 <= int(c64["worst_state"]) < (1 << n)


def test_report_json_has_device_times(tmp_path):
    path = tmp_path / "report.json"
    out = run("qftG", 12, "--report-json", path, "--seed", 1)
    assert "Measured state (k)" in out
    rep = json.loads(path.read_text())
    ops, cnt = oracle.qft_ops(12)
    assert rep["qubits"] == 12 and rep["operation_count"] == cnt and rep["strategy"] == "hip"
    assert len(rep["operations"]) == cnt
    assert all(o["seconds"] >= 0 for o in rep["operations"])
    assert sum(o["seconds"] for o in rep["operations"]) > 0


def test_program_errors_on_the_device():
    out = run("grover", 4, 3, "--strategy", "dense", expect=2)
    assert "grover error:" in out and "Available strategies:" in out and "hip" in field(out, "Available strategies")
    out = run("qftG", 5, "--fuse-qubits", 9, expect=2)
    assert "qftG error:" in out


def test_small_campaign_end_to_end(tmp_path):
    out_dir = tmp_path / "campaign"
    rc = campaign.main(["--qubits", "8", "--workloads", "qft_high_entropy,grover_multi,grover_lazy_single_n2",
                        "--output-dir", str(out_dir), "--seed", "5", "--timeout", "120"])
    assert rc == 0
    with open(out_dir / "runs.csv", newline="") as fh:
        runs = list(csv.DictReader(fh))
    assert len(runs) == 9 and all(r["exit_code"] == "0" for r in runs)
    for r in runs:
        if r["workload"] == "grover_multi":
            assert int(r["measured_state"]) in (32, 128, 224)
        if r["workload"] == "grover_lazy_single_n2":
            assert r["measured_state"] == "128"
        assert float(r["time_sec"]) > 0 and float(r["max_memory_mb"]) > 1
    with open(out_dir / "errors.csv", newline="") as fh:
        errors = list(csv.DictReader(fh))
    assert len(errors) == 9 and all(e["exit_code"] == "0" for e in errors)
    for e in errors:
        bound = {"hip-exact": 0.0, "hip": 1e-12, "hip-complex64": 1e-5}[e["strategy"]]
        assert float(e["max_abs_amplitude_error"]) <= bound, e
    with open(out_dir / "summary.csv", newline="") as fh:
        summary = list(csv.DictReader(fh))
    assert {(s["scenario"], s["strategy"]) for s in summary} == {
        (w, s) for w in ("qft_high_entropy", "grover_multi") for s in campaign.STRATEGIES}
    assert os.path.isfile(out_dir / "manifest.json")
