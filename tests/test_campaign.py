"""Campaign runner host logic (tmfqsfullstate_amd/campaign.py), on the CPU:
the command lines it builds (scripts/run_thesis_experiments.py:216-280 in the
reference), the parsing of program output, resume, the combined and summary
tables, and the wall-time / peak-RSS measurement -- with stand-in programs in
place of the device programs, which the GPU tests run for real."""

import csv
import json
import os
import stat
import sys
import textwrap

import pytest

from tmfqsfullstate_amd import campaign

FAKE = {
    # Prints what the real grover programs print; allocates ~64 MiB first so
    # the peak-RSS column has something to see.
    "grover": """
        import sys
        block = bytearray(64 << 20)
        for i in range(0, len(block), 4096):
            block[i] = 1
        print("Grover mode: normal")
        print("Resolved strategy: hip")
        print("Grover search result: " + sys.argv[2].split(",")[0])
    """,
    "groverLazy": """
        import sys
        print("Grover mode: lazy")
        print("Grover search result: " + sys.argv[2].split(",")[-1])
    """,
    "qftG": """
        import os, sys
        if os.environ.get("FAKE_QFT_FAIL") and "random-phase" in sys.argv:
            print("qftG error: boom")
            sys.exit(2)
        if os.environ.get("FAKE_QFT_SLEEP"):
            import time
            time.sleep(30)
        print("Input family: pattern")
        print("Measured state (k): 8")
        print("Calculated r (N/k): 2")
    """,
    "experiment_compare": """
        import sys
        a = sys.argv
        n, w = a[a.index("--qubits") + 1], a[a.index("--workload") + 1]
        err = "1e-7" if "complex64" in a else "2e-15"
        print("qubits,workload,strategy,accuracy_kind,bitwise_equal,max_abs_amplitude_error,"
              "max_abs_component_error,rel_l2,rmse_amplitude,max_abs_probability_error,"
              "total_probability_diff,worst_state")
        print(f"{n},{w},hip,x,false,{err},{err},1e-9,1e-10,1e-11,1e-12,5")
    """,
}


@pytest.fixture()
def fake_bin(tmp_path):
    d = tmp_path / "bin"
    d.mkdir()
    for name, body in FAKE.items():
        p = d / name
        p.write_text(f"#!{sys.executable}\n" + textwrap.dedent(body))
        p.chmod(p.stat().st_mode | stat.S_IXUSR)
    return d


def test_qubit_and_name_parsing():
    assert campaign.parse_qubits("20-23") == [20, 21, 22, 23]
    assert campaign.parse_qubits("5,7") == [5, 7]
    with pytest.raises(Exception):
        campaign.parse_qubits("9-3")
    parse = campaign.parse_names(campaign.WORKLOADS)
    assert parse("qft_pattern,grover_multi") == ["qft_pattern", "grover_multi"]
    with pytest.raises(Exception):
        parse("qft_pattern,nope")


def test_marked_states_and_programs():
    # N/8, N/2, 7N/8 of the reference campaign.
    assert campaign.marked_states(6, "grover_single_n8") == "8"
    assert campaign.marked_states(6, "grover_lazy_single_n2") == "32"
    assert campaign.marked_states(6, "grover_single_7n8") == "56"
    assert campaign.marked_states(6, "grover_multi") == "8,32,56"
    assert campaign.marked_states(6, "qft_high_entropy") == ""
    assert [campaign.binary_for(w) for w in ("qft_pattern", "grover_multi", "grover_lazy_multi")] == [
        "qftG", "grover", "groverLazy"]
    assert campaign.family_for("grover_lazy_single_n8") == "grover_lazy"
    assert campaign.scenario_for_summary("grover_lazy_single_7n8") == "grover_lazy_single"
    assert campaign.scenario_for_summary("grover_multi") == "grover_multi"


def test_command_lines(tmp_path):
    cmd = campaign.resource_command(tmp_path, campaign.Job(7, "qft_high_entropy", "hip-complex64"), seed=3)
    assert cmd[1:] == ["7", "--input-family", "random-phase", "--strategy", "hip", "--precision", "complex64",
                       "--seed", "3"]
    cmd = campaign.resource_command(tmp_path, campaign.Job(5, "grover_multi", "hip-exact"), seed=None)
    assert os.path.basename(cmd[0]) == "grover" and cmd[1:] == ["5", "4,16,28", "--strategy", "hip", "--exact"]
    cmd = campaign.compare_command(tmp_path, campaign.Job(9, "grover_lazy_multi", "hip"))
    assert os.path.basename(cmd[0]) == "experiment_compare"
    assert cmd[1:] == ["--qubits", "9", "--workload", "grover_lazy_multi", "--strategy", "hip"]


def test_output_parsing():
    assert campaign.parse_measured_state("Grover mode: lazy\nGrover search result: 17\n") == "17"
    assert campaign.parse_measured_state("Measured state (k): 4\nCalculated r (N/k): 8\n") == "4"
    assert campaign.parse_measured_state("qftG error: x\n") == ""
    row = campaign.parse_compare_stdout("noise\nqubits,workload,bitwise_equal\n4,qft_pattern,true\n\n")
    assert row == {"qubits": "4", "workload": "qft_pattern", "bitwise_equal": "true"}
    assert campaign.parse_compare_stdout("only one line\n") == {}


def test_run_measured_reports_exit_rss_and_timeout(tmp_path, fake_bin, monkeypatch):
    out, err = tmp_path / "o", tmp_path / "e"
    code, secs, rss = campaign.run_measured([str(fake_bin / "grover"), "4", "3"], out, err, None)
    assert code == 0 and secs > 0 and rss > 60.0, (code, secs, rss)
    assert "Grover search result: 3" in out.read_text()
    monkeypatch.setenv("FAKE_QFT_SLEEP", "1")
    code, secs, _ = campaign.run_measured([str(fake_bin / "qftG"), "4"], out, err, 1.0)
    assert code == -9 and secs < 20


def read(path):
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def test_campaign_tables_and_resume(tmp_path, fake_bin, monkeypatch):
    out = tmp_path / "run"
    monkeypatch.setenv("FAKE_QFT_FAIL", "1")
    argv = ["--qubits", "4-5", "--bin-dir", str(fake_bin), "--output-dir", str(out), "--seed", "11"]
    assert campaign.main(argv) == 0
    runs, errors = read(out / "runs.csv"), read(out / "errors.csv")
    assert tuple(runs[0]) == campaign.RESOURCE_FIELDS
    assert tuple(errors[0]) == campaign.ERROR_FIELDS
    n_jobs = 2 * len(campaign.WORKLOADS) * len(campaign.STRATEGIES)
    assert len(runs) == n_jobs
    failed = [r for r in runs if r["exit_code"] != "0"]
    assert {r["workload"] for r in failed} == {"qft_high_entropy"} and len(failed) == 6
    grover = next(r for r in runs if r["run_id"] == "q5_grover_multi_hip")
    assert grover["marked_states"] == "4,16,28" and grover["measured_state"] == "4"
    assert "--seed 11" in grover["command"] and float(grover["max_memory_mb"]) > 60.0
    lazy = next(r for r in runs if r["run_id"] == "q4_grover_lazy_multi_hip-complex64")
    assert lazy["measured_state"] == "14" and lazy["family"] == "grover_lazy"
    ref = [e for e in errors if e["strategy"] == "hip-exact"]
    assert len(ref) == 2 * len(campaign.WORKLOADS)
    assert all(e["bitwise_equal"] == "true" and e["accuracy_kind"] == "reference" for e in ref)
    c64 = next(e for e in errors if e["strategy"] == "hip-complex64" and e["workload"] == "grover_multi")
    assert c64["accuracy_kind"] == "lossy" and float(c64["max_abs_amplitude_error"]) == 1e-7
    assert c64["worst_state"] == "5"

    combined = read(out / "combined.csv")
    assert len(combined) == len(runs)
    assert next(r for r in combined if r["run_id"] == "q4_qft_pattern_hip")["rel_l2"] == "1e-9"
    summary = read(out / "summary.csv")
    scen = {(r["qubits"], r["scenario"], r["strategy"]) for r in summary}
    assert ("5", "grover_single", "hip") in scen and ("4", "qft_pattern", "hip-exact") in scen
    assert not any(s[1] == "qft_high_entropy" for s in scen)  # every run failed
    ref_row = next(r for r in summary if r["scenario"] == "grover_multi" and r["strategy"] == "hip-exact" and r["qubits"] == "4")
    assert float(ref_row["relative_time_x"]) == pytest.approx(1.0)
    assert float(ref_row["memory_saving_percent"]) == pytest.approx(0.0)
    fused = next(r for r in summary if r["scenario"] == "grover_single" and r["strategy"] == "hip" and r["qubits"] == "5")
    assert float(fused["worst_error"]) == 2e-15
    manifest = json.loads((out / "manifest.json").read_text())
    assert manifest["qubits"] == [4, 5] and manifest["seed"] == 11
    assert manifest["strategies"]["hip-complex64"] == ["--strategy", "hip", "--precision", "complex64"]

    # Resume: only the failed jobs run again.
    monkeypatch.delenv("FAKE_QFT_FAIL")
    assert campaign.main(argv) == 0
    runs2 = read(out / "runs.csv")
    assert len(runs2) == n_jobs + 6
    assert all(r["exit_code"] == "0" for r in runs2[n_jobs:])
    summary2 = read(out / "summary.csv")
    assert any(r["scenario"] == "qft_high_entropy" for r in summary2)


def test_campaign_stops_on_failure_and_needs_programs(tmp_path, fake_bin, monkeypatch):
    monkeypatch.setenv("FAKE_QFT_FAIL", "1")
    argv = ["--qubits", "4", "--workloads", "qft_high_entropy", "--bin-dir", str(fake_bin),
            "--output-dir", str(tmp_path / "a"), "--stop-on-failure"]
    assert campaign.main(argv) == 1
    assert len(read(tmp_path / "a" / "runs.csv")) == 1
    assert campaign.main(["--qubits", "4", "--bin-dir", str(tmp_path), "--output-dir", str(tmp_path / "b")]) == 2
