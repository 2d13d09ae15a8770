"""Experiment campaign over the command-line programs: the reference's
experiment matrix runner (scripts/run_thesis_experiments.py in the
reference) for the device backend.

For every (qubits, workload, strategy) it runs one resource job -- the
grover / groverLazy / qftG program, timed on the wall clock, with the child's
peak resident set from ``os.wait4`` (the reference wraps ``/usr/bin/time`` or
psrecord, scripts/run_thesis_experiments.py:346-387) -- and one error job --
``experiment_compare`` against the exact complex128 run.  Rows are appended
to ``runs.csv`` and ``errors.csv`` with the reference's column sets, so a
campaign resumes where it stopped (rows with exit code 0 are kept) and the
reference's plotting scripts read the output; ``combined.csv``,
``summary.csv`` and ``manifest.json`` follow the same layout
(run_thesis_experiments.py:176-205, 503-563).

The strategies are device configurations in place of the reference's
dense / blosc / zfp:

  hip-exact       complex128, the dense backend's arithmetic bit for bit
                  (the reference row: errors are zero by construction)
  hip             complex128 with fused phase ladders (within 1e-12)
  hip-complex64   complex64 storage

``max_memory_mb`` is host memory; the state lives in device memory (16 or
8 bytes per amplitude), which the program's own process RSS does not show.

Run as ``python -m tmfqsfullstate_amd.campaign --qubits 20-28``.
"""

from __future__ import annotations

import argparse
import csv
import json
import os
import re
import shlex
import signal
import socket
import subprocess
import sys
import time
from dataclasses import dataclass
from datetime import datetime
from pathlib import Path
from typing import Iterable

PKG = Path(__file__).resolve().parent
DEFAULT_BIN = PKG / "bin"
DEFAULT_QUBITS = list(range(20, 29))
REFERENCE_STRATEGY = "hip-exact"
STRATEGIES = (REFERENCE_STRATEGY, "hip", "hip-complex64")
STRATEGY_FLAGS = {
    "hip-exact": ["--strategy", "hip", "--exact"],
    "hip": ["--strategy", "hip"],
    "hip-complex64": ["--strategy", "hip", "--precision", "complex64"],
}
ACCURACY_KIND = {"hip-exact": "reference", "hip": "fused", "hip-complex64": "lossy"}
WORKLOADS = (
    "qft_pattern",
    "qft_high_entropy",
    "grover_single_n8",
    "grover_single_n2",
    "grover_single_7n8",
    "grover_multi",
    "grover_lazy_single_n8",
    "grover_lazy_single_n2",
    "grover_lazy_single_7n8",
    "grover_lazy_multi",
)
# Column sets of scripts/run_thesis_experiments.py:39-86 (kept identical so
# the reference's readers of these files work unchanged).
RESOURCE_FIELDS = (
    "run_id", "qubits", "workload", "family", "strategy", "marked_states", "zfp_mode", "zfp_precision",
    "chunk_states", "cache_slots", "nthreads", "exit_code", "time_sec", "max_memory_mb", "measured_state",
    "stdout_path", "stderr_path", "profile_path", "command",
)
ERROR_FIELDS = (
    "qubits", "workload", "strategy", "accuracy_kind", "exit_code", "bitwise_equal", "max_abs_amplitude_error",
    "max_abs_component_error", "rel_l2", "rmse_amplitude", "max_abs_probability_error", "total_probability_diff",
    "worst_state", "stdout_path", "stderr_path", "command",
)
SUMMARY_FIELDS = (
    "qubits", "scenario", "strategy", "avg_time_sec", "avg_max_memory_mb", "memory_saving_percent",
    "relative_time_x", "avg_error", "worst_error",
)
METRIC_FIELDS = ERROR_FIELDS[5:13]


@dataclass(frozen=True)
class Job:
    qubits: int
    workload: str
    strategy: str

    @property
    def key(self) -> tuple[str, str, str]:
        return (str(self.qubits), self.workload, self.strategy)


def parse_qubits(value: str) -> list[int]:
    """"20-28" (inclusive range) or "20,22,24"."""
    if "-" in value:
        lo, hi = (int(v) for v in value.split("-", 1))
        if lo > hi:
            raise argparse.ArgumentTypeError("qubit range start must be <= end")
        return list(range(lo, hi + 1))
    return [int(v) for v in value.split(",") if v]


def parse_names(allowed: Iterable[str]):
    allowed = tuple(allowed)

    def parse(value: str) -> list[str]:
        names = [v for v in value.split(",") if v]
        bad = [v for v in names if v not in allowed]
        if bad or not names:
            raise argparse.ArgumentTypeError(f"unknown name(s) {bad}; choose from {', '.join(allowed)}")
        return names
    return parse


def family_for(workload: str) -> str:
    if workload.startswith("qft"):
        return "qft"
    return "grover_lazy" if workload.startswith("grover_lazy") else "grover"


def binary_for(workload: str) -> str:
    if workload.startswith("qft"):
        return "qftG"
    return "groverLazy" if workload.startswith("grover_lazy") else "grover"


def marked_states(qubits: int, workload: str) -> str:
    """The campaign's marked states N/8, N/2, 7N/8 (run_thesis_experiments.py:216-230)."""
    if workload.startswith("qft"):
        return ""
    n = 1 << qubits
    if workload.endswith("multi"):
        values = [n // 8, n // 2, (7 * n) // 8]
    elif workload.endswith("7n8"):
        values = [(7 * n) // 8]
    elif workload.endswith("n2"):
        values = [n // 2]
    else:
        values = [n // 8]
    return ",".join(str(v) for v in values)


def resource_command(bin_dir: Path, job: Job, seed: int | None) -> list[str]:
    cmd = [str(bin_dir / binary_for(job.workload)), str(job.qubits)]
    if job.workload == "qft_pattern":
        cmd += ["--input-family", "pattern"]
    elif job.workload == "qft_high_entropy":
        cmd += ["--input-family", "random-phase"]
    else:
        cmd.append(marked_states(job.qubits, job.workload))
    cmd += STRATEGY_FLAGS[job.strategy]
    if seed is not None:
        cmd += ["--seed", str(seed)]
    return cmd


def compare_command(bin_dir: Path, job: Job) -> list[str]:
    return [str(bin_dir / "experiment_compare"), "--qubits", str(job.qubits), "--workload", job.workload,
            *STRATEGY_FLAGS[job.strategy]]


def parse_measured_state(stdout: str) -> str:
    for pattern in (r"Grover search result:\s*(\d+)", r"Measured state \(k\):\s*(\d+)"):
        m = re.search(pattern, stdout)
        if m:
            return m.group(1)
    return ""


def parse_compare_stdout(stdout: str) -> dict[str, str]:
    """The header and row experiment_compare prints last."""
    lines = [line for line in stdout.splitlines() if line.strip()]
    if len(lines) < 2:
        return {}
    return next(csv.DictReader(lines[-2:]), {})


def run_measured(command: list[str], stdout_path: Path, stderr_path: Path,
                 timeout: float | None) -> tuple[int, float, float]:
    """Runs command; returns (exit code, wall seconds, peak RSS in MiB).

    The exit code is the process's own, or -signal when it was killed
    (-9 after a time-out)."""
    with stdout_path.open("w") as out, stderr_path.open("w") as err:
        start = time.perf_counter()
        proc = subprocess.Popen(command, stdout=out, stderr=err, start_new_session=True)
        while True:
            pid, status, usage = os.wait4(proc.pid, os.WNOHANG if timeout else 0)
            if pid == proc.pid:
                break
            if time.perf_counter() - start > timeout:
                os.killpg(proc.pid, signal.SIGKILL)
                pid, status, usage = os.wait4(proc.pid, 0)
                break
            time.sleep(0.05)
        elapsed = time.perf_counter() - start
    proc.returncode = os.waitstatus_to_exitcode(status)  # reaped here; keep Popen from waiting again
    return proc.returncode, elapsed, usage.ru_maxrss / 1024.0


def run_resource_job(job: Job, out_dir: Path, bin_dir: Path, seed: int | None,
                     timeout: float | None) -> dict[str, object]:
    run_id = f"q{job.qubits}_{job.workload}_{job.strategy}"
    stdout_path = out_dir / "logs" / f"{run_id}.stdout.log"
    stderr_path = out_dir / "logs" / f"{run_id}.stderr.log"
    stdout_path.parent.mkdir(parents=True, exist_ok=True)
    command = resource_command(bin_dir, job, seed)
    print(f"[resource] {run_id}", flush=True)
    code, seconds, rss_mb = run_measured(command, stdout_path, stderr_path, timeout)
    return {
        "run_id": run_id, "qubits": job.qubits, "workload": job.workload, "family": family_for(job.workload),
        "strategy": job.strategy, "marked_states": marked_states(job.qubits, job.workload),
        "zfp_mode": "", "zfp_precision": "", "chunk_states": "", "cache_slots": "", "nthreads": "",
        "exit_code": code, "time_sec": f"{seconds:.6f}", "max_memory_mb": f"{rss_mb:.6f}",
        "measured_state": parse_measured_state(stdout_path.read_text(errors="replace")),
        "stdout_path": str(stdout_path), "stderr_path": str(stderr_path), "profile_path": "",
        "command": shlex.join(command),
    }


def reference_error_row(qubits: int, workload: str) -> dict[str, object]:
    """The exact run compared with itself (run_thesis_experiments.py:467-485)."""
    row: dict[str, object] = {f: 0 for f in METRIC_FIELDS}
    row.update({"qubits": qubits, "workload": workload, "strategy": REFERENCE_STRATEGY,
                "accuracy_kind": "reference", "exit_code": 0, "bitwise_equal": "true",
                "stdout_path": "", "stderr_path": "", "command": "reference row"})
    return row


def run_error_job(job: Job, out_dir: Path, bin_dir: Path, timeout: float | None) -> dict[str, object]:
    run_id = f"q{job.qubits}_{job.workload}_{job.strategy}_error"
    stdout_path = out_dir / "logs" / f"{run_id}.stdout.log"
    stderr_path = out_dir / "logs" / f"{run_id}.stderr.log"
    stdout_path.parent.mkdir(parents=True, exist_ok=True)
    command = compare_command(bin_dir, job)
    print(f"[error] {run_id}", flush=True)
    code, _, _ = run_measured(command, stdout_path, stderr_path, timeout)
    parsed = parse_compare_stdout(stdout_path.read_text(errors="replace")) if code == 0 else {}
    row: dict[str, object] = {"qubits": job.qubits, "workload": job.workload, "strategy": job.strategy,
                              "accuracy_kind": ACCURACY_KIND[job.strategy], "exit_code": code,
                              "stdout_path": str(stdout_path), "stderr_path": str(stderr_path),
                              "command": shlex.join(command)}
    for f in METRIC_FIELDS:
        row[f] = parsed.get(f, "")
    return row


def append_csv(path: Path, fields: Iterable[str], row: dict[str, object]) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    header = not path.exists() or path.stat().st_size == 0
    with path.open("a", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(fields), extrasaction="ignore")
        if header:
            w.writeheader()
        w.writerow(row)
        fh.flush()
        os.fsync(fh.fileno())


def read_rows(path: Path) -> list[dict[str, str]]:
    if not path.exists():
        return []
    with path.open(newline="") as fh:
        return list(csv.DictReader(fh))


def write_rows(path: Path, fields: Iterable[str], rows: list[dict[str, object]]) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(fields), extrasaction="ignore")
        w.writeheader()
        w.writerows(rows)


def completed(path: Path) -> set[tuple[str, str, str]]:
    return {(r["qubits"], r["workload"], r["strategy"]) for r in read_rows(path) if r.get("exit_code") == "0"}


def generate_combined(out_dir: Path) -> None:
    """runs.csv rows joined with their errors.csv metrics (run_thesis_experiments.py:503-515).

    exit_code stays the resource run's (the reference's join lets the error
    job's overwrite it, so a failed run whose comparison passed would count
    in the summary); a failed error job leaves its metric columns empty."""
    errors = {(r["qubits"], r["workload"], r["strategy"]): r for r in read_rows(out_dir / "errors.csv")}
    merged = []
    for row in read_rows(out_dir / "runs.csv"):
        e = errors.get((row["qubits"], row["workload"], row["strategy"]), {})
        if e.get("exit_code") != "0":
            e = {}
        m = dict(row)
        for f in ERROR_FIELDS:
            if f not in ("qubits", "workload", "strategy", "exit_code", "stdout_path", "stderr_path", "command"):
                m[f] = e.get(f, "")
        merged.append(m)
    write_rows(out_dir / "combined.csv", list(RESOURCE_FIELDS) + [f for f in ERROR_FIELDS if f not in RESOURCE_FIELDS],
               merged)


def scenario_for_summary(workload: str) -> str:
    if workload.startswith("grover_lazy_single"):
        return "grover_lazy_single"
    if workload.startswith("grover_single"):
        return "grover_single"
    return workload


def _mean(values: list[float]) -> float | None:
    return sum(values) / len(values) if values else None


def generate_summary(out_dir: Path) -> None:
    """Per (qubits, scenario, strategy): mean time and memory, both relative to
    the exact run's, and the mean and worst amplitude error; the single-marked
    scenarios only once all three marked positions ran
    (run_thesis_experiments.py:526-563)."""
    groups: dict[tuple[str, str, str], list[dict[str, str]]] = {}
    for row in read_rows(out_dir / "combined.csv"):
        if row.get("exit_code") == "0":
            groups.setdefault((row["qubits"], scenario_for_summary(row["workload"]), row["strategy"]), []).append(row)
    summary = []
    for (qubits, scenario, strategy), rows in sorted(groups.items(), key=lambda kv: (int(kv[0][0]), kv[0][1], kv[0][2])):
        if scenario in ("grover_single", "grover_lazy_single") and len(rows) < 3:
            continue
        times = [float(r["time_sec"]) for r in rows if r.get("time_sec")]
        mems = [float(r["max_memory_mb"]) for r in rows if r.get("max_memory_mb")]
        errs = [float(r.get("max_abs_amplitude_error") or 0.0) for r in rows]
        if not times or not mems:
            continue
        ref = groups.get((qubits, scenario, REFERENCE_STRATEGY), [])
        ref_time = _mean([float(r["time_sec"]) for r in ref if r.get("time_sec")])
        ref_mem = _mean([float(r["max_memory_mb"]) for r in ref if r.get("max_memory_mb")])
        t, m = _mean(times), _mean(mems)
        summary.append({
            "qubits": qubits, "scenario": scenario, "strategy": strategy,
            "avg_time_sec": f"{t:.6f}", "avg_max_memory_mb": f"{m:.6f}",
            "memory_saving_percent": "" if not ref_mem else f"{100.0 * (1.0 - m / ref_mem):.6f}",
            "relative_time_x": "" if not ref_time else f"{t / ref_time:.6f}",
            "avg_error": f"{sum(errs) / len(errs):.17g}", "worst_error": f"{max(errs):.17g}",
        })
    write_rows(out_dir / "summary.csv", SUMMARY_FIELDS, summary)


def git_metadata() -> dict[str, str]:
    def read(cmd: list[str]) -> str:
        try:
            return subprocess.check_output(cmd, cwd=PKG, text=True, stderr=subprocess.DEVNULL).strip()
        except (OSError, subprocess.SubprocessError):
            return ""
    return {"commit": read(["git", "rev-parse", "HEAD"]), "branch": read(["git", "rev-parse", "--abbrev-ref", "HEAD"]),
            "status_short": read(["git", "status", "--short"])}


def write_manifest(out_dir: Path, args: argparse.Namespace) -> None:
    manifest = {
        "started_at": datetime.now().isoformat(timespec="seconds"),
        "host": socket.gethostname(),
        "bin_dir": str(args.bin_dir),
        "qubits": args.qubits,
        "strategies": {s: STRATEGY_FLAGS[s] for s in args.strategies},
        "workloads": list(args.workloads),
        "profiler": "wait4",
        "seed": args.seed,
        "timeout_sec": args.timeout,
        "git": git_metadata(),
    }
    (out_dir / "manifest.json").write_text(json.dumps(manifest, indent=2) + "\n")


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(description="Run the experiment matrix over the command-line programs.")
    p.add_argument("--qubits", type=parse_qubits, default=DEFAULT_QUBITS, help="range like 20-28 or list like 20,21")
    p.add_argument("--workloads", type=parse_names(WORKLOADS), default=list(WORKLOADS))
    p.add_argument("--strategies", type=parse_names(STRATEGIES), default=list(STRATEGIES))
    p.add_argument("--output-dir", default="", help="default: results/campaign/<timestamp> under the working directory")
    p.add_argument("--bin-dir", type=Path, default=DEFAULT_BIN)
    p.add_argument("--seed", type=int, default=None, help="measurement seed passed to every program")
    p.add_argument("--timeout", type=float, default=None, help="seconds per program run")
    p.add_argument("--stop-on-failure", action="store_true")
    p.add_argument("--skip-errors", action="store_true")
    args = p.parse_args(argv)

    out_dir = Path(args.output_dir) if args.output_dir else Path("results/campaign") / datetime.now().strftime("%Y%m%d-%H%M%S")
    out_dir.mkdir(parents=True, exist_ok=True)
    for prog in ("grover", "groverLazy", "qftG", "experiment_compare"):
        if not (args.bin_dir / prog).is_file():
            print(f"missing program {args.bin_dir / prog}: build the package first", file=sys.stderr)
            return 2
    write_manifest(out_dir, args)
    runs_csv, errors_csv = out_dir / "runs.csv", out_dir / "errors.csv"
    done_runs, done_errors = completed(runs_csv), completed(errors_csv)

    for q in args.qubits:
        for workload in args.workloads:
            for strategy in args.strategies:
                job = Job(q, workload, strategy)
                if job.key not in done_runs:
                    row = run_resource_job(job, out_dir, args.bin_dir, args.seed, args.timeout)
                    append_csv(runs_csv, RESOURCE_FIELDS, row)
                    if row["exit_code"] != 0 and args.stop_on_failure:
                        return 1
                    if row["exit_code"] == 0:
                        done_runs.add(job.key)
                if args.skip_errors or job.key in done_errors:
                    continue
                if strategy == REFERENCE_STRATEGY:
                    row = reference_error_row(q, workload)
                else:
                    row = run_error_job(job, out_dir, args.bin_dir, args.timeout)
                append_csv(errors_csv, ERROR_FIELDS, row)
                if row["exit_code"] != 0 and args.stop_on_failure:
                    return 1
                if row["exit_code"] == 0:
                    done_errors.add(job.key)
            generate_combined(out_dir)
            generate_summary(out_dir)
    generate_combined(out_dir)
    generate_summary(out_dir)
    print(f"Experiment outputs: {out_dir}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
