"""HBM traffic per kernel launch from two rocprofv3 counter runs.

    rocprofv3 --pmc FETCH_SIZE -f csv -d <fetch_dir> -o fetch -- python -u bench.py <legs>
    rocprofv3 --pmc WRITE_SIZE -f csv -d <write_dir> -o write -- python -u bench.py <legs>
    python -m tmfqsfullstate_amd.pmc_traffic --fetch <fetch_dir> --write <write_dir> --qubits 30 \\
        --out profiles/traffic.json --markdown <summary.md>

FETCH_SIZE and WRITE_SIZE are KiB per dispatch.  Both are scaled by what
they read for the library's copy kernel (k_copy<double>, which moves exactly
2^n * 16 bytes each way in the bench's copy probe), and averaged per kernel
and precision: workload qft<n>_c128 collects the double instantiations,
qft<n>_c64 the float ones.  The output records the hash of the kernel
sources it was measured on (kernel_sources_sha), so a reader can tell when
a later kernel change has made it stale.
"""

from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import re
import sys
from collections import defaultdict

PKG = os.path.dirname(os.path.abspath(__file__))


def kernel_sources_sha() -> str:
    """SHA-256 (first 16 hex digits) of the HIP kernel sources (csrc/*.hip, *.h)."""
    csrc = os.path.join(PKG, "csrc")
    h = hashlib.sha256()
    for f in sorted(os.listdir(csrc)):
        if f.endswith((".hip", ".h")):
            h.update(f.encode())
            with open(os.path.join(csrc, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def short_name(kernel: str) -> tuple[str, str | None]:
    """('k_qft_block', 'c128') from 'void tmfqs_hip::k_qft_block<double, 5, false>(...)'."""
    head = kernel.replace("(anonymous namespace)::", "").split("(", 1)[0]
    m = re.search(r"(\w+)\s*(<.*)?$", head.replace("void ", ""))
    name = m.group(1) if m else head
    targs = head[len(head.split("<", 1)[0]):]
    prec = "c128" if "double" in targs else ("c64" if "float" in targs else None)
    return name, prec


def read_counter(run_dir: str, counter: str) -> dict[tuple[str, str | None], list[float]]:
    """Per (kernel, precision): the counter's values (bytes) of every dispatch."""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *counter_collection.csv under {run_dir}")
    out: dict[tuple[str, str | None], list[float]] = defaultdict(list)
    for path in files:
        with open(path, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                out[short_name(row["Kernel_Name"])].append(float(row["Counter_Value"]) * 1024.0)
    return out


def summarize(fetch_dir: str, write_dir: str, qubits: int, commands: list[str]) -> dict:
    fetch = read_counter(fetch_dir, "FETCH_SIZE")
    write = read_counter(write_dir, "WRITE_SIZE")
    copy_key = ("k_copy", "c128")
    if copy_key not in fetch or copy_key not in write:
        raise ValueError("the runs hold no k_copy<double> dispatch to calibrate against")
    copy_bytes = float((1 << qubits) * 16)
    mean = lambda v: sum(v) / len(v)
    fscale = copy_bytes / mean(fetch[copy_key])
    wscale = copy_bytes / mean(write[copy_key])
    result: dict = {
        "kernel_sources_sha": kernel_sources_sha(),
        "calibration": {
            "copy_kernel_bytes_read": copy_bytes,
            "copy_kernel_bytes_written": copy_bytes,
            "fetch_scale": fscale,
            "write_scale": wscale,
            "commands": commands,
        },
    }
    for key in sorted(set(fetch) & set(write), key=lambda k: (str(k[1]), k[0])):
        name, prec = key
        if prec is None:
            continue
        rd, wr = mean(fetch[key]) * fscale, mean(write[key]) * wscale
        result.setdefault(f"qft{qubits}_{prec}", {})[name] = {
            "dispatches": len(fetch[key]),
            "read_bytes_per_launch": rd,
            "write_bytes_per_launch": wr,
            "bytes_per_launch": rd + wr,
        }
    return result


def markdown(result: dict, qubits: int) -> str:
    cal = result["calibration"]
    lines = [f"# HBM traffic per launch, {qubits}-qubit bench legs (rocprofv3 PMC)", "",
             "Commands (repository root, one MI355X, one counter per run):", ""]
    lines += [f"    {c}" for c in cal["commands"]]
    lines += ["", f"Calibration: k_copy<double> moves {cal['copy_kernel_bytes_read']:.0f} B each way; FETCH_SIZE reads "
              f"{cal['fetch_scale']:.4f} x and WRITE_SIZE {cal['write_scale']:.4f} x below that, so every kernel's counters "
              f"are scaled by those factors.  Kernel sources: {result['kernel_sources_sha']}.", ""]
    for wl in sorted(k for k in result if k.startswith("qft")):
        state = (1 << qubits) * (16 if wl.endswith("c128") else 8)
        lines += [f"## {wl}", "", "| Kernel | Dispatches | Read B / launch | Written B / launch | Read / state | Written / state |",
                  "|---|---|---|---|---|---|"]
        for name, v in sorted(result[wl].items()):
            lines.append(f"| {name} | {v['dispatches']} | {v['read_bytes_per_launch']:.4e} | {v['write_bytes_per_launch']:.4e} | "
                         f"{v['read_bytes_per_launch'] / state:.3f} | {v['write_bytes_per_launch'] / state:.3f} |")
        lines.append("")
    return "\n".join(lines) + "\n"


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--fetch", required=True, help="rocprofv3 -d directory of the FETCH_SIZE run")
    p.add_argument("--write", required=True, help="rocprofv3 -d directory of the WRITE_SIZE run")
    p.add_argument("--qubits", type=int, default=30, help="state size of the bench legs (the copy probe's state)")
    p.add_argument("--command", action="append", default=[], help="the profiled commands, recorded in the output")
    p.add_argument("--out", required=True, help="JSON output (profiles/traffic.json)")
    p.add_argument("--markdown", help="also write a Markdown summary here")
    a = p.parse_args(argv)
    result = summarize(a.fetch, a.write, a.qubits, a.command)
    with open(a.out, "w") as fh:
        json.dump(result, fh, indent=1)
        fh.write("\n")
    if a.markdown:
        with open(a.markdown, "w") as fh:
            fh.write(markdown(result, a.qubits))
    print(json.dumps({k: v for k, v in result.items() if k != "calibration"}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
