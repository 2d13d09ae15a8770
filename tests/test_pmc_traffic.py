"""The PMC traffic summarizer (tmfqsfullstate_amd/pmc_traffic.py) on
synthetic rocprofv3 counter files: kernel names, precision split, copy
calibration and the recorded kernel-source hash."""

import csv
import json

import pytest

from tmfqsfullstate_amd import pmc_traffic as P

FIELDS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_run(path, counter, rows):
    path.mkdir(parents=True)
    with open(path / "x_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, kib) in enumerate(rows):
            w.writerow({"Correlation_Id": i, "Dispatch_Id": i, "Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": kib})


def test_short_names():
    assert P.short_name("void tmfqs_hip::k_qft_block<double, 5, false>(tmfqs_hip::V2T<double>::t*, int)") == (
        "k_qft_block", "c128")
    assert P.short_name("void tmfqs_hip::(anonymous namespace)::k_random_phase<HIP_vector_type<float, 2u> >(unsigned int const*)") == (
        "k_random_phase", "c64")
    assert P.short_name("__amd_rocclr_copyBuffer") == ("__amd_rocclr_copyBuffer", None)


def test_summarize_scales_by_the_copy_kernel(tmp_path):
    n = 10
    state = (1 << n) * 16 / 1024.0  # KiB
    copy = "void tmfqs_hip::k_copy<double>(tmfqs_hip::V2T<double>::t const*, tmfqs_hip::V2T<double>::t*, unsigned long)"
    blk = "void tmfqs_hip::k_qft_block<double, 5, false>(double2*)"
    blk64 = "void tmfqs_hip::k_qft_block<float, 5, false>(float2*)"
    # FETCH reads half the true bytes, WRITE the true bytes (as on gfx950).
    write_run(tmp_path / "f", "FETCH_SIZE", [(copy, state / 2), (blk, state / 2), (blk, state / 2 * 1.5), (blk64, state / 4),
                                            ("__amd_rocclr_copyBuffer", 3.0)])
    write_run(tmp_path / "w", "WRITE_SIZE", [(copy, state), (blk, state), (blk, state), (blk64, state / 2)])
    out = tmp_path / "traffic.json"
    assert P.main(["--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"), "--qubits", str(n), "--out", str(out),
                   "--command", "cmd a", "--markdown", str(tmp_path / "t.md")]) == 0
    d = json.loads(out.read_text())
    assert d["kernel_sources_sha"] == P.kernel_sources_sha()
    assert d["calibration"]["fetch_scale"] == pytest.approx(2.0) and d["calibration"]["write_scale"] == pytest.approx(1.0)
    b = d[f"qft{n}_c128"]["k_qft_block"]
    assert b["dispatches"] == 2
    assert b["read_bytes_per_launch"] == pytest.approx((1 << n) * 16 * 1.25)
    assert b["bytes_per_launch"] == pytest.approx((1 << n) * 16 * 2.25)
    assert d[f"qft{n}_c64"]["k_qft_block"]["bytes_per_launch"] == pytest.approx((1 << n) * 16)
    md = (tmp_path / "t.md").read_text()
    assert "cmd a" in md and "| k_qft_block | 2 |" in md


def test_summarize_needs_the_copy_kernel(tmp_path):
    write_run(tmp_path / "f", "FETCH_SIZE", [("void tmfqs_hip::k_qft_block<double, 5, false>(x)", 1.0)])
    write_run(tmp_path / "w", "WRITE_SIZE", [("void tmfqs_hip::k_qft_block<double, 5, false>(x)", 1.0)])
    with pytest.raises(ValueError):
        P.summarize(str(tmp_path / "f"), str(tmp_path / "w"), 10, [])
