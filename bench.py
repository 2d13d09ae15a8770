#!/usr/bin/env python3
"""Benchmark of the dense state-vector hot path on one MI355X.

Workload (the headline step): QFT on 30 qubits in complex128 -- makeQft(30),
480 circuit operations -- applied through sim::Simulator's timed region
(beginOperationBatch .. endOperationBatch, which waits for the device) to a
register already resident in HBM that holds QuantumState::randomPhase(30, 7).
One step = one QFT-30 circuit.  value = QFT-30 circuits per second.

Also measured in the same run and reported beside the headline:
  * Grover-30, complex128, materialized diffusion, marked {2^29 - 1}: the full
    25735-iteration circuit through sim::Simulator::run, its marked-state
    probability and the Mt19937(123456) measurement;
  * complex64: QFT-30 time and fidelity against the complex128 result
    (device comparison);
  * dense five-qubit blocks (8 random 32 x 32 unitaries) on the 30-qubit
    state, complex64 on FP32 MFMA against complex128;
  * batched circuits: the committed seed-42 workload (256 circuits of 10-22
    qubits, 50-200 gates) as one autotuned batch graph launch, against the same
    circuits run one at a time;
  * adjoint gradients: <Z_0> + |0><0| after 8 ring layers on 24 qubits (192
    controlled-phase angles from mt19937(3)), one forward + one backward sweep,
    against 384 forward circuits of plain central differences on the same GPU
    (time) and the parameter-shift rule (accuracy), complex128 and complex64;
  * the streaming-copy ceiling of this GPU;
  * roofline of the dominant kernel from live HIP-event timing on the engine's
    stream;
  * cpu_baseline: the reference's own dense backend (compiled from its
    sources into oracle/_ref) timed on this host on a bounded sample of
    QFT-30 operations on a 2^30 state, extrapolated by operation type.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30]
"""

import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QFT-30/Grover-30 wall s + HBM GB/s; adjoint grads/s vs plain FD at ≤1e-9 err"
HBM_PEAK_GBPS = 8000.0
FP64_VECTOR_PEAK_ADDS = 39.3e12  # MI355X FP64 vector 78.6 TFLOP/s counts an FMA as two flops


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def read_traffic(kernel, workload):
    """HBM bytes per launch from the committed PMC summary (two rocprofv3
    --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel's
    readings), or None when it is absent or was measured on other kernel
    sources than these."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        d = json.load(open(path))
        from tmfqsfullstate_amd.pmc_traffic import kernel_sources_sha

        if d.get("kernel_sources_sha") != kernel_sources_sha():
            log("profiles/traffic.json was measured on other kernel sources: traffic reported as null")
            return None
        return d[workload][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def time_qft(S, n, precision, steps, warmup, seed=7):
    t0 = time.perf_counter()
    reg = S.Register.random_phase(n, seed, precision=precision)
    log(f"randomPhase({n}, {seed}) generated and loaded in {time.perf_counter() - t0:.1f} s")
    ops, cnt = S.qft_ops(n)
    hs = reg.hip_state()
    for _ in range(warmup):
        reg.run(ops, cnt)
    hs.synchronize()
    hs.profile(True)
    hs.profile_reset()
    exec_s = []
    t0 = time.perf_counter()
    for _ in range(steps):
        exec_s.append(reg.run(ops, cnt))
    hs.synchronize()
    wall = time.perf_counter() - t0
    prof = hs.profile_read()
    hs.profile(False)
    return reg, wall, exec_s, prof, cnt


def time_adjoint(T, S, precision, n=24, layers=8, seed=3, steps=3):
    """Adjoint gradient of <Z_0> + |0><0| for makeRingLayers(n, layers, seed)
    against plain central differences (h = 1e-6) and the parameter-shift rule,
    both from forward runs of the same engine on the same GPU."""
    ops, cnt = S.ring_ops(n, layers, seed)
    terms = [(T.DIAG_Z, 1.0, 1 << 0), (T.DIAG_PROJECTOR, 1.0, 0)]
    params = [i for i in range(cnt) if ops[i].kind == T.OP_CPHASE]
    out = {"qubits": n, "layers": layers, "angles": len(params), "gates": cnt,
           "precision": "complex128" if precision == T.C128 else "complex64"}
    with T.HipState(precision=precision) as s:
        s.init_basis(n, 0)
        s.adjoint(ops, cnt, terms)  # warmup: allocates the checkpoint buffers
        times = []
        for k in range(steps):
            s.init_basis(n, 0)
            s.synchronize()
            if k == steps - 1:
                s.profile(True)
                s.profile_reset()
            t0 = time.perf_counter()
            e, g = s.adjoint(ops, cnt, terms)
            times.append(time.perf_counter() - t0)
        prof = s.profile_read()
        s.profile(False)
        adj_s = sorted(times)[len(times) // 2]

        def energy():
            s.init_basis(n, 0)
            s.run(ops, cnt)
            return s.expectation(terms)

        def two_point(delta):
            vals = []
            for i in params:
                th = ops[i].theta
                ops[i].theta = th + delta
                ep = energy()
                ops[i].theta = th - delta
                em = energy()
                ops[i].theta = th
                vals.append((ep, em))
            return vals

        t0 = time.perf_counter()
        fd = np.array([(a - b) / 2e-6 for a, b in two_point(1e-6)])
        fd_s = time.perf_counter() - t0
        ps = np.array([0.5 * (a - b) for a, b in two_point(np.pi / 2)])
    out.update({
        "expectation": e,
        "adjoint_s": adj_s,
        "adjoint_runs_s": times,
        "grads_per_s": len(params) / adj_s,
        "fd_s": fd_s,
        "fd_forward_circuits": 2 * len(params),
        "fd_grads_per_s": len(params) / fd_s,
        "speedup_vs_fd": fd_s / adj_s,
        "max_abs_diff_vs_param_shift": float(np.max(np.abs(g - ps))),
        "max_abs_diff_vs_central_fd": float(np.max(np.abs(g - fd))),
        "kernels": prof,
    })
    if precision != T.C128:
        # central differences at h = 1e-6 on FP32 amplitudes measure the FP32
        # rounding of two energies (~1e-7 / 2e-6), not the gradient: the
        # complex64 gradients are compared with the complex128 adjoint instead,
        # and the 1e-9 agreement target applies to complex128
        out["fd_on_fp32_state_not_an_accuracy_figure"] = out.pop("max_abs_diff_vs_central_fd")
        out["param_shift_on_fp32_state"] = out.pop("max_abs_diff_vs_param_shift")
    return out, g


def time_grover(T, S, n, marked, precision, materialized=True):
    """makeGrover(n, marked, materialized) through Simulator::run's timed
    region from QuantumState::basis(n); returns (register, report)."""
    ops, cnt = S.grover_ops(n, marked, materialized)
    reg = S.Register.basis(n, 0, precision=precision)
    hs = reg.hip_state()
    hs.profile(True)
    hs.profile_reset()
    secs = reg.run(ops, cnt)
    prof = hs.profile_read()
    hs.profile(False)
    p = sum(abs(reg.amplitude(m)) ** 2 for m in marked)
    m = len(set(marked))
    iters = (cnt - n) // (m + 1)
    gbytes = sum(v["bytes"] for v in prof.values())
    gms = sum(v["ms"] for v in prof.values())
    k = math.floor(math.pi / 4 * math.sqrt(2.0 ** n / m))
    closed = math.sin((2 * k + 1) * math.asin(math.sqrt(m / 2.0 ** n))) ** 2
    return reg, {
        "qubits": n,
        "marked": list(marked),
        "precision": "complex128" if precision == T.C128 else "complex64",
        "diffusion": "materialized" if materialized else "lazy",
        "iterations": iters,
        "wall_s": secs,
        "iterations_per_s": iters / secs,
        "p_marked": p,
        "p_marked_closed_form": closed,
        "p_marked_error": abs(p - closed),
        "measured_seed_123456": reg.measure_seed(123456),
        "hbm_bytes": gbytes,
        "hbm_GBps_over_kernel_time": gbytes / (gms * 1e-3) / 1e9 if gms > 0 else None,
        "kernels": prof,
    }


def time_qft32(T, S, n=32, seed=11):
    """QFT-32 in complex64 on QuantumState::randomPhase(32, 11) against the
    complex128 result of the same engine (the FP64 oracle at this size)."""
    for prec in (T.C64, T.C128):  # first launches (kernel attributes, code objects) on a small state
        w = S.Register.random_phase(20, seed, precision=prec)
        w.run(*S.qft_ops(20))
        del w
    ops, cnt = S.qft_ops(n)
    t0 = time.perf_counter()
    r64 = S.Register.random_phase(n, seed, precision=T.C64)
    gen_s = time.perf_counter() - t0
    secs64 = r64.run(ops, cnt)
    log(f"QFT-{n} complex64: {secs64 * 1e3:.1f} ms (input generation + load {gen_s:.1f} s)")
    r128 = S.Register.random_phase(n, seed, precision=T.C128)
    secs128 = r128.run(ops, cnt)
    log(f"QFT-{n} complex128: {secs128 * 1e3:.1f} ms")
    cmp = r128.hip_state().compare(r64.hip_state())
    return {
        "qubits": n,
        "seed": seed,
        "complex64_s": secs64,
        "complex128_s": secs128,
        "speedup_vs_complex128": secs128 / secs64,
        "fidelity": cmp["fidelity"],
        "fidelity_error": 1.0 - cmp["fidelity"],
        "max_abs_error": cmp["max_abs_amplitude_error"],
        "state_GiB_complex64": (8 << n) / 2 ** 30,
    }


def time_blocks(T, n=30, gates=8, seed=5):
    """Dense five-qubit blocks (random 32 x 32 unitaries on random qubits) on a
    2^n state: complex64 runs them on FP32 MFMA inside the tiled pass,
    complex128 on FP64 VALU with the dense backend's accumulation order."""
    rng = np.random.default_rng(seed)
    mats, qubits = [], []
    for _ in range(gates):
        z = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
        q, r = np.linalg.qr(z)
        mats.append(q * (np.diag(r) / np.abs(np.diag(r))))
        qubits.append([int(x) for x in rng.choice(n, 5, replace=False)])
    out, states = {}, []
    for prec in (T.C64, T.C128):
        s = T.HipState(precision=prec)
        s.init_basis(n, 0)
        for q in range(n):
            s.h(q)
        s.synchronize()
        s.profile(True)
        s.profile_reset()
        t0 = time.perf_counter()
        for u, qs in zip(mats, qubits):
            s.gate(u, qs)
        s.synchronize()
        secs = time.perf_counter() - t0
        prof = s.profile_read()
        s.profile(False)
        flops = gates * 8.0 * 32 * (1 << n)  # 8 real flops per complex MAC, 32 MACs per amplitude
        out["complex64" if prec == T.C64 else "complex128"] = {
            "seconds": secs,
            "gate_passes": sum(v["launches"] for v in prof.values()),
            "TFLOPs_dense_blocks": flops / secs / 1e12,
            "kernels": prof,
        }
        states.append(s)
    cmp = states[1].compare(states[0])
    out.update({"qubits": n, "gates": gates, "fidelity_complex64_vs_complex128": cmp["fidelity"],
                "max_abs_error": cmp["max_abs_amplitude_error"],
                "speedup_complex64": out["complex128"]["seconds"] / out["complex64"]["seconds"]})
    for st in states:
        st.close()
    return out


def time_chunked(T, S, n=30, chunk_qubits=24, window_bits=4):
    """QFT-n complex128 on a register held in pinned host memory as 2^(n-c)
    chunks, streamed through two device windows of 2^(c+g) amplitudes.  Input
    |0..0>: the transfers and kernels do not depend on the data, and the
    output is checked against the closed form (every amplitude 2^(-n/2))."""
    from tmfqsfullstate_amd.chunked import ChunkedState

    ops, cnt = S.qft_ops(n)
    t0 = time.perf_counter()
    s = ChunkedState(n, chunk_qubits, window_bits)
    alloc_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.run(ops, cnt)
    secs = time.perf_counter() - t0
    info = s.info()
    head = s.export(0, 1 << 20)
    tail = s.export((1 << n) - (1 << 20), 1 << 20)
    err = max(float(np.max(np.abs(head - 2.0 ** (-n / 2)))), float(np.max(np.abs(tail - 2.0 ** (-n / 2)))))
    s.close()
    moved = info["bytes_h2d"] + info["bytes_d2h"]
    return {"qubits": n, "chunk_qubits": chunk_qubits, "window_bits": window_bits, "input": "basis state 0",
            "host_bytes": 16.0 * (1 << n), "window_bytes": 16.0 * (1 << (chunk_qubits + window_bits)),
            "seconds": secs, "alloc_s": alloc_s, "segments": info["segments"], "windows": info["windows"],
            "bytes_h2d": info["bytes_h2d"], "bytes_d2h": info["bytes_d2h"],
            "host_link_GBps": moved / secs / 1e9, "max_abs_error_vs_closed_form_first_last_2^20": err}


def composed_circuit(T, n, seed=9, depth=400, width=5):
    """A deep run of gates on `width` random qubits (random 1-qubit unitaries,
    CX, H, controlled phases, phases) after a Hadamard layer on all n qubits:
    the shape gate-run composition is for."""
    rng = np.random.default_rng(seed)
    sub = [int(q) for q in rng.choice(n, width, replace=False)]
    seq, mats, quals = [], [], []
    for q in range(n):
        seq.append((T.OP_H, q, 0, 0.0, 0, 0))
    for _ in range(depth):
        k = int(rng.integers(0, 5))
        a, b = (int(x) for x in rng.choice(sub, 2, replace=False))
        if k == 0:
            z = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
            u, _ = np.linalg.qr(z)
            seq.append((T.OP_MATRIX, len(quals), 0, 0.0, 2 * 4 * len(mats), 1))
            quals.append(a)
            mats.append(u)
        elif k == 1:
            seq.append((T.OP_CX, a, b, 0.0, 0, 0))
        elif k == 2:
            seq.append((T.OP_H, a, 0, 0.0, 0, 0))
        elif k == 3:
            seq.append((T.OP_CPHASE, a, b, float(rng.uniform(0, 2 * np.pi)), 0, 0))
        else:
            seq.append((T.OP_PHASE, a, 0, float(rng.uniform(0, 2 * np.pi)), 0, 0))
    arr = (T.Op * len(seq))()
    for i, (kind, q0, q1, th, st, k) in enumerate(seq):
        arr[i].kind, arr[i].q0, arr[i].q1, arr[i].theta, arr[i].state, arr[i].k = kind, q0, q1, th, st, k
    pool = np.zeros(8 * len(mats))
    for i, u in enumerate(mats):
        pool[8 * i:8 * i + 8:2] = u.real.ravel()
        pool[8 * i + 1:8 * i + 8:2] = u.imag.ravel()
    return arr, len(seq), pool, np.array(quals, dtype=np.uint32)


def time_composition(T, n=30, reps=2):
    """Gate-run composition on the deep five-qubit run: device time for each
    composition width (0 = every gate applied in register stages), the
    autotuner's choice, and the error of the composed result against the
    uncomposed complex128 one."""
    arr, cnt, pool, quals = composed_circuit(T, n)
    out = {"qubits": n, "gates": cnt}
    L = T.lib()
    for prec in (T.C128, T.C64):
        name = "complex128" if prec == T.C128 else "complex64"
        sweep = {}
        states = {}
        with T.HipState(precision=prec) as s:
            for fq in (0, 3, 4, 5):
                T.check(L.tmfqs_hip_set_tuning(s._h, 0, fq))
                best = None
                for _ in range(reps):
                    s.init_basis(n, 0)
                    s.synchronize()
                    s.profile(True)
                    s.profile_reset()
                    s.run(arr, cnt, pool, quals)
                    s.synchronize()
                    p = s.profile_read()
                    s.profile(False)
                    ms = sum(v["ms"] for v in p.values())
                    best = ms if best is None else min(best, ms)
                sweep[fq] = {"device_ms": best, "passes": sum(v["launches"] for k, v in p.items() if k.startswith("k_")),
                             "composed_runs": p.get("gate_runs_composed", {}).get("launches", 0)}
                if fq in (0, 5):
                    states[fq] = s.clone()
            s.init_basis(n, 0)
            t, f, ms = s.tune(arr, cnt, pool, quals)
        cmp = states[0].compare(states[5])
        for st in states.values():
            st.close()
        out[name] = {"device_ms_by_fuse_qubits": sweep, "tuned": {"tile_qubits": t, "fuse_qubits": f, "device_ms": ms},
                     "speedup_fuse5_vs_off": sweep[0]["device_ms"] / sweep[5]["device_ms"],
                     "max_abs_diff_fuse5_vs_off": cmp["max_abs_amplitude_error"]}
    return out


def time_batch(T, precision, steps=3):
    """The committed seed-42 workload (256 circuits, 10-22 qubits) as one
    batch (autotuned tile width, one graph launch per run) against the same
    circuits run one at a time through the engine."""
    from tmfqsfullstate_amd.batch import Batch, load_circuit_list

    circuits = load_circuit_list(os.path.join(ROOT, "tests", "golden", "batch_circuits_seed42.txt"))
    gates = sum(c[2] for c in circuits)
    terms = [(T.DIAG_Z, 1.0, 1 << 0), (T.DIAG_PROJECTOR, 1.0, 0)]  # Z on qubit 0 plus |0><0|
    with Batch(circuits, precision=precision) as b:
        tuned = b.tune()
        b.run()
        runs = [b.run() for _ in range(steps)]
        info = b.info()
        b.adjoint(terms)
        t0 = time.perf_counter()
        _, grads = b.adjoint(terms)
        adjoint_s = time.perf_counter() - t0
    batch_s = sorted(runs)[len(runs) // 2]
    with T.HipState(precision=precision) as s:
        for n, ops, cnt in circuits[:8]:  # warm up the single-circuit path
            s.init_basis(n, 0)
            s.run(ops, cnt)
        s.synchronize()
        t0 = time.perf_counter()
        for n, ops, cnt in circuits:
            s.init_basis(n, 0)
            s.run(ops, cnt)
            s.synchronize()
        seq_s = time.perf_counter() - t0
    return {
        "circuits": len(circuits),
        "gates": gates,
        "precision": "complex128" if precision == T.C128 else "complex64",
        "batch_s": batch_s,
        "batch_runs_s": runs,
        "circuits_per_s": len(circuits) / batch_s,
        "gates_per_s": gates / batch_s,
        "sequential_s": seq_s,
        "speedup_vs_sequential": seq_s / batch_s,
        "tuned_tile_qubits": tuned,
        "adjoint_s": adjoint_s,
        "adjoint_gradients": int(sum(len(g) for g in grads)),
        "rounds": info["rounds"],
        "work_items": info["items"],
        "graph": bool(info["graph"]),
        "state_GiB": info["state_bytes"] / 2 ** 30,
    }


def cpu_baseline(n, gpu_seconds):
    """The reference's dense backend on this host, one thread (it has no
    threads), against the GPU's times of the same workloads:

    * QFT-n and Grover-n (minutes and hours on a CPU) from a bounded sample:
      per-operation times of makeQft(n)'s operation kinds and of 17 Grover
      iterations on one dense 2^n-state register, extrapolated by operation
      count (every Grover iteration after the first is the same sweep; the
      first also computes the cached amplitude sum);
    * Grover-20 and QFT-26 run in full through sim::Simulator (its timed
      region, as on the GPU).
    """
    import oracle as O

    reference = O.ref_available()
    kind = "reference" if reference else "port"
    marked = (1 << (n - 1)) - 1
    cp_pairs = [(1, 0), (5, 2), (9, 4), (13, 6), (17, 8), (21, 10), (25, 12), (29, 14), (29, 28), (20, 19), (27, 1),
                (15, 3)]
    seq = [(O.OP_H, q % n) for q in (0, 10, 20, 29)]
    seq += [(O.OP_CPHASE, c % n, t % n, math.pi / 2 ** ((c % n) - (t % n))) for c, t in cp_pairs if c % n > t % n]
    seq += [(O.OP_SWAP, a % n, b % n) for a, b in ((0, 29), (5, 24), (14, 15)) if a % n != b % n]
    nq = len(seq)
    grover_iters_sampled = 17
    seq += [(O.OP_PHASE_FLIP, 0, 0, 0.0, marked), (O.OP_DIFFUSE_MAT,)] * grover_iters_sampled
    t0 = time.perf_counter()
    total_s, per = O.time_ops_each(n, seq, reference)
    kinds = [t[0] for t in seq[:nq]]
    per_op = {name: float(np.mean([p for p, k in zip(per[:nq], kinds) if k == code]))
              for name, code in (("h", O.OP_H), ("cphase", O.OP_CPHASE), ("swap", O.OP_SWAP))}
    it = per[nq::2] + per[nq + 1::2]
    first_iter, steady_iter = float(it[0]), float(np.mean(it[1:]))
    nh, ncp, nsw = n, n * (n - 1) // 2, n // 2
    qft_s = nh * per_op["h"] + ncp * per_op["cphase"] + nsw * per_op["swap"]
    iters = O.grover_iterations(n, 1)
    # the H layer of Grover starts from a basis state, where the reference
    # skips all-zero pairs: the dense-state H time bounds it from above
    grover_s = nh * per_op["h"] + first_iter + (iters - 1) * steady_iter
    sample_s = time.perf_counter() - t0
    workloads = {
        f"qft{n}_c128": {"cpu_s": qft_s, "cpu_measured": False, "gpu_s": gpu_seconds.get("qft")},
        f"grover{n}_c128": {"cpu_s": grover_s, "cpu_measured": False, "gpu_s": gpu_seconds.get("grover"),
                            "iterations": iters, "cpu_sampled_iterations": grover_iters_sampled,
                            "cpu_first_iteration_s": first_iter, "cpu_iteration_s": steady_iter},
    }
    full = {}
    if reference:
        from tmfqsfullstate_amd import sim as S
        g20 = O.grover_ops(20, [524287], True)
        full["grover20_c128"] = (O.ref_run(20, g20)[3], gpu_seconds.get("grover20"))
        init26 = S.random_phase(26, 7)
        full["qft26_c128"] = (O.ref_run(26, O.qft_ops(26), init=init26)[3], gpu_seconds.get("qft26"))
        del init26
    for name, (cs, gs) in full.items():
        workloads[name] = {"cpu_s": cs, "cpu_measured": True, "gpu_s": gs}
    for w in workloads.values():
        w["speedup"] = w["cpu_s"] / w["gpu_s"] if w.get("gpu_s") else None
    return {
        "value": 1.0 / qft_s,
        "unit": f"QFT-{n} circuits/s",
        "cores": 1,
        "kind": kind,
        "projected_qft_seconds": qft_s,
        "projected_grover_seconds": grover_s,
        "per_op_seconds": per_op,
        "workloads": workloads,
        "sample": (f"one dense 2^{n}-state complex128 register (distinct real amplitudes, so no zero skips): "
                   f"{len([k for k in kinds if k == O.OP_H])} H, {len([k for k in kinds if k == O.OP_CPHASE])} "
                   f"controlled-phase and {len([k for k in kinds if k == O.OP_SWAP])} swap (3 CX each) operations "
                   f"of makeQft({n}) and {grover_iters_sampled} Grover iterations (phase flip + materialized "
                   f"diffusion) timed per operation ({total_s:.1f} s of CPU work), extrapolated to makeQft({n})'s "
                   f"{nh} H + {ncp} controlled-phase + {nsw} swap operations and makeGrover({n})'s {nh} H + "
                   f"{iters} iterations; Grover-20 and QFT-26 (randomPhase(26, 7)) run in full through "
                   f"sim::Simulator" + ("" if reference else " (skipped: no reference build)") +
                   f"; {sample_s:.0f} s in all; single thread, {os.cpu_count()} host CPUs, {cpu_model()}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-grover", action="store_true")
    ap.add_argument("--skip-c64", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-chunked", action="store_true")
    ap.add_argument("--qft32", action="store_true", help="also run QFT-32 complex64 vs complex128 (96 GiB of HBM)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); the state-vector engine runs on "
              f"a single MI355X in this build.", file=sys.stderr)
        sys.exit(2)

    import tmfqsfullstate_amd as T
    from tmfqsfullstate_amd import sim as S

    n = args.qubits
    N = 1 << n
    result = {}

    # ---- headline: QFT-n complex128
    log(f"QFT-{n} complex128: generating randomPhase({n}, 7) and loading it")
    reg, wall, exec_s, prof, nops = time_qft(S, n, T.C128, args.steps, args.warmup)
    ms_per_step = 1000.0 * wall / args.steps
    hs = reg.hip_state()
    copy_ms = hs.copy_probe(3)
    copy_gbps = 2.0 * 16 * N / (copy_ms * 1e-3) / 1e9
    tp = reg.total_probability()
    # host hand-over of the whole state through the C-ABI (pinned, double-
    # buffered staging over PCIe): reported beside, never part of `value`
    t0 = time.perf_counter()
    host = hs.export()
    export_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    hs.load(n, host)
    hs.synchronize()
    load_s = time.perf_counter() - t0
    del host
    state_io = {"bytes": 16.0 * N, "export_s": export_s, "export_GBps": 16.0 * N / export_s / 1e9,
                "load_s": load_s, "load_GBps": 16.0 * N / load_s / 1e9}
    log(f"state I/O: export {state_io['export_GBps']:.1f} GB/s, load {state_io['load_GBps']:.1f} GB/s (host buffers)")
    step_bytes = sum(v["bytes"] for v in prof.values()) / args.steps
    launches = {k: v["launches"] / args.steps for k, v in prof.items()}
    dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    dname, dstat = dom
    avg_ms = dstat["ms"] / dstat["launches"]
    bytes_per_launch = dstat["bytes"] / dstat["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    log(f"QFT-{n}: {ms_per_step:.2f} ms/step, kernels {launches}, copy ceiling {copy_gbps:.0f} GB/s")
    qft = {
        "wall_s": ms_per_step / 1000.0,
        "execution_seconds": exec_s,
        "ops": nops,
        "kernel_launches_per_circuit": launches,
        "hbm_bytes_per_circuit": step_bytes,
        "hbm_GBps": step_bytes / (ms_per_step * 1e-3) / 1e9,
        "hbm_fraction": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        "gate_equivalent_GBps_not_roofline": (n + n * (n - 1) // 2 + 3 * (n // 2)) * 2.0 * 16 * N
        / (ms_per_step * 1e-3) / 1e9,
        "total_probability_after": tp,
        "state_io_host_buffers": state_io,
        "kernels": prof,
    }
    roofline = {
        "kernel": dname,
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBPS,
        "traffic": read_traffic(dname, f"qft{n}_c128"),
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": avg_ms,
        "copy_kernel_GBps": copy_gbps,
    }

    # ---- complex64 QFT and fidelity against complex128
    if not args.skip_c64:
        log(f"QFT-{n} complex64")
        reg64, wall64, exec64, prof64, _ = time_qft(S, n, T.C64, args.steps, args.warmup)
        ms64 = 1000.0 * wall64 / args.steps
        # same number of QFTs applied to both (warmup + steps): compare states on the device
        cmp = hs.compare(reg64.hip_state())
        result["complex64"] = {
            "qft_wall_s": ms64 / 1000.0,
            "speedup_vs_complex128": ms_per_step / ms64,
            "fidelity_vs_complex128": cmp["fidelity"],
            "fidelity_error": 1.0 - cmp["fidelity"],
            "max_abs_error_vs_complex128": cmp["max_abs_amplitude_error"],
            "kernels": prof64,
        }
        del reg64
    del reg, hs

    # ---- Grover-n complex128, full circuit
    if not args.skip_grover:
        marked = (1 << (n - 1)) - 1
        ops, cnt = S.grover_ops(n, [marked], True)
        log(f"Grover-{n} complex128: {cnt} ops")
        greg = S.Register.basis(n, 0, precision=T.C128)  # QuantumState::basis(n) materialised
        gh = greg.hip_state()
        gh.profile(True)
        gh.profile_reset()
        gsecs = greg.run(ops, cnt)  # Simulator::run's timed region
        gprof = gh.profile_read()
        gh.profile(False)
        a = greg.amplitude(marked)
        meas = greg.measure_seed(123456)
        iters = (cnt - n) // 2
        sweeps = sum(v["launches"] for k, v in gprof.items()
                     if k in ("k_diffuse_chain", "k_ew_chain", "k_tile", "k_hadamard"))
        gbytes = sum(v["bytes"] for v in gprof.values())
        gms = sum(v["ms"] for v in gprof.values())
        result["grover"] = {
            "qubits": n,
            "marked": marked,
            "iterations": iters,
            "wall_s": gsecs,
            "iterations_per_s": iters / gsecs,
            "p_marked": abs(a) ** 2,
            "measured_seed_123456": meas,
            "full_state_sweeps": sweeps,
            "hbm_bytes": gbytes,
            "hbm_GBps_over_kernel_time": gbytes / (gms * 1e-3) / 1e9 if gms > 0 else None,
            "hbm_GBps_over_wall": gbytes / gsecs / 1e9,
            "one_sweep_per_iteration_equivalent_GBps_not_hbm": iters * 2.0 * 16 * N / gsecs / 1e9,
            # Grover's state stays real (every imaginary part +-0), so each
            # diffusion step is one FP64 add per amplitude (the real part; the
            # imaginary part is the chain applied to +-0, computed once on the
            # host): with 1024 steps per sweep the sweeps are bound by this
            # rate, not by HBM bandwidth
            "diffusion_fp64_adds_per_s": (iters * 1.0 * N / (gprof["k_diffuse_chain"]["ms"] * 1e-3)
                                          if "k_diffuse_chain" in gprof else None),
            "diffusion_hbm_GBps": (gprof["k_diffuse_chain"]["bytes"] / (gprof["k_diffuse_chain"]["ms"] * 1e-3) / 1e9
                                   if "k_diffuse_chain" in gprof else None),
            "fp64_vector_peak_adds_per_s": FP64_VECTOR_PEAK_ADDS,
            "kernels": gprof,
        }
        if result["grover"]["diffusion_fp64_adds_per_s"]:
            result["grover"]["fraction_of_fp64_vector_peak"] = result["grover"]["diffusion_fp64_adds_per_s"] / FP64_VECTOR_PEAK_ADDS
        log(f"Grover-{n}: {gsecs:.3f} s, P={abs(a) ** 2:.15f}, measured {meas}, diffusion sweeps at "
            f"{result['grover'].get('fraction_of_fp64_vector_peak', 0):.0%} of the FP64 add rate and "
            f"{(result['grover']['diffusion_hbm_GBps'] or 0) / 8000:.0%} of HBM peak")
        del greg, gh

    # ---- Grover-n with three marked states, complex64 against complex128
    if not args.skip_grover and not args.skip_c64:
        marked3 = [1, (1 << (n - 1)) + 5, (1 << n) - 1]
        log(f"Grover-{n} with 3 marked states, complex64 and complex128")
        g64, rep64 = time_grover(T, S, n, marked3, T.C64)
        g128, rep128 = time_grover(T, S, n, marked3, T.C128)
        cmp = g128.hip_state().compare(g64.hip_state())
        rep64["speedup_vs_complex128"] = rep128["wall_s"] / rep64["wall_s"]
        rep64["fidelity_vs_complex128"] = cmp["fidelity"]
        rep64["fidelity_error"] = 1.0 - cmp["fidelity"]
        rep64["max_abs_error_vs_complex128"] = cmp["max_abs_amplitude_error"]
        result["grover3_complex64"] = rep64
        result["grover3_complex128"] = rep128
        log(f"Grover-{n} x3: complex64 {rep64['wall_s']:.3f} s, complex128 {rep128['wall_s']:.3f} s, "
            f"P {rep64['p_marked']:.12f} / {rep128['p_marked']:.12f}")
        del g64, g128
    if not args.skip_grover:
        g20, rep20 = time_grover(T, S, 20, [524287], T.C128)
        result["grover20"] = rep20
        del g20
        # lazy diffusion (the affine overlay): no full-state sweep per iteration
        glz, replz = time_grover(T, S, n, [(1 << (n - 1)) - 1], T.C128, materialized=False)
        result["grover_lazy"] = replz
        log(f"Grover-{n} lazy: {replz['wall_s']:.3f} s, P={replz['p_marked']:.15f}, "
            f"measured {replz['measured_seed_123456']}")
        del glz

    if args.qft32:
        result["qft32_complex64"] = time_qft32(T, S)

    # ---- dense five-qubit blocks (FP32 MFMA in complex64)
    if not args.skip_c64:
        log(f"five-qubit blocks on {n} qubits: complex64 (MFMA) and complex128")
        blk = time_blocks(T, n)
        log(f"blocks: complex64 {blk['complex64']['seconds'] * 1e3:.1f} ms, complex128 "
            f"{blk['complex128']['seconds'] * 1e3:.1f} ms, fidelity {blk['fidelity_complex64_vs_complex128']:.12f}")
        result["five_qubit_blocks"] = blk

    # ---- gate-run composition and the circuit autotuner
    if not args.skip_c64:
        log(f"gate-run composition: a deep five-qubit run on {n} qubits")
        comp = time_composition(T, n)
        log("composition: " + "; ".join(
            f"{p}: " + ", ".join(f"k={k} {v['device_ms']:.1f} ms" for k, v in comp[p]["device_ms_by_fuse_qubits"].items()) +
            f" (tuned k={comp[p]['tuned']['fuse_qubits']}, tile {comp[p]['tuned']['tile_qubits']})"
            for p in ("complex128", "complex64")))
        result["composition"] = comp

    # ---- adjoint gradients: n = 24, 8 ring layers, 192 angles
    if not args.skip_adjoint:
        log("adjoint n=24 complex128 (and central differences / parameter shift on the same GPU)")
        adj, g128 = time_adjoint(T, S, T.C128)
        log(f"adjoint: {adj['adjoint_s'] * 1e3:.2f} ms, FD {adj['fd_s']:.2f} s, speedup {adj['speedup_vs_fd']:.1f}x, "
            f"|g - shift| {adj['max_abs_diff_vs_param_shift']:.2e}")
        if not args.skip_c64:
            adj64, g64 = time_adjoint(T, S, T.C64, steps=3)
            adj64["max_abs_diff_vs_complex128_adjoint"] = float(np.max(np.abs(g64 - g128)))
            adj64["speedup_vs_complex128_adjoint"] = adj["adjoint_s"] / adj64["adjoint_s"]
            adj["complex64"] = adj64
        result["adjoint"] = adj

    # ---- batched uneven circuits: the seed-42 workload
    if not args.skip_batch:
        log("batched circuits: 256 circuits of 10-22 qubits (seed 42), complex128")
        bt = time_batch(T, T.C128)
        log(f"batch: {bt['batch_s'] * 1e3:.1f} ms per batch ({bt['circuits_per_s']:.0f} circuits/s), "
            f"sequential {bt['sequential_s'] * 1e3:.1f} ms, tile {bt['tuned_tile_qubits']}; "
            f"batched adjoint {bt['adjoint_s'] * 1e3:.1f} ms for {bt['adjoint_gradients']} gradients")
        if not args.skip_c64:
            bt["complex64"] = time_batch(T, T.C64)
        result["batch"] = bt

    # ---- chunked register in host memory
    if not args.skip_chunked:
        log(f"QFT-{n} complex128 on a chunked register in pinned host memory")
        ch = time_chunked(T, S, n)
        log(f"chunked: {ch['seconds']:.2f} s, {ch['segments']} segments, {ch['windows']} windows, "
            f"{ch['host_link_GBps']:.1f} GB/s over the host link, max abs error {ch['max_abs_error_vs_closed_form_first_last_2^20']:.1e}")
        result["chunked"] = ch

    # ---- QFT-26 complex128 (the reference's largest QFT campaign size)
    if not args.skip_cpu:
        log("QFT-26 complex128")
        r26, wall26, _, _, _ = time_qft(S, 26, T.C128, args.steps, args.warmup)
        result["qft26"] = {"qubits": 26, "wall_s": wall26 / args.steps}
        del r26

    cpu = None
    if not args.skip_cpu:
        log("CPU baseline (reference dense backend, one thread)")
        gpu_s = {"qft": ms_per_step / 1000.0,
                 "grover": result.get("grover", {}).get("wall_s"),
                 "grover20": result.get("grover20", {}).get("wall_s"),
                 "qft26": result.get("qft26", {}).get("wall_s")}
        cpu = cpu_baseline(n, gpu_s)
        log(f"CPU: projected QFT-{n} {cpu['projected_qft_seconds']:.1f} s, Grover-{n} "
            f"{cpu['projected_grover_seconds']:.0f} s; " +
            ", ".join(f"{k} {v['cpu_s']:.2f} s ({(v['speedup'] or 0):.0f}x)" for k, v in cpu["workloads"].items()))

    value = 1000.0 / ms_per_step
    line = {
        "metric": METRIC,
        "value": value,
        "unit": f"QFT-{n} circuits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: QuantumState::randomPhase({n}, 7) (mt19937 seed 7)",
        "config": {"workload": f"qft{n}_c128", "qubits": n, "circuit": f"makeQft({n})", "ops": nops,
                   "precision": "complex128", "strategy": "hip"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "qft": qft,
    }
    line.update(result)
    if cpu is not None:
        line["speedup_vs_cpu_reference"] = value / cpu["value"]
        g = cpu["workloads"].get(f"grover{n}_c128", {})
        if g.get("speedup"):
            line["grover_speedup_vs_cpu_reference"] = g["speedup"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
