"""Writes the golden fixtures in tests/golden/ from the reference build.

Runs the reference's own dense path (compiled from its sources into
oracle/_ref/libtmfqs_ref.so by oracle/Makefile) at small sizes and stores its
inputs and outputs as data.  Run from the repository root:

    make -C oracle && python tests/golden/make_golden.py

Fixtures (all in golden.npz):
  draws_<seed>            first 16 Mt19937RandomSource(seed).nextUnitDouble()
  random_phase_12_7       QuantumState::randomPhase(12, 7) amplitudes
  qft_12_7                makeQft(12) applied to random_phase_12_7
  grover_<n>_amps         final amplitudes of makeGrover(n, {marked}), materialized
  grover_<n>_meta         [marked, measured(seed 123456), total probability]
  gate_k<k>_in / _mat / _qubits / _out   applyGate of a seeded random unitary
                          on a seeded random 8-qubit state (k = 1..5)
  cx_cp_swap_amps / ops   a fixed mixed circuit (H, X, CX, CP, SWAP) at n = 6
"""

import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402


def random_unitary(rng, dim):
    z = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def interleave_matrix(u):
    out = np.empty(2 * u.size)
    out[0::2] = u.real.ravel()
    out[1::2] = u.imag.ravel()
    return out


def main():
    if not O.ref_available():
        raise SystemExit("reference build missing: make -C oracle")
    fx = {}
    for seed in (123456, 100, 7, 3, 42):
        fx[f"draws_{seed}"] = O.ref_unit_draws(seed, 16)
    init = O.ref_random_phase(12, 7)
    fx["random_phase_12_7"] = init
    fx["qft_12_7"] = O.ref_run(12, O.qft_ops(12), init=init)[0]
    for n, marked in ((10, 511), (12, 2047), (14, 8191)):
        amps, tp, meas, _ = O.ref_run(n, O.grover_ops(n, [marked], True), measure_seed=123456)
        fx[f"grover_{n}_amps"] = amps
        fx[f"grover_{n}_meta"] = np.array([marked, meas, tp], dtype=np.float64)
    rng = np.random.default_rng(2024)
    n = 8
    for k in range(1, 6):
        state = rng.standard_normal(2 << n)
        state /= np.linalg.norm(state)
        qubits = rng.choice(n, size=k, replace=False).astype(np.uint32)
        mat = interleave_matrix(random_unitary(rng, 1 << k))
        ops = O.make_ops([(O.OP_MATRIX, 0, 0, 0.0, 0, k)])
        out = O.ref_run(n, ops, init=state, mat_pool=mat, mat_qubits=qubits)[0]
        fx[f"gate_k{k}_in"] = state
        fx[f"gate_k{k}_mat"] = mat
        fx[f"gate_k{k}_qubits"] = qubits
        fx[f"gate_k{k}_out"] = out
    mixed = [
        (O.OP_H, 0), (O.OP_H, 3), (O.OP_X, 5), (O.OP_CX, 0, 4), (O.OP_CPHASE, 3, 1, 0.7),
        (O.OP_SWAP, 1, 5), (O.OP_H, 2), (O.OP_CPHASE, 2, 0, -1.3), (O.OP_CX, 5, 2), (O.OP_SWAP, 0, 3),
    ]
    fx["cx_cp_swap_ops"] = np.array([[t[0], t[1], t[2] if len(t) > 2 else 0, t[3] if len(t) > 3 else 0.0]
                                     for t in mixed], dtype=np.float64)
    fx["cx_cp_swap_amps"] = O.ref_run(6, mixed, init=O.ref_random_phase(6, 11))[0]
    out = os.path.join(os.path.dirname(__file__), "golden.npz")
    np.savez_compressed(out, **fx)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
