"""Writes batch_circuits_seed42.txt: the batched-circuit workload.

256 independent circuits; qubit counts uniform in [10, 22], depths uniform in
[50, 200]; each gate drawn uniformly from {H, CX, controlledPhase(theta
uniform in [0, 2 pi)), swap} on uniformly drawn distinct qubits; numpy
Generator(PCG64(42)), drawn in circuit order.  The list is committed so that
every run and every machine uses exactly these circuits.

Format: '#' comment lines; 'C <qubits> <gates>' opens a circuit; then one
gate per line: 'H q', 'CX c t', 'CP c t theta' (theta as repr of the float),
'SW a b'.  Run: python tests/golden/make_batch_circuits.py
"""

import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def generate(seed=42, count=256):
    rng = np.random.Generator(np.random.PCG64(seed))
    circuits = []
    for _ in range(count):
        n = int(rng.integers(10, 23))
        depth = int(rng.integers(50, 201))
        gates = []
        for _ in range(depth):
            kind = int(rng.integers(0, 4))
            if kind == 0:
                gates.append(("H", int(rng.integers(n))))
            else:
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                if kind == 1:
                    gates.append(("CX", a, b))
                elif kind == 2:
                    gates.append(("CP", a, b, float(rng.uniform(0.0, 2.0 * math.pi))))
                else:
                    gates.append(("SW", a, b))
        circuits.append((n, gates))
    return circuits


def write(path, circuits):
    with open(path, "w") as f:
        f.write("# batched circuits: numpy PCG64(42), 256 circuits, qubits U[10,22], depth U[50,200],\n")
        f.write("# gates U{H, CX, CP(theta U[0,2pi)), SW} on distinct random qubits (tests/golden/make_batch_circuits.py)\n")
        for n, gates in circuits:
            f.write(f"C {n} {len(gates)}\n")
            for g in gates:
                if g[0] == "CP":
                    f.write(f"CP {g[1]} {g[2]} {g[3]!r}\n")
                else:
                    f.write(" ".join(str(x) for x in g) + "\n")


if __name__ == "__main__":
    write(os.path.join(HERE, "batch_circuits_seed42.txt"), generate())
