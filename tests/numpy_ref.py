"""Brute-force numpy models used to pin the oracle (independent of its index
arithmetic): the state is a tensor with one axis per qubit, axis q = qubit q
(qubit 0 is the most significant bit of the basis index), and a k-qubit gate
is a tensordot over its target axes with targets[0] as the matrix's MSB."""

import numpy as np

H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2.0)
X = np.array([[0, 1], [1, 0]], dtype=complex)
CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)


def cphase(theta):
    return np.diag([1, 1, 1, np.exp(1j * theta)])


def apply(state, n, mat, targets):
    k = len(targets)
    t = state.reshape((2,) * n)
    u = mat.reshape((2,) * (2 * k))
    out = np.tensordot(u, t, axes=(list(range(k, 2 * k)), list(targets)))
    # tensordot puts the gate's output axes first; move them back.
    rest = [q for q in range(n) if q not in targets]
    order = list(targets) + rest
    inv = np.argsort(order)
    return np.transpose(out, inv).reshape(-1)


def random_unitary(rng, dim):
    z = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def interleave(c):
    out = np.empty(2 * c.size)
    out[0::2] = c.real.ravel()
    out[1::2] = c.imag.ravel()
    return out


def grover_p_marked(n, m, k):
    theta = np.arcsin(np.sqrt(m / 2.0 ** n))
    return np.sin((2 * k + 1) * theta) ** 2
