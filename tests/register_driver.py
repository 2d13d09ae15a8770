"""Drives a HipState with the exact backend call sequence of the reference's
QuantumRegister (src/tmfqs/register/quantum_register.cpp): the cached
amplitude sum recurrence (:176-189, :405-420, :447-467) and the lazy affine
overlay (:152-231, :423-444).  Test-side code: it lets backend-level tests run
whole Grover circuits before the C++ API is involved."""

import tmfqsfullstate_amd as T


class Register:
    def __init__(self, st, n):
        self.st = st
        self.n = n
        self.N = 1 << n
        self.sum = None
        self.overlay = False
        self.scale = 1 + 0j
        self.bias = 0j

    def _apply(self, v):
        if not self.overlay:
            return v
        s, b = self.scale, self.bias
        mr = s.real * v.real - s.imag * v.imag
        mi = s.real * v.imag + s.imag * v.real
        return complex(mr + b.real, mi + b.imag)

    def _remove(self, v):
        if not self.overlay:
            return v
        s, b = self.scale, self.bias
        nr, ni = v.real - b.real, v.imag - b.imag
        d = s.real * s.real + s.imag * s.imag
        return complex((nr * s.real + ni * s.imag) / d, (ni * s.real - nr * s.imag) / d)

    def _flush(self):
        if not self.overlay:
            return
        saved = self.sum
        self.st.affine(self.scale, self.bias)
        self.overlay, self.scale, self.bias, self.sum = False, 1 + 0j, 0j, saved

    def amplitude(self, s):
        return self._apply(self.st.amplitude(s))

    def logical_sum(self):
        if self.sum is None:
            if self.overlay:
                raise NotImplementedError("sum under overlay not needed by these tests")
            self.sum = self.st.amplitude_sum()
        return self.sum

    def gate(self, fn, *args):
        self._flush()
        fn(*args)
        self.sum = None

    def phase_flip(self, s):
        if self.sum is not None or self.overlay:
            lg = self.amplitude(s)
            if self.sum is not None:
                self.sum = complex(self.sum.real - 2.0 * lg.real, self.sum.imag - 2.0 * lg.imag)
            if self.overlay:
                self.st.set_amplitude(s, self._remove(complex(-lg.real, -lg.imag)))
                return
        self.st.phase_flip(s)

    def _recurrence(self, m):
        if self.sum is not None:
            self.sum = complex(2.0 * m.real * float(self.N) - self.sum.real,
                               2.0 * m.imag * float(self.N) - self.sum.imag)

    def inversion(self, materialized):
        s = self.logical_sum()
        m = complex(s.real / float(self.N), s.imag / float(self.N))
        self._recurrence(m)
        if materialized:
            self._flush()
            self.st.inversion_about_mean(m)
        else:
            sc, bi = (self.scale, self.bias) if self.overlay else (1 + 0j, 0j)
            self.scale = complex(-sc.real, -sc.imag)
            self.bias = complex(m.real * 2.0 - bi.real, m.imag * 2.0 - bi.imag)
            self.overlay = True

    def run(self, ops, count):
        st = self.st
        for i in range(count):
            op = ops[i]
            k = op.kind
            if k == T.OP_H:
                self.gate(st.h, op.q0)
            elif k == T.OP_X:
                self.gate(st.x, op.q0)
            elif k == T.OP_CX:
                self.gate(st.cx, op.q0, op.q1)
            elif k == T.OP_CPHASE:
                self.gate(st.cphase, op.q0, op.q1, op.theta)
            elif k == T.OP_SWAP:
                self.gate(st.swap, op.q0, op.q1)
            elif k == T.OP_PHASE_FLIP:
                self.phase_flip(op.state)
            elif k == T.OP_DIFFUSE_LAZY:
                self.inversion(False)
            elif k == T.OP_DIFFUSE_MAT:
                self.inversion(True)
        st.end_batch()

    def amplitudes(self):
        a = self.st.amplitudes()
        if self.overlay:
            s, b = self.scale, self.bias
            a = s * a + b
        return a

    def measure(self, rnd):
        if self.overlay:
            return self.st.sample_affine(rnd, self.scale, self.bias)
        return self.st.sample(rnd)
