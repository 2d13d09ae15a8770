"""GPU parity of the C-ABI backend against the CPU oracle.

Tolerances: complex128 max-abs amplitude error <= 1e-12
against the reference dense backend; per-operation sweeps and every fused
operation except the QFT phase-ladder fusion reproduce it bit for bit
(np.array_equal); measured indices are bit-exact.  complex64: fidelity error
<= 1e-6 against the FP64 oracle.
"""

import numpy as np
import pytest

import oracle as O
import tmfqsfullstate_amd as T
from tests import numpy_ref as R
from tests.register_driver import Register

pytestmark = pytest.mark.gpu

TOL = 1e-12


def st(precision=T.C128, **cfg):
    return T.HipState(precision=precision, **cfg)


def cplx(inter):
    return inter[0::2] + 1j * inter[1::2]


def rand_state(n, seed):
    return O.random_phase(n, seed)


MODES = [dict(fusion=0), dict(fusion=1, exact_phases=1), dict(fusion=1)]


@pytest.mark.parametrize("mode", MODES, ids=["unfused", "fused-exact", "fused"])
def test_single_ops_bitexact(mode):
    n = 10
    init = rand_state(n, 1)
    ops = []
    for q in range(n):
        ops.append([(O.OP_H, q)])
        ops.append([(O.OP_X, q)])
    for c, t in [(0, 9), (9, 0), (3, 4), (5, 2), (8, 7)]:
        ops.append([(O.OP_CX, c, t)])
        ops.append([(O.OP_CPHASE, c, t, 0.37 + c)])
        ops.append([(O.OP_SWAP, c, t)])
    for seq in ops:
        s = st(**mode).load(n, init)
        s.run_gates(seq)
        got = s.export()
        want = O.run_circuit(n, seq, init=init).amplitudes()
        assert np.array_equal(got, want), seq


@pytest.mark.parametrize("mode", MODES, ids=["unfused", "fused-exact", "fused"])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_matrix_gates_bitexact(mode, k):
    n = 9
    rng = np.random.default_rng(k)
    init = rand_state(n, k)
    u = R.random_unitary(rng, 1 << k)
    qubits = [int(q) for q in rng.choice(n, size=k, replace=False)]
    s = st(**mode).load(n, init)
    s.gate(u, qubits)
    want = O.apply_gate(init, n, R.interleave(u), qubits)
    assert np.array_equal(s.export(), want)


@pytest.mark.parametrize("mode", MODES, ids=["unfused", "fused-exact", "fused"])
def test_random_circuits(mode):
    rng = np.random.default_rng(7)
    for n in (6, 13):
        init = rand_state(n, n)
        seq = []
        for _ in range(400):
            kind = int(rng.integers(0, 5))
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            seq.append((kind, a, b, float(rng.uniform(0, 2 * np.pi))))
        s = st(**mode).load(n, init)
        s.run_gates(seq)
        got = s.export()
        want = O.run_circuit(n, seq, init=init).amplitudes()
        assert np.array_equal(got, want), (mode, n)


@pytest.mark.parametrize("precision", [T.C128, T.C64], ids=["c128", "c64"])
@pytest.mark.parametrize("n", [11, 13, 16])
def test_relabeling_passes(precision, n):
    # A gate run ending in swaps leaves its data under a bit relabeling that
    # one pass puts back: k_involution when the relabeling is its own inverse
    # (disjoint swaps; complex128 from 11 qubits, complex64 from 13) and
    # k_bitswap passes otherwise (a 3-cycle).  Data moves only, so the fused
    # run equals the oracle bit for bit in complex128 and the unfused
    # (one sweep per swap) run in complex64.
    rng = np.random.default_rng(100 + n)
    init = rand_state(n, n + 1)
    for trial in range(4):
        q = [int(x) for x in rng.permutation(n)]
        seq = [(O.OP_H, a, 0, 0.0) for a in range(n)]
        if trial < 3:
            pairs = 1 + 2 * trial if trial < 2 else n // 2
            seq += [(O.OP_SWAP, q[2 * i], q[2 * i + 1], 0.0) for i in range(pairs)]
        else:
            seq += [(O.OP_SWAP, q[0], q[1], 0.0), (O.OP_SWAP, q[1], q[2], 0.0)]
        s = st(precision, fusion=1).load(n, init)
        s.profile(True)
        s.run_gates(seq)
        got = s.export()  # runs the queued gates
        prof = s.profile_read()
        if trial < 3 and n >= (11 if precision == T.C128 else 13):
            assert prof["k_involution"]["launches"] == 1, (n, trial, prof)
        if trial == 3:
            assert prof["k_bitswap"]["launches"] >= 2, (n, prof)
        if precision == T.C128:
            want = O.run_circuit(n, seq, init=init).amplitudes()
        else:
            u = st(precision, fusion=0).load(n, init)
            u.run_gates(seq)
            want = u.export()
        assert np.array_equal(got, want), (n, trial)


QFT_MODES = MODES + [dict(fusion=1, qft_plan=1)]


@pytest.mark.parametrize("mode", QFT_MODES, ids=["unfused", "fused-exact", "fused", "fused-in-place"])
@pytest.mark.parametrize("n", [1, 2, 3, 9, 12, 13, 14, 16, 20])
def test_qft(mode, n):
    init = rand_state(n, 7)
    ops, cnt = O.qft_ops(n)
    s = st(**mode).load(n, init)
    reg = Register(s, n)
    reg.run(ops, cnt)
    got = s.export()
    want = O.run_circuit(n, (ops, cnt), init=init).amplitudes()
    if mode.get("fusion", 1) == 0 or mode.get("exact_phases", 0):
        assert np.array_equal(got, want)
    else:
        assert np.abs(cplx(got) - cplx(want)).max() <= TOL
    dft = np.fft.ifft(cplx(init), norm="ortho")
    assert np.abs(cplx(got) - dft).max() <= TOL


@pytest.mark.parametrize("precision", [T.C128, T.C64], ids=["c128", "c64"])
def test_qft_tail_pair_shapes(precision):
    # The fused tail exchanges 64 x 64 tile pairs {M, rev(M)} over the n - 12
    # middle bits: n = 12, 13 have one or two palindromic tiles, odd n - 12 has
    # a middle bit, and every n here also runs the eight-target block passes.
    # n = 13 .. 21 also start with a partial block pass of the leading
    # (n - 6) mod 8 targets (K = 7, 1, 3, 5, 6, 7), and must run on the QFT
    # kernels alone.
    for n in (12, 13, 14, 15, 17, 19, 20, 21):
        init = rand_state(n, 3)
        s = st(precision, qft_plan=1).load(n, init)
        ops, cnt = O.qft_ops(n)
        s.profile(True)
        Register(s, n).run(ops, cnt)
        got = s.amplitudes()
        used = set(s.profile_read())
        if n >= 13:
            assert used == {"k_qft_block", "k_qft_tail_reverse"}, (n, used)
        want = np.fft.ifft(cplx(init), norm="ortho")
        if precision == T.C128:
            assert np.abs(got - want).max() <= TOL, n
        else:
            assert np.abs(got - want).max() < 1e-5, n
            assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6, n


@pytest.mark.parametrize("precision", [T.C128, T.C64], ids=["c128", "c64"])
def test_qft_ten_target_plan_shapes(precision):
    # The default plan: a leading pass of (n - 10) mod 10 targets (K = 3..9
    # here, none at n = 20), ten-target passes down to bit 10, and the last ten
    # targets fused with the bit reversal out of place.  n = 13 (complex128)
    # and 14 (complex64) have no bits between the tail's passengers and its
    # targets; complex64 n = 13 keeps the in-place plan.
    lo = 13 if precision == T.C128 else 14
    for n in (13, 14, 15, 16, 19, 20, 21, 23, 25):
        init = rand_state(n, 5)
        s = st(precision).load(n, init)
        ops, cnt = O.qft_ops(n)
        s.profile(True)
        Register(s, n).run(ops, cnt)
        got = s.amplitudes()
        prof = s.profile_read()
        if n >= lo:
            assert set(prof) == {"k_qft_wide_block", "k_qft_wide_tail"}, (n, set(prof))
            assert prof["k_qft_wide_block"]["launches"] == (n - 10 + 9) // 10, n
            assert prof["k_qft_wide_tail"]["launches"] == 1
        want = np.fft.ifft(cplx(init), norm="ortho")
        if precision == T.C128:
            assert np.abs(got - want).max() <= TOL, n
            if n <= 16:
                ref = cplx(O.run_circuit(n, (ops, cnt), init=init).amplitudes())
                assert np.abs(got - ref).max() <= TOL, n
        else:
            assert np.abs(got - want).max() < 1e-5, n
            assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6, n
        # the state moved into the handle's second buffer: later operations
        # and a second QFT (back into the first buffer) see it there
        s.h(0)
        Register(s, n).run(ops, cnt)
        got2 = s.amplitudes()
        h0 = want.reshape(2, -1)
        h0 = np.concatenate([(h0[0] + h0[1]), (h0[0] - h0[1])]) / np.sqrt(2)
        want2 = np.fft.ifft(h0, norm="ortho")
        assert np.abs(got2 - want2).max() <= (TOL if precision == T.C128 else 1e-5), n


def _op_tuples(ops, cnt):
    return [(ops[i].kind, ops[i].q0, ops[i].q1, ops[i].theta) for i in range(cnt)]


@pytest.mark.parametrize("plan", [0, 1], ids=["ten-target", "in-place"])
@pytest.mark.parametrize("precision", [T.C128, T.C64], ids=["c128", "c64"])
def test_qft_inside_longer_gate_runs(precision, plan):
    # Two QFTs with other gates before, between and after them, all in one
    # flush: each QFT is still planned as a unit and runs on the QFT kernels.
    n = 16
    rng = np.random.default_rng(21)

    def others(k):
        seq = []
        for _ in range(k):
            kind = int(rng.choice([O.OP_H, O.OP_CX, O.OP_CPHASE, O.OP_SWAP]))
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            seq.append((kind, a, b, float(rng.uniform(0, 2 * np.pi))))
        return seq

    qft = _op_tuples(*O.qft_ops(n))
    seq = others(12) + qft + others(7) + qft + others(5)
    init = rand_state(n, 9)
    s = st(precision, qft_plan=plan).load(n, init)
    s.profile(True)
    s.run_gates(seq)
    got = s.amplitudes()
    prof = s.profile_read()
    if plan:
        assert prof["k_qft_tail_reverse"]["launches"] == 2
        assert prof["k_qft_block"]["launches"] == 4  # a partial (two targets) and a full block per QFT
    else:
        assert prof["k_qft_wide_tail"]["launches"] == 2
        assert prof["k_qft_wide_block"]["launches"] == 2  # one six-target pass per QFT
    want = cplx(O.run_circuit(n, seq, init=init).amplitudes())
    if precision == T.C128:
        assert np.abs(got - want).max() <= TOL
    else:
        assert np.abs(got - want).max() < 1e-5
        assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6


def test_qft_known_answers():
    # test_circuit.cpp:76-106
    for src in (0, 5):
        s = st().init_basis(3, src)
        reg = Register(s, 3)
        ops, cnt = O.qft_ops(3)
        reg.run(ops, cnt)
        y = np.arange(8)
        assert np.allclose(s.amplitudes(), np.exp(2j * np.pi * src * y / 8) / np.sqrt(8), atol=1e-6)


@pytest.mark.parametrize("case", ["real", "mixed", "complex"])
def test_diffusion_chains_bitexact(case):
    # Phase flips and materialized diffusions (300 steps: one sweep at the
    # default block, three at grover_block=128) from a loaded state.  "real": every imaginary part is +0 or -0, so every
    # wavefront takes the real-only chain.  "mixed": two imaginary parts
    # +-0.25 cancel exactly in the amplitude sum, so Im(2m) stays 0 and the
    # two wavefronts holding them take the full chain beside real ones.
    # "complex": Im(2m) != 0, the full chain everywhere.  Bit for bit against
    # the oracle, signs of zeros included.
    n = 14
    rng = np.random.default_rng(11)
    inter = np.zeros(2 << n)
    inter[0::2] = rng.uniform(-1, 1, 1 << n) / 128
    inter[1::2][rng.random(1 << n) < 0.5] = -0.0
    if case == "mixed":
        inter[2 * 100 + 1], inter[2 * 9000 + 1] = 0.25, -0.25
    if case == "complex":
        inter[2 * 100 + 1] = 0.25
    seq = []
    for it in range(150):
        seq.append((O.OP_PHASE_FLIP, 0, 0, 0.0, 77 + (it % 3)))
        seq.append((O.OP_DIFFUSE_MAT,))
        seq.append((O.OP_DIFFUSE_MAT,))
    ops, cnt = O.make_ops(seq)
    want = O.run_circuit(n, (ops, cnt), init=inter).amplitudes()
    for block in (0, 128):
        s = st(grover_block=block).load(n, inter)
        reg = Register(s, n)
        reg.run(ops, cnt)
        got = s.export()
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), block
        if case != "complex":
            assert np.all(got[1::2][np.r_[0:100, 101:9000, 9001:1 << n]] == 0.0)


@pytest.mark.parametrize("materialized", [True, False])
@pytest.mark.parametrize("mode", MODES, ids=["unfused", "fused-exact", "fused"])
def test_grover_vs_oracle(materialized, mode):
    for n, marked in ((10, [511]), (12, [3, 2047, 4000])):
        ops, cnt = O.grover_ops(n, marked, materialized)
        s = st(**mode).init_basis(n, 0)
        reg = Register(s, n)
        reg.run(ops, cnt)
        got = reg.amplitudes()
        oreg = O.run_circuit(n, (ops, cnt))
        want = cplx(oreg.amplitudes())
        assert np.abs(got - want).max() <= TOL
        if materialized:  # bit for bit, signs of zeros included
            assert np.array_equal(s.export().view(np.uint64), oreg.amplitudes().view(np.uint64))
        for seed in (123456, 1, 2, 3):
            rnd = O.mt_unit_draws(seed, 1)[0]
            assert reg.measure(rnd) == oreg.measure(rnd)
        k = O.grover_iterations(n, len(set(marked)))
        p = sum(abs(got[m]) ** 2 for m in set(marked))
        assert abs(p - R.grover_p_marked(n, len(set(marked)), k)) < 1e-10


@pytest.mark.parametrize("n", [0, 1, 5, 12, 15])
def test_amplitude_sum_bitexact(n):
    # signed sequential sums: random phases (partial sums cross zero), a
    # uniform superposition (one binade per chunk), a sparse state
    states = [rand_state(n, 2)]
    u = np.zeros(2 << n)
    u[0::2] = 2.0 ** (-n / 2)
    states.append(u)
    sp = np.zeros(2 << n)
    sp[2 * ((1 << n) - 1)] = -0.75
    sp[1] = 0.25
    states.append(sp)
    for init in states:
        s = st().load(n, init)
        re, im = O.ctypes.c_double(), O.ctypes.c_double()
        O.lib().orc_amplitude_sum(init.ctypes.data_as(O._dp), n, O.ctypes.byref(re), O.ctypes.byref(im))
        got = s.amplitude_sum()
        assert got.real == re.value and got.imag == im.value


@pytest.mark.parametrize("n", [4, 11, 14, 17])
def test_sample_and_total_probability_bitexact(n):
    init = rand_state(n, 3)
    s = st().load(n, init)
    ops, cnt = O.qft_ops(n)
    Register(s, n).run(ops, cnt)
    amps = s.export()
    assert s.total_probability() == O.total_probability(amps, n)
    rng = np.random.default_rng(n)
    for rnd in list(rng.random(40)) + [0.0, 1e-300, 0.5, float(np.nextafter(1.0, 0.0))]:
        try:
            want = O.sample(amps, n, rnd)
        except O.OracleError as e:
            with pytest.raises(T.HipError) as got:
                s.sample(rnd)
            assert got.value.code == T.ERR_RUNTIME and e.code == O.RUNTIME_ERROR
            continue
        assert s.sample(rnd) == want, rnd


def test_sample_boundaries_exact():
    # rnd exactly at cumulative boundaries of the sequential chain
    n = 12
    init = rand_state(n, 9)
    s = st().load(n, init)
    p = init[0::2] ** 2 + init[1::2] ** 2
    cum = 0.0
    bounds = []
    for i, pi in enumerate(p):
        cum += pi
        if i % 97 == 0:
            bounds.append(cum)
    for b in bounds:
        for r in (b, np.nextafter(b, 0), np.nextafter(b, 2)):
            if r >= 1.0:
                continue
            assert s.sample(float(r)) == O.sample(init, n, float(r))


@pytest.mark.parametrize("n,basis", [(13, 5000), (14, 12288), (5, 31)])
def test_sample_zero_draw_before_leading_zero_chunks(n, basis):
    # rnd == 0 is a legal draw: the sequential chain stops at state 0, whose
    # cumulative probability 0 already satisfies rnd <= cum, even when whole
    # leading chunks of the state hold zero probability
    s = st().init_basis(n, basis)
    amps = s.export()
    want = O.sample(amps, n, 0.0)
    assert want == 0
    assert s.sample(0.0) == want
    oreg = O.Register(n, amps)
    oreg.inversion(False, complex(0.0, 0.0))  # overlay scale -1, bias 0
    assert s.sample_affine(0.0, complex(-1, 0), complex(0, 0)) == oreg.measure(0.0) == 0
    # a later draw still lands on the basis state
    assert s.sample(0.5) == O.sample(amps, n, 0.5) == basis


@pytest.mark.parametrize("fusion", [0, 1])
def test_complex64_amplitude_reads_match_storage(fusion):
    n = 6
    s = st(T.C64, fusion=fusion).init_basis(n, 0)
    s.set_amplitude(5, 0.1 - 0.3j)
    want = complex(np.float32(0.1), np.float32(-0.3))
    assert s.amplitude(5) == want
    assert s.probability(5) == want.real * want.real + want.imag * want.imag
    s.phase_flip(5)
    assert s.amplitude(5) == -want
    got = s.export()
    assert complex(got[10], got[11]) == s.amplitude(5)


def test_affine_overlay_reads():
    n = 8
    init = rand_state(n, 4)
    s = st().load(n, init)
    scale, bias = complex(-0.7, 0.2), complex(0.01, -0.03)
    oreg = O.Register(n, init)
    oreg.inversion(False, complex(0.005, -0.015))  # overlay scale -1, bias 2m
    got_tp = s.total_probability_affine(complex(-1.0, 0.0), complex(0.01, -0.03))
    assert got_tp == oreg.total_probability()
    for rnd in (0.1, 0.5, 0.93):
        assert s.sample_affine(rnd, complex(-1, 0), complex(0.01, -0.03)) == oreg.measure(rnd)
    s.affine(scale, bias)
    want = init.copy()
    O.lib().orc_affine(want.ctypes.data_as(O._dp), n, scale.real, scale.imag, bias.real, bias.imag)
    assert np.array_equal(s.export(), want)


def test_state_io_and_clone():
    n = 7
    s = st().init_uniform(n, [5, 3, 3, 100])
    a = s.amplitudes()
    assert np.allclose(a[[3, 5, 100]], 1 / np.sqrt(3)) and np.count_nonzero(a) == 3
    c = s.clone()
    s.h(0)
    assert np.array_equal(c.amplitudes(), a)  # clone is independent
    s.set_amplitude(9, 0.25 - 0.5j)
    assert s.amplitude(9) == 0.25 - 0.5j
    assert abs(s.probability(9) - 0.3125) < 1e-15
    part = s.export(8, 4)
    full = s.export()
    assert np.array_equal(part, full[16:24])


def test_errors():
    s = st()
    with pytest.raises(T.HipError) as e:
        s.h(0)  # not initialised
    assert e.value.code == T.ERR_LOGIC
    s.init_basis(2, 0)
    with pytest.raises(T.OutOfRange):
        s.h(2)
    with pytest.raises(T.OutOfRange):
        s.amplitude(4)
    with pytest.raises(T.InvalidArgument):
        s.cx(1, 1)
    with pytest.raises(T.InvalidArgument):
        s.gate(R.CX, [1, 1])
    with pytest.raises(T.OutOfRange):
        s.gate(R.H, [2])
    with pytest.raises(T.InvalidArgument):
        s.sample(1.0)
    with pytest.raises(T.InvalidArgument):
        s.init_uniform(3, [])
    with pytest.raises(T.OutOfRange):
        s.init_uniform(3, [0, 8])
    with pytest.raises(T.InvalidArgument):
        s.load(2, np.zeros(6))


def test_complex64_fidelity():
    for n in (12, 18):
        init = rand_state(n, 11)
        s = st(T.C64).load(n, init)
        ops, cnt = O.qft_ops(n)
        Register(s, n).run(ops, cnt)
        got = s.amplitudes()
        want = np.fft.ifft(cplx(init), norm="ortho")
        fid = abs(np.vdot(want, got)) ** 2
        assert 1 - fid <= 1e-6, (n, 1 - fid)
        assert np.abs(got - want).max() < 1e-5
    n = 14
    ops, cnt = O.grover_ops(n, [77, 9000], True)
    s = st(T.C64).init_basis(n, 0)
    reg = Register(s, n)
    reg.run(ops, cnt)
    got = reg.amplitudes()
    want = cplx(O.run_circuit(n, (ops, cnt)).amplitudes())
    fid = abs(np.vdot(want, got)) ** 2
    assert 1 - fid <= 1e-6


@pytest.mark.parametrize("n", [9, 12, 14])
def test_complex64_five_qubit_blocks(n):
    # complex64 k = 5 matrices: FP32 MFMA on tiles of >= 10 bits (n = 12, 14),
    # FP32 VALU below; against the complex128 oracle
    rng = np.random.default_rng(n)
    init = rand_state(n, n)
    s = st(T.C64).load(n, init)
    want = init.copy()
    for _ in range(6):
        u = R.random_unitary(rng, 32)
        qubits = [int(q) for q in rng.choice(n, size=5, replace=False)]
        s.gate(u, qubits)
        want = O.apply_gate(want, n, R.interleave(u), qubits)
    got = s.amplitudes()
    w = cplx(want)
    assert np.abs(got - w).max() < 2e-6
    assert 1.0 - abs(np.vdot(w, got)) ** 2 / (np.vdot(w, w).real * np.vdot(got, got).real) <= 1e-6


@pytest.mark.parametrize("n", [11, 12, 14, 17])
def test_complex128_five_qubit_blocks_mfma(n):
    # complex128 k = 5 matrices on tiles of >= 11 bits run on FP64 MFMA
    # (FP64 accumulation in another order than the dense backend's): within
    # 1e-12 of the oracle; the exact mode keeps the reference's VALU order and
    # its bits.  Matrices mixed with other gates in the same tiled passes.
    rng = np.random.default_rng(100 + n)
    init = rand_state(n, n)
    seq, mats, quals = [], [], []
    for i in range(8):
        qubits = [int(q) for q in rng.choice(n, size=5, replace=False)]
        seq.append(("u", R.random_unitary(rng, 32), qubits))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        seq.append(("h", a))
        seq.append(("cp", a, b, float(rng.uniform(0, 6.3))))
    want = init.copy()
    for g in seq:
        if g[0] == "u":
            want = O.apply_gate(want, n, R.interleave(g[1]), g[2])
        elif g[0] == "h":
            want = O.run_circuit(n, [(O.OP_H, g[1])], init=want).amplitudes()
        else:
            want = O.run_circuit(n, [(O.OP_CPHASE, g[1], g[2], g[3])], init=want).amplitudes()
    for exact in (0, 1):
        s = st(exact_phases=exact).load(n, init)
        s.begin_batch()
        for g in seq:
            if g[0] == "u":
                s.gate(g[1], g[2])
            elif g[0] == "h":
                s.h(g[1])
            else:
                s.cphase(g[1], g[2], g[3])
        s.end_batch()
        got = s.export()
        if exact:
            assert np.array_equal(got, want)
        else:
            assert np.abs(cplx(got) - cplx(want)).max() <= 1e-12


def test_qft_large_vs_fft():
    # n = 24: a size-independent property (the DFT) rather than the oracle.
    n = 24
    rng = np.random.default_rng(0)
    ph = rng.uniform(0, 2 * np.pi, 1 << n)
    init = np.empty(2 << n)
    init[0::2] = np.cos(ph) * 2.0 ** (-n / 2)
    init[1::2] = np.sin(ph) * 2.0 ** (-n / 2)
    s = st().load(n, init)
    ops, cnt = O.qft_ops(n)
    Register(s, n).run(ops, cnt)
    got = s.amplitudes()
    want = np.fft.ifft(cplx(init), norm="ortho")
    assert np.abs(got - want).max() <= TOL


@pytest.mark.parametrize("precision,n", [(T.C128, 31), (T.C64, 32)], ids=["c128-n31", "c64-n32"])
def test_qft_largest_sizes_basis_closed_form(precision, n):
    # The largest QFTs a test can afford (32 GiB states), checked against the
    # closed form QFT|s> = sum_y e^{2 pi i s y / N} |y> / sqrt(N) on three
    # slices: n = 31 starts with a one-target partial pass, n = 32 with a
    # two-target one, and the tail exchanges 2^(n-12) tiles.
    from tmfqsfullstate_amd import sim as S

    N = 1 << n
    src = 0x2468ACE1 % N
    print(f"QFT-{n} of |{src}>", flush=True)
    with T.HipState(precision=precision) as s:
        s.init_basis(n, src)
        ops, cnt = S.qft_ops(n)
        s.run(ops, cnt)
        s.synchronize()
        assert abs(s.total_probability() - 1.0) < (1e-12 if precision == T.C128 else 1e-5)
        for first in (0, N // 2 - 1000, N - 4096):
            got = cplx(s.export(first, 4096))
            y = np.arange(first, first + 4096, dtype=np.uint64)
            ph = (np.uint64(src) * y) % np.uint64(N)  # exact: src * y < 2^64
            want = np.exp(2j * np.pi * (ph.astype(np.float64) / N)) / np.sqrt(N)
            err = np.abs(got - want).max() * np.sqrt(N)  # relative to the amplitude size
            assert err <= (1e-10 if precision == T.C128 else 1e-5), (first, err)


@pytest.mark.parametrize("precision", [T.C128, T.C64], ids=["c128", "c64"])
def test_load_export_round_trip_across_staging_chunks(precision):
    # n = 24: 256 MiB complex128 / 128 MiB complex64, several 64 MiB staging
    # chunks each way; partial exports start and end inside chunks.
    n = 24
    rng = np.random.default_rng(3)
    inter = rng.uniform(-1, 1, 2 << n)
    if precision == T.C64:
        inter = inter.astype(np.float32).astype(np.float64)  # exactly representable
    s = st(precision=precision).load(n, inter)
    assert np.array_equal(s.export(), inter)
    per = (64 << 20) // (16 if precision == T.C128 else 8)  # states per staging chunk
    for first, count in ((per - 5, 11), (per // 2, per + 7), ((1 << n) - 3, 3), (0, 1)):
        got = s.export(first, count)
        assert np.array_equal(got, inter[2 * first:2 * (first + count)]), (first, count)
    with pytest.raises(T.OutOfRange):  # a range past the last state
        s.export(per // 2, 1 << n)
    s.close()
