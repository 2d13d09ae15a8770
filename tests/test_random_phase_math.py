"""randomPhase arithmetic the device kernels use, checked on the host
(csrc/phase_math.h, csrc/mt_jump.cpp):

* the jump-ahead reaches exactly the outputs numpy's own MT19937 (legacy
  seeding = std::mt19937's init_genrand) gives after discarding J words;
* the double-double cos/sin is within 2^-74 (relative) of 120-bit mpmath
  values, including arguments next to multiples of pi/2 and the three-part
  pi/2 it reduces with sums to pi/2 within 2^-158;
* outside the flag band the rounded value equals glibc's sincos -- the
  function the reference's compiled cos/sin pair calls -- and the band holds
  about 9% of values.
"""

import ctypes
import math
import re
from pathlib import Path

import mpmath
import numpy as np
import pytest
from numpy.random import MT19937

import tmfqsfullstate_amd as tm

mpmath.mp.prec = 120
_libm = ctypes.CDLL("libm.so.6")
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_libm.sincos.restype = None


def glibc_sincos(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def numpy_words(seed, skip, count):
    b = MT19937(0)
    b._legacy_seeding(seed)
    if skip:
        b.random_raw(skip)
    return b.random_raw(count).astype(np.uint32)


def test_numpy_legacy_seeding_is_std_mt19937():
    # std::mt19937's default seed 5489 gives 3499211612 first ([rand.predef]:
    # the 10000th output is 4123659995).
    w = numpy_words(5489, 0, 10000)
    assert w[0] == 3499211612 and w[-1] == 4123659995


@pytest.mark.parametrize("seed", [7, 123456, 0, 2**32 - 1])
@pytest.mark.parametrize("jump", [0, 1, 397, 623, 624, 625, 4096, 2**20, 3 * 2**21 + 17, 2**31 + 12345])
def test_jump_ahead_matches_sequential_generation(seed, jump):
    got = tm.mt_jump_outputs(seed, jump, 700)
    if jump > 2**26:
        # 2^31 words through numpy would take minutes: check against a
        # composed jump instead (2^31 then 12345) and its tail agreement.
        a = tm.mt_jump_outputs(seed, jump - 12345, 12345 + 700)
        assert np.array_equal(got, a[12345:])
        return
    assert np.array_equal(got, numpy_words(seed, jump, 700))


def test_pio2_parts():
    # The constants as phase_math.h spells them.
    text = (Path(tm.__file__).parent / "csrc" / "phase_math.h").read_text()
    part = {m.group(1): float.fromhex(m.group(2))
            for m in re.finditer(r"constexpr double (kPio2\w+|kTwoOverPi) = (-?0x[0-9a-fp.+-]+);", text)}
    assert set(part) == {"kPio2Hi", "kPio2Mid", "kPio2Lo", "kTwoOverPi"}
    err = abs(mpmath.mpf(part["kPio2Hi"]) + mpmath.mpf(part["kPio2Mid"]) + mpmath.mpf(part["kPio2Lo"]) - mpmath.pi / 2)
    assert err < mpmath.mpf(2) ** -158
    assert abs(mpmath.mpf(part["kTwoOverPi"]) - 2 / mpmath.pi) < mpmath.mpf(2) ** -53


def phases(count, seed):
    """generate_canonical<double, 53> * 2 acos(-1) from mt19937 words."""
    w = numpy_words(seed, 0, 2 * count).astype(np.float64)
    r = (w[0::2] + w[1::2] * 4294967296.0) * 2.0**-64
    r = np.where(r >= 1.0, np.nextafter(1.0, 0.0), r)
    return r * (2.0 * math.acos(-1.0))


def special_phases():
    xs = [0.0, 5e-324, 1e-300, 1e-20, 1e-8, 0.5, math.pi / 4, 2 * math.acos(-1.0) - 2**-50]
    for k in range(1, 5):
        c = k * math.pi / 2
        v = c
        for _ in range(4):
            xs.append(v)
            v = np.nextafter(v, 0.0)
        v = c
        for _ in range(4):
            v = np.nextafter(v, 10.0)
            xs.append(v)
    return [x for x in xs if 0.0 <= x < 2 * math.pi]


def test_double_double_sincos_accuracy():
    xs = list(phases(3000, 99)) + special_phases()
    worst = 0.0
    for x in xs:
        out, _ = tm.phase_eval(float(x))
        mx = mpmath.mpf(float(x))
        for got_hi, got_lo, ref in ((out[0], out[1], mpmath.cos(mx)), (out[2], out[3], mpmath.sin(mx))):
            if ref == 0:
                assert got_hi == 0.0 and got_lo == 0.0
                continue
            rel = abs((mpmath.mpf(got_hi) + mpmath.mpf(got_lo) - ref) / ref)
            worst = max(worst, float(rel))
    assert worst < 2.0**-74, worst


def test_unflagged_values_equal_glibc_sincos():
    xs = np.concatenate([phases(200000, 4242), np.array(special_phases())])
    flagged = 0
    for x in xs:
        out, (fc, fs) = tm.phase_eval(float(x))
        gs, gc = glibc_sincos(float(x))
        if not fc:
            assert out[0] == gc, (x, out[0], gc)
        if not fs:
            assert out[2] == gs, (x, out[2], gs)
        flagged += fc + fs
    frac = flagged / (2 * len(xs))
    assert 0.05 < frac < 0.14, frac


def test_flags_catch_glibc_misrounding():
    # Every phase where glibc's sincos is not the double nearest the exact
    # value must be flagged (the band is what makes the device bit-exact).
    seen = 0
    for x in phases(60000, 7):
        out, (fc, fs) = tm.phase_eval(float(x))
        gs, gc = glibc_sincos(float(x))
        mx = mpmath.mpf(float(x))
        for flag, g, ref in ((fc, gc, mpmath.cos(mx)), (fs, gs, mpmath.sin(mx))):
            nearest = float(ref)  # mpmath rounds to nearest on conversion
            if g != nearest:
                seen += 1
                assert flag, (x, g, nearest)
    assert seen > 20  # glibc misrounds ~0.15% of values: the check exercised it


def test_phase_eval_rejects_out_of_range():
    with pytest.raises(tm.HipError):
        tm.phase_eval(-1.0)
    with pytest.raises(tm.HipError):
        tm.phase_eval(float("nan"))
