"""tmfqs_hip_expectation_device, the entry point that allocates nothing,
against tmfqs_hip_expectation on the same states: the same kernels, so the
same bits; complex128 also within 1e-12 of the CPU oracle's expectation.

Two calls in a row share the device scratch array, the second on a smaller
state (fewer workgroup partials than the first left behind)."""

import ctypes

import numpy as np
import pytest
import torch

import oracle as O
import tmfqsfullstate_amd as T

pytestmark = pytest.mark.gpu


def terms_for(n):
    return [(T.DIAG_Z, 1.0, 1 << 0), (T.DIAG_Z, -0.375, (1 << 3) | (1 << (n - 1))), (T.DIAG_PROJECTOR, 1.0, 0),
            (T.DIAG_PROJECTOR, 0.5, (1 << n) - 3)]


@pytest.mark.parametrize("precision", [T.C128, T.C64], ids=["complex128", "complex64"])
def test_expectation_device_matches_handle_entry_point(precision):
    dtype = np.float64 if precision == T.C128 else np.float32
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream(0).cuda_stream
    for call, n in enumerate((20, 11)):
        print(f"expectation_device call {call}: n={n} precision={precision}", flush=True)
        init = O.random_phase(n, 5 + call)
        terms = terms_for(n)
        with T.HipState(precision=precision) as s:
            s.load(n, init)
            want = s.expectation(terms)
            held = s.export()  # the amplitudes as the state holds them (complex64: FP32 values)
        amps = torch.from_numpy(held.astype(dtype)).to("cuda:0")
        assert amps.numel() == 2 << n
        T.expectation_device(amps.data_ptr(), n, terms, out.data_ptr() + 8 * call, precision, 0, stream)
        torch.cuda.synchronize(0)
        got = out[call].item()
        assert got == want, (call, got, want)
        if precision == T.C128:
            assert abs(got - O.expectation(init, n, terms)) <= 1e-12
        else:
            assert abs(got - O.expectation(init, n, terms)) <= 1e-6


def test_expectation_device_rejects_bad_arguments():
    out = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    amps = torch.zeros(2 << 4, dtype=torch.float64, device="cuda:0")
    L = T.lib()
    with pytest.raises(T.InvalidArgument):
        T.expectation_device(amps.data_ptr(), 64, [], out.data_ptr())
    with pytest.raises(T.InvalidArgument):
        T.expectation_device(amps.data_ptr(), 4, [], out.data_ptr(), precision=7)
    with pytest.raises(T.InvalidArgument):
        T.expectation_device(amps.data_ptr(), 4, [], out.data_ptr(), device=99)
    with pytest.raises(T.OutOfRange):
        T.expectation_device(amps.data_ptr(), 4, [(T.DIAG_Z, 1.0, 1 << 4)], out.data_ptr())
    arr, cnt = T.make_terms([])
    rc = L.tmfqs_hip_expectation_device(0, None, T.C128, None, 4, arr, cnt, ctypes.c_void_p(out.data_ptr()))
    assert rc == T.ERR_INVALID_ARGUMENT
