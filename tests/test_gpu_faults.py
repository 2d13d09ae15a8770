"""Error paths forced from the host.  TMFQS_HIP_FAULT=<kind>@<k> makes the
k-th device allocation, kernel launch or synchronization after a handle (or
batch) is created report an error instead of calling HIP, so every kernel that
does run exits normally.  The call raises the mapped exception, the handle can
still be destroyed, and the next handle computes the right state.

Tolerance: complex128 amplitudes within 1e-12 of the DFT of the basis state
(the QFT of |3>)."""

import numpy as np
import pytest

import tmfqsfullstate_amd as T
from tmfqsfullstate_amd import sim as S
from tmfqsfullstate_amd.batch import Batch

pytestmark = pytest.mark.gpu

N = 14
TOL = 1e-12


def want():
    e = np.zeros(1 << N, dtype=complex)
    e[3] = 1
    return np.fft.ifft(e, norm="ortho")


def cplx(x):
    return x[0::2] + 1j * x[1::2]


def run_handle():
    ops, cnt = S.qft_ops(N)
    with T.HipState() as s:
        s.init_basis(N, 3)
        s.run(ops, cnt)
        return cplx(s.export())


def run_register():
    ops, cnt = S.qft_ops(N)
    reg = S.Register.basis(N, 3)
    try:
        reg.run(ops, cnt)
        return reg.amplitudes()
    finally:
        reg.close()


def run_batch():
    ops, cnt = S.qft_ops(N)
    arr = (T.Op * cnt)()
    for i in range(cnt):
        arr[i] = ops[i]
    with Batch([(N, arr, cnt)], init_states=[3]) as b:
        b.run()
        return cplx(b.export(0))


@pytest.mark.parametrize("path", [run_handle, run_register, run_batch], ids=["handle", "register", "batch"])
@pytest.mark.parametrize("kind,code", [("alloc", T.ERR_OUT_OF_MEMORY), ("launch", None), ("sync", None)])
def test_injected_fault_raises_and_next_handle_works(monkeypatch, path, kind, code):
    ref = want()
    raised = 0
    for k in range(1, 7):
        print(path.__name__, kind, k, flush=True)
        monkeypatch.setenv("TMFQS_HIP_FAULT", f"{kind}@{k}")
        try:
            got = path()
        except T.HipError as e:
            raised += 1
            if code is not None:
                assert e.code == code, e
            else:
                assert e.code in (T.ERR_DEVICE, T.ERR_RUNTIME), e
            continue
        # past the last call of this kind: nothing was injected
        assert np.abs(got - ref).max() <= TOL
    monkeypatch.delenv("TMFQS_HIP_FAULT")
    assert raised >= 1
    assert np.abs(path() - ref).max() <= TOL
