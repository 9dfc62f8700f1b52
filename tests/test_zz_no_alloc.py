"""th_shell_spectrum_noalloc through the C-ABI: the same shell spectrum and mode counts as th_shell_spectrum
(counts bit for bit, sums to FP64 rounding) on two calls in a row, the second on a smaller grid, in FP64 and
FP32; a shell count above its static scratch is refused with an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from tarang_amd import Context
    return Context(0)


def _spectra(plan, A, n, noalloc):
    from tarang_amd import abi
    L = abi.lib()
    size = L.th_shell_ek_size(plan.h)
    Sk, cnt = np.zeros(size), np.zeros(size)
    fn = L.th_shell_spectrum_noalloc if noalloc else L.th_shell_spectrum
    rc = fn(plan.h, C.c_void_p(A.data_ptr()), None, n, Sk.ctypes.data_as(C.POINTER(C.c_double)),
            cnt.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, L.th_last_error()
    return Sk, cnt


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_shell_spectrum_matches(ctx, prec):
    from tarang_amd import Plan
    for N in [(64, 64, 64), (24, 16, 32)]:            # the second call on a smaller grid
        print(f"noalloc shell spectrum N={N} {prec}", flush=True)
        p = Plan(ctx, N, precision=prec)
        rng = np.random.default_rng(sum(N))
        A = p.to_dev(rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape))
        for n in (0, 2):
            ref, rcnt = _spectra(p, A, n, False)
            got, gcnt = _spectra(p, A, n, True)
            assert np.array_equal(gcnt, rcnt)
            assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), (N, n)


def test_noalloc_refuses_more_shells_than_its_scratch(ctx):
    from tarang_amd import Plan, abi
    L = abi.lib()
    p = Plan(ctx, (256, 8, 8), kfactor=(4.0, 4.0, 4.0), dealias=False)   # shell_ek_size 514 > 512
    assert L.th_shell_ek_size(p.h) > 512
    A = p.empty(1)[0]
    Sk = np.zeros(L.th_shell_ek_size(p.h))
    rc = L.th_shell_spectrum_noalloc(p.h, C.c_void_p(A.data_ptr()), None, 0, Sk.ctypes.data_as(C.POINTER(C.c_double)),
                                     None)
    assert rc != 0
    assert b"exceeds the maximum" in L.th_last_error()
