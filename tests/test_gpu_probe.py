"""GPU parity of th_dft16_probe: the 16-point DFT by the vector-ALU codelet and by the FP32 MFMA
DFT-as-GEMM both equal numpy's FFT, (W / 4)^reps x with W = exp(sign 2 pi i k n / 16), to FP32 accuracy
(max error <= 5e-6 of the largest output magnitude: a few FP32 ulps of the 16- or 32-term sums per
application)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from tarang_amd import Context
    return Context(0)


def ref(x, sign, reps):
    y = x.astype(np.complex128)
    for _ in range(reps):
        y = (np.fft.fft(y, axis=0) if sign < 0 else 16 * np.fft.ifft(y, axis=0)) / 4
    return y


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("sign,reps", [(-1, 1), (1, 1), (-1, 3)])
def test_dft16_probe_matches_numpy(ctx, method, sign, reps):
    import torch
    nvec = 4096
    rng = np.random.default_rng(7 + reps)
    x = (rng.standard_normal((16, nvec)) + 1j * rng.standard_normal((16, nvec))).astype(np.complex64)
    xd = torch.from_numpy(x).to("cuda:0")
    yd = torch.empty_like(xd)
    print(f"dft16 probe method={method} sign={sign} reps={reps}", flush=True)
    ctx.dft16_probe(xd, yd, method, sign=sign, reps=reps)
    y = yd.cpu().numpy().astype(np.complex128)
    r = ref(x, sign, reps)
    assert np.abs(y - r).max() <= 5e-6 * np.abs(r).max()


def test_dft16_probe_rejects_bad_sizes(ctx):
    import torch
    x = torch.zeros((16, 300), dtype=torch.complex64, device="cuda:0")
    with pytest.raises(RuntimeError, match="multiple of 256"):
        ctx.dft16_probe(x, torch.empty_like(x), 0)
