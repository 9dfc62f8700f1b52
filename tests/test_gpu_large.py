"""Parity at the measured size and over a long seeded run.

* 512^3 decaying-turbulence workload (seed 1, E = 0.5, nu = 1e-3, dt from CFL 0.5): one fused RK4 step in
  FP64 equals the oracle's step (time_advance.cc:100-155) to <= 1e-12 per component and <= 1e-10 in E(k);
  the FP32-mixed step of the same state is within 1e-5 relative in E(k).
* 64^3 seeded random-phase field, FP64, RK4, dt = 1e-3, nu = 0.01, 100 steps, fused and plain paths:
  E(k) and Pi(R) over unit spheres <= 1e-10 relative of the oracle's run.
* Fused MHD (Elsasser) step in FP32-mixed at 32^3 and 128^3: <= 2e-5 of the FP64 oracle per component,
  the kinetic and magnetic E(k) <= 1e-5."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pyoracle import Oracle  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from tarang_amd import Context
    return Context(0)


def H(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def rel(a, b):
    m = np.abs(b).max()
    return np.abs(a - b).max() / (m if m > 0 else 1.0)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spec_err(Eg, Er):
    sel = Er >= 1e-10 * Er.max()
    return float((np.abs(Eg - Er)[sel] / Er[sel]).max())


@pytest.mark.timeout(900)
def test_512_fused_step_fp64_and_fp32_against_oracle(ctx):
    import torch
    from tarang_amd import Plan, State
    N = (512, 512, 512)
    nu = 1e-3
    p64 = Plan(ctx, N, precision="fp64")
    V = p64.ic_random(1, 0.5, 4.0)
    speed2 = None
    for a in V:
        b = a.clone()
        p64.inverse(b)
        r = torch.view_as_real(b).reshape(N[0], N[1], -1)[:, :, :N[2]]
        speed2 = r * r if speed2 is None else speed2 + r * r
        del b, r
    dt = 0.5 * (2 * math.pi / N[0]) / float(torch.sqrt(speed2.max()))
    del speed2
    log(f"512^3 step: dt = {dt:.6e}")
    Vh = [H(a) for a in V]
    s64 = State(p64)
    s64.upload(V)
    s64.step(dt, nu=nu, steps=1)
    g64 = s64.download()
    Eg64 = p64.vector_spectrum(g64)
    g64 = [H(a) for a in g64]
    s64.close()
    p32 = Plan(ctx, N, precision="fp32")
    s32 = State(p32)
    s32.upload([a.to(torch.complex64) for a in V])
    del V
    s32.step(dt, nu=nu, steps=1)
    Eg32 = p64.vector_spectrum([a.to(torch.complex128) for a in s32.download()])
    s32.close()
    torch.cuda.empty_cache()
    o = Oracle(N)
    o.rk4_ns(Vh, nu, dt, steps=1)
    Er = o.vector_spectrum(Vh)
    errs = [rel(g64[i], Vh[i]) for i in range(3)]
    e64, e32 = spec_err(Eg64, Er), spec_err(Eg32, Er)
    log(f"512^3 FP64 fused vs oracle: components {max(errs):.2e}, E(k) {e64:.2e}; FP32-mixed E(k) {e32:.2e}")
    assert max(errs) < 1e-12
    assert e64 <= 1e-10
    assert e32 <= 1e-5


@pytest.mark.timeout(600)
@pytest.mark.parametrize("path", ["fused", "plain"])
def test_random_field_100_steps_fp64_64(ctx, path):
    from tarang_amd import Plan, State
    N = (64, 64, 64)
    nu, dt, steps = 0.01, 1e-3, 100
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_random(7)
    g = [p.to_dev(v) for v in V]
    if path == "fused":
        s = State(p)
        s.upload(g)
        s.step(dt, nu=nu, steps=steps)
        g = s.download()
    else:
        p.step_plain(0, g, dt, nu=nu, steps=steps)
    o.rk4_ns(V, nu, dt, steps=steps)
    Er, Eg = o.vector_spectrum(V), p.vector_spectrum(g)
    err = spec_err(Eg, Er)
    radii = np.concatenate([np.arange(0, N[0] // 3 + 2, dtype=float), [10000.0]])
    fo = o.flux_ns(V, radii)
    fg = p.flux_ns(g, radii)
    sel = np.abs(fo) >= 1e-10 * np.abs(fo).max()
    ferr = float((np.abs(fg - fo)[sel] / np.abs(fo)[sel]).max())
    log(f"64^3 random 100 steps {path}: E(k) {err:.2e}, Pi {ferr:.2e}")
    assert err <= 1e-10
    assert ferr <= 1e-10


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n", [32, 128])
def test_fused_mhd_fp32_mixed(ctx, n):
    from tarang_amd import Plan, State, abi
    N = (n, n, n)
    nu, eta, dt, steps = 1e-3, 1e-3, 2e-3, 3
    o = Oracle(N)
    p = Plan(ctx, N, precision="fp32")
    s = State(p, abi.TH_MHD)
    U, B = o.ic_random(2), o.ic_random(3)
    s.upload([p.to_dev(a) for a in U + B])
    o.rk4_mhd(U, B, nu, eta, dt, steps=steps)
    s.step(dt, nu=nu, eta=eta, steps=steps)
    out = [H(a) for a in s.download()]
    errs = [rel(out[i], ref) for i, ref in enumerate(U + B)]
    p64 = Plan(ctx, N)
    Eu = spec_err(p64.vector_spectrum([p64.to_dev(a) for a in out[:3]]), o.vector_spectrum(U))
    Eb = spec_err(p64.vector_spectrum([p64.to_dev(a) for a in out[3:]]), o.vector_spectrum(B))
    log(f"MHD fp32-mixed {n}^3: components {max(errs):.2e}, E_u(k) {Eu:.2e}, E_b(k) {Eb:.2e}")
    assert max(errs) < 2e-5
    assert Eu <= 1e-5 and Eb <= 1e-5
