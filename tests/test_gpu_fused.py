"""GPU parity of the fused, dealias-pruned path (th_state / th_step / th_flux_shell_tr).

The fused step assumes a 2/3-dealiased state, which the reference keeps after every step
(time_advance.cc:162) and after its ICs.  FP64 runs match the oracle to <= 1e-12 relative per
few steps and <= 1e-10 in E(k) and Pi(k) after 100 steps (config 1); FP32-mixed runs match the
FP64 oracle to <= 1e-5 relative in E(k)."""
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


@pytest.mark.parametrize("N", [(16, 16, 16), (48, 32, 24)])
def test_upload_download_roundtrip(ctx, N):
    from tarang_amd import Plan, State
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p)
    rng = np.random.default_rng(0)
    V = [rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape) for _ in range(3)]
    s.upload([p.to_dev(v) for v in V])
    out = s.download()
    for i in range(3):
        assert np.array_equal(H(out[i]), o.dealias(V[i].copy()))


@pytest.mark.parametrize("N", [(16, 16, 16), (32, 32, 32), (48, 32, 24), (64, 64, 64), (64, 64, 128)])
def test_fused_rk4_ns_fp64(ctx, N):
    from tarang_amd import Plan, State
    log(f"fused ns fp64 N={N}")
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p)
    V = o.ic_random(21)
    s.upload([p.to_dev(v) for v in V])
    nu, dt = 0.01, 2e-3
    o.rk4_ns(V, nu, dt, steps=3)
    s.step(dt, nu=nu, steps=3)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V[i]) < 1e-12


def test_fused_rk4_ns_fp32_mixed(ctx):
    from tarang_amd import Plan, State
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N, precision="fp32")
    s = State(p)
    V = o.ic_random(22)
    s.upload([p.to_dev(v) for v in V])
    o.rk4_ns(V, 0.01, 2e-3, steps=5)
    s.step(2e-3, nu=0.01, steps=5)
    out = s.download()
    Ek = o.vector_spectrum(V)
    p64 = Plan(ctx, N)
    gEk = p64.vector_spectrum([p64.to_dev(H(a)) for a in out])
    sel = Ek >= 1e-10 * Ek.max()
    err = (np.abs(gEk - Ek)[sel] / Ek[sel]).max()
    log(f"fp32-mixed E(k) max rel err {err:.2e}")
    assert err < 1e-5


def test_fused_mhd_fp64(ctx):
    from tarang_amd import Plan, State, abi
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p, abi.TH_MHD)
    U, B = o.ic_random(2), o.ic_random(3)
    s.upload([p.to_dev(a) for a in U + B])
    o.rk4_mhd(U, B, 0.01, 0.02, 2e-3, steps=3)
    s.step(2e-3, nu=0.01, eta=0.02, steps=3)
    out = s.download()
    for i, ref in enumerate(U + B):
        assert rel(H(out[i]), ref) < 1e-12


@pytest.mark.parametrize("N", [(48, 32, 24), (32, 32, 48), (512, 16, 16)])
def test_fused_mhd_fp64_spilling_shapes(ctx, N):
    """The FP64 MHD kernels that spill registers to scratch (z pass at N3 = 3 * 2^k, x-forward at N1 = 512)
    still equal the oracle's Time_advance(IncVF&) steps."""
    from tarang_amd import Plan, State, abi
    log(f"fused mhd fp64 N={N}")
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p, abi.TH_MHD)
    U, B = o.ic_random(4), o.ic_random(6)
    s.upload([p.to_dev(a) for a in U + B])
    o.rk4_mhd(U, B, 0.01, 0.02, 2e-3, steps=2)
    s.step(2e-3, nu=0.01, eta=0.02, steps=2)
    out = s.download()
    for i, ref in enumerate(U + B):
        assert rel(H(out[i]), ref) < 1e-12, i


@pytest.mark.parametrize("N", [(48, 32, 24), (32, 32, 48)])
def test_fused_flux_mhd_fp64_spilling_shapes(ctx, N):
    """th_flux_mhd_fused through the energy-transfer z pass (MODE 2) at N3 = 3 * 2^k, where its FP64 build
    spills, equals the oracle's sphere loop."""
    from tarang_amd import Plan, State, abi
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p, abi.TH_MHD)
    U, B = o.ic_random(2, kp=3.0), o.ic_random(3, kp=3.0)
    s.upload([p.to_dev(a) for a in U + B])
    radii = np.concatenate([np.arange(10, dtype=float), [10000.0]])
    got = s.flux_mhd(radii)
    ref = o.flux_mhd(U, B, radii, elsasser=True)
    for r, k in enumerate(["u2u", "b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in", "u_out_b_out", "zp", "zm"]):
        sc = np.abs(ref[k]).max()
        assert np.abs(got[r] - ref[k]).max() <= 1e-11 * sc, k


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-5)])
def test_fused_flux_and_shell_transfer(ctx, prec, tol):
    """Pi(R) from shell givers + corner bins equals the reference's sphere loop (Compute_flux) and
    T(n,m) equals Compute_shell_tr."""
    from tarang_amd import Plan, State
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    s = State(p)
    V = o.ic_random(5, kp=3.0)
    s.upload([p.to_dev(v) for v in V])
    radii = np.concatenate([np.arange(12, dtype=float), [10000.0]])
    Pi, T = s.flux_shell_tr(radii)
    f_ref = o.flux_ns(V, radii)
    T_ref = o.shell_tr_ns(V, radii)
    scale_f = np.abs(f_ref).max()
    assert np.abs(Pi - f_ref).max() < tol * scale_f
    assert np.abs(T - T_ref).max() < tol * np.abs(T_ref).max()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-5)])
def test_fused_flux_mhd(ctx, prec, tol):
    """th_flux_mhd_fused: the nine MHD fluxes of Compute_flux(IncVF& W) (compute_flux.cc:148-262) from giver
    shells through the fused pipeline equal the reference's sphere loop (the oracle), every row including
    the in-sphere and outside-giver ones and the Elsasser pair."""
    from tarang_amd import Plan, State, abi
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    s = State(p, abi.TH_MHD)
    U, B = o.ic_random(2, kp=3.0), o.ic_random(3, kp=3.0)
    s.upload([p.to_dev(a) for a in U + B])
    radii = np.concatenate([np.arange(12, dtype=float), [10000.0]])
    got = s.flux_mhd(radii)
    ref = o.flux_mhd(U, B, radii, elsasser=True)
    names = ["u2u", "b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in", "u_out_b_out", "zp", "zm"]
    for r, k in enumerate(names):
        sc = np.abs(ref[k]).max()
        assert sc > 0, k
        assert np.abs(got[r] - ref[k]).max() <= tol * sc, (k, np.abs(got[r] - ref[k]).max() / sc)


def test_config1_fused_taylor_green_64(ctx):
    """Config 1 through the fused path: TG 64^3 FP64, RK4, dt = 1e-3, 100 steps, nu = 0.01."""
    from tarang_amd import Plan, State
    N = (64, 64, 64)
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p)
    V = o.ic_taylor_green(1, 1.0)
    s.upload([p.to_dev(v) for v in V])
    o.rk4_ns(V, 0.01, 1e-3, steps=100)
    s.step(1e-3, nu=0.01, steps=100)
    out = s.download()
    Ek = o.vector_spectrum(V)
    gEk = p.vector_spectrum(out)
    sel = Ek >= 1e-10 * Ek.max()
    err = (np.abs(gEk - Ek)[sel] / Ek[sel]).max()
    log(f"config1 fused max rel E(k) err {err:.3e}")
    assert err <= 1e-10
    radii = np.concatenate([np.arange(0, 22, dtype=float), [10000.0]])
    s2 = State(p)
    s2.upload(out)
    Pi, _ = s2.flux_shell_tr(radii)
    fo = o.flux_ns(V, radii)
    sel = np.abs(fo) >= 1e-10 * np.abs(fo).max()
    ferr = (np.abs(Pi - fo)[sel] / np.abs(fo)[sel]).max()
    log(f"config1 fused max rel flux err {ferr:.3e}")
    assert ferr <= 1e-10


def _rb_fields(o, amp=0.1, seed=4, warm=3):
    """Non-trivial dealiased RB fields: the oracle's RK4 run from V = 0 and a random temperature."""
    V = o.zeros(3)
    T = o.dealias(o.ic_random_scalar(seed, amp))
    o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=warm)
    return V, T


@pytest.mark.parametrize("N", [(32, 32, 32), (48, 32, 24), (64, 48, 32)])
def test_fused_rb_scft_fp64(ctx, N):
    """Fused SCFT/RB step (DCT/DST x passes, RB products, force, projection, IF-RK4 for V and T)
    equals the oracle's Time_advance(IncSF&) step."""
    from tarang_amd import Plan, State, abi
    import math
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    log(f"fused rb fp64 N={N}")
    o = Oracle(N, kfactor=kf, basis="SCFT")
    V, T = _rb_fields(o)
    p = Plan(ctx, N, kfactor=kf, basis="SCFT")
    s = State(p, abi.TH_RB)
    s.upload([p.to_dev(a) for a in V + [T]])
    back = s.download()
    for i, ref in enumerate(V + [T]):
        assert np.array_equal(H(back[i]), ref), i
    o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=3)
    s.step(1e-2, Ra=1e5, Pr=1.0, steps=3)
    out = s.download()
    for i, ref in enumerate(V + [T]):
        assert rel(H(out[i]), ref) < 1e-12, i
    assert abs(p.nusselt(out[:3], out[3], 1e5, 1.0) - o.nusselt(V, T, 1e5, 1.0)) < 1e-12


def test_fused_rb_scft_fp32_mixed(ctx):
    from tarang_amd import Plan, State, abi
    import math
    N = (32, 32, 32)
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    o = Oracle(N, kfactor=kf, basis="SCFT")
    V, T = _rb_fields(o)
    p = Plan(ctx, N, kfactor=kf, basis="SCFT", precision="fp32")
    s = State(p, abi.TH_RB)
    s.upload([p.to_dev(a) for a in V + [T]])
    o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=3)
    s.step(1e-2, Ra=1e5, Pr=1.0, steps=3)
    out = s.download()
    for i, ref in enumerate(V + [T]):
        assert rel(H(out[i]), ref) < 2e-5, i


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_chained_step_after_other_uses_of_the_state(ctx, prec, tol):
    """The NS step chains each stage's x-inverse into the previous x-forward, so its first stage needs the
    x-inverse of S in the pipeline buffer.  Every operation that changes S or uses those buffers in between
    (an upload, a flux evaluation, per-pass timing, tuning, forcing switched on and off) must leave the next
    step equal to the oracle's."""
    from tarang_amd import Plan, State
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    s = State(p)
    V = o.ic_random(13)
    s.upload([p.to_dev(v) for v in V])
    nu, dt = 0.01, 2e-3
    radii = np.concatenate([np.arange(8, dtype=float), [10000.0]])

    def step_and_check(what):
        s.step(dt, nu=nu, steps=2)
        o.rk4_ns(V, nu, dt, steps=2)
        out = s.download()
        for i in range(3):
            assert rel(H(out[i]), V[i]) < tol, (what, i)

    step_and_check("first")
    W = o.ic_random(14)                      # a new state uploaded between steps
    V[:] = [w.copy() for w in W]
    s.upload([p.to_dev(v) for v in V])
    step_and_check("upload")
    s.flux_shell_tr(radii)
    step_and_check("flux")
    for stage in (1, 2, 3, 4):
        s.profile_passes(dt, nu=nu, stage=stage, nrep=1)
    step_and_check("profile")
    s.tune(dt, nu=nu, nrep=1)
    step_and_check("tune")
    s.set_force(1, 1.5, 2.5, 0.0, 0.0)       # energy supply 0: the forced step is the free step
    step_and_check("forced, zero supply")
    s.set_force(0, 0, 0, 0, 0)
    step_and_check("forcing off")


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("N", [(32, 32, 32), (64, 16, 16)])
def test_chained_mhd_step_after_other_uses_of_the_state(ctx, prec, tol, N):
    """The MHD step chains the next stage's Elsasser x-inverse (Z+ = U + B, Z- = U - B) into each x-forward;
    an upload, the fused MHD fluxes, per-pass timing and tuning between steps must leave the next step equal
    to the oracle's Time_advance(IncVF&)."""
    from tarang_amd import Plan, State, abi
    log(f"chained mhd {prec} N={N}")
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    s = State(p, abi.TH_MHD)
    U, B = o.ic_random(21), o.ic_random(22)
    s.upload([p.to_dev(a) for a in U + B])
    nu, eta, dt = 0.01, 0.02, 2e-3
    radii = np.concatenate([np.arange(8, dtype=float), [10000.0]])

    def step_and_check(what):
        s.step(dt, nu=nu, eta=eta, steps=2)
        o.rk4_mhd(U, B, nu, eta, dt, steps=2)
        out = s.download()
        for i, ref in enumerate(U + B):
            assert rel(H(out[i]), ref) < tol, (what, i)

    step_and_check("first")
    U2, B2 = o.ic_random(23), o.ic_random(24)
    U[:] = [u.copy() for u in U2]
    B[:] = [b.copy() for b in B2]
    s.upload([p.to_dev(a) for a in U + B])
    step_and_check("upload")
    s.flux_mhd(radii)
    step_and_check("flux")
    for stage in (1, 2, 3, 4):
        s.profile_passes(dt, nu=nu, eta=eta, stage=stage, nrep=1)
    step_and_check("profile")
    s.tune(dt, nu=nu, eta=eta, nrep=1)
    step_and_check("tune")


def test_plan_destroyed_before_its_state(ctx):
    """th_plan_destroy on a plan with a live state defers the release to th_state_destroy (the order
    a garbage collector may pick for a reference cycle), and the state keeps working meanwhile."""
    from tarang_amd import Plan, State, abi
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(5)
    p = Plan(ctx, N)
    s = State(p)
    s.upload([p.to_dev(a) for a in V])
    lib = abi.lib()
    assert lib.th_plan_destroy(p.h) == 0          # deferred: the state is alive
    assert lib.th_plan_destroy(p.h) == 0          # a second call is a no-op
    s.step(1e-3, nu=1e-2, steps=1)
    o.rk4_ns(V, 1e-2, 1e-3, steps=1)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V[i]) < 1e-12, i
    assert lib.th_state_destroy(s.h) == 0         # releases the plan too
    s.h = None
    p.h = None


def test_profile_passes_leaves_the_state(ctx):
    """th_profile_passes replays single RK stages for timing; the state itself (which stage 4 writes)
    is restored, so what follows a measurement sees the state the timed steps left."""
    from tarang_amd import Plan, State
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p)
    s.upload([p.to_dev(a) for a in o.ic_random(9)])
    before = [H(a) for a in s.download()]
    for stage in (1, 2, 3, 4):
        s.profile_passes(1e-3, nu=1e-2, stage=stage, nrep=2)
    after = [H(a) for a in s.download()]
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
