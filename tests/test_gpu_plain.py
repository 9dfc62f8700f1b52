"""GPU parity of the plain HIP path (reference operation sequence) against the FP64 oracle.

Tolerances: FP64 plans match the oracle to a few ulps of the largest value (transforms 1e-13,
nonlinear terms 1e-12 relative to max |.|); FP32-mixed plans to 2e-6 relative.  Binning and
mode counts are compared bit for bit."""
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


GRIDS = [(8, 8, 8), (16, 12, 24), (32, 32, 32), (64, 48, 16)]


@pytest.mark.parametrize("N", GRIDS)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_four_transforms(ctx, N, prec, tol):
    from tarang_amd import Plan
    log(f"four transforms N={N} {prec}")
    o = Oracle(N, dealias=False)
    p = Plan(ctx, N, precision=prec, dealias=False)
    rng = np.random.default_rng(1)
    A = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    ref = A.copy()
    o.inverse(ref)
    d = p.to_dev(A)
    p.inverse(d)
    assert rel(H(d), ref) < tol
    ref2 = ref.copy()
    o.forward(ref2)
    p.forward(d)
    assert rel(H(d), ref2) < tol


@pytest.mark.parametrize("N", [(8, 8, 8), (16, 12, 24), (32, 32, 16)])
@pytest.mark.parametrize("parity", [0, 1])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 3e-6)])
def test_scft_transforms(ctx, N, parity, prec, tol):
    from tarang_amd import Plan
    log(f"scft transforms N={N} parity={parity} {prec}")
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    o = Oracle(N, kfactor=kf, basis="SCFT", dealias=False)
    p = Plan(ctx, N, kfactor=kf, basis="SCFT", precision=prec, dealias=False)
    rng = np.random.default_rng(2)
    A = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    if parity:
        A[0] = 0
    ref = A.copy()
    o.inverse(ref, parity)
    d = p.to_dev(A)
    p.inverse(d, parity)
    assert rel(H(d), ref) < tol
    ref2 = ref.copy()
    o.forward(ref2, parity)
    p.forward(d, parity)
    assert rel(H(d), ref2) < tol


def test_golden_fixtures_on_gpu(ctx, golden_dir):
    """numpy rfftn golden vectors directly against the GPU transforms (FP64)."""
    from tarang_amd import Plan
    G = np.load(f"{golden_dir}/four_fft.npz")
    for tag in ("a", "c"):
        N = tuple(int(v) for v in G[f"{tag}_N"])
        p = Plan(ctx, N, dealias=False)
        A = np.zeros(p.shape, np.complex128)
        A.view(np.float64).reshape(N[0], N[1], -1)[:, :, :N[2]] = G[f"{tag}_u"]
        d = p.to_dev(A)
        p.forward(d)
        assert rel(H(d), G[f"{tag}_fwd"]) < 1e-13
        d = p.to_dev(G[f"{tag}_spec"])
        p.inverse(d)
        got = H(d).view(np.float64).reshape(N[0], N[1], -1)[:, :, :N[2]]
        assert np.abs(got - G[f"{tag}_inv"]).max() < 1e-12 * np.abs(G[f"{tag}_inv"]).max()


@pytest.mark.parametrize("basis", ["FOUR", "SCFT"])
def test_spectral_ops_bitwise(ctx, basis):
    """Derivatives, dealias, real products, 1/K^2 and exp(f K^2) (FP64) match the oracle to <= 1 ulp."""
    from tarang_amd import Plan
    N = (16, 24, 12)
    kf = (1.0, 0.5, 2.0)
    o = Oracle(N, kfactor=kf, basis=basis)
    p = Plan(ctx, N, kfactor=kf, basis=basis)
    rng = np.random.default_rng(3)
    A = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    B = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    dA, dB = p.to_dev(A), p.to_dev(B)
    for axis in (1, 2, 3):
        for par in (0, 1):
            assert rel(H(p.deriv(axis, dA, par)), o.deriv(axis, A, par)) <= 1e-16
    assert np.array_equal(H(p.dealias(dA.clone())), o.dealias(A.copy()))
    assert np.array_equal(H(p.real_mult(dA, dB)), o.real_mult(A, B))
    assert rel(H(p.divide_ksqr(dA.clone())), o.divide_ksqr(A.copy())) < 1e-15
    assert rel(H(p.exp_ksqr(dA.clone(), -0.01)), o.exp_ksqr(A.copy(), -0.01)) < 1e-15


@pytest.mark.parametrize("N", [(16, 16, 16), (24, 32, 16)])
def test_nlin_pressure_rhs_ns(ctx, N):
    from tarang_amd import Plan
    log(f"nlin ns N={N}")
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_random(11)
    nl = o.nlin_ns([v.copy() for v in V])
    dV = [p.to_dev(v) for v in V]
    dn = p.nlin_ns(dV)
    for i in range(3):
        assert rel(H(dn[i]), nl[i]) < 1e-12
    pr = o.pressure(nl)
    dp = p.pressure(dn)
    assert rel(H(dp), pr) < 1e-12
    o.rhs(nl, pr)
    p.rhs(dn, dp)
    for i in range(3):
        assert rel(H(dn[i]), nl[i]) < 1e-12


def test_nlin_mhd_rb_et(ctx):
    from tarang_amd import Plan
    N = (16, 16, 16)
    o = Oracle(N)
    p = Plan(ctx, N)
    U, B = o.ic_random(2), o.ic_random(3)
    dU, dB = [p.to_dev(a) for a in U], [p.to_dev(a) for a in B]
    nU, nB = o.nlin_mhd([a.copy() for a in U], [a.copy() for a in B])
    gU, gB = p.nlin_mhd(dU, dB)
    for i in range(3):
        assert rel(H(gU[i]), nU[i]) < 1e-12 and rel(H(gB[i]), nB[i]) < 1e-12
    G = [o.fill_shell(a, 2.0, 4.0) for a in U]
    gG = [p.fill_shell(a, 2.0, 4.0) for a in [p.to_dev(u) for u in U]]
    for i in range(3):
        assert np.array_equal(H(gG[i]), G[i])
    ne = o.et_nlin(U, G)
    ge = p.et_nlin([p.to_dev(a) for a in U], gG)
    for i in range(3):
        assert rel(H(ge[i]), ne[i]) < 1e-12
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    orb = Oracle(N, kfactor=kf, basis="SCFT")
    prb = Plan(ctx, N, kfactor=kf, basis="SCFT")
    rng = np.random.default_rng(4)
    V = [orb.dealias(rng.standard_normal(orb.shape) + 1j * rng.standard_normal(orb.shape)) for _ in range(3)]
    T = orb.ic_random_scalar(4, 1e-2)
    nV, nT = orb.nlin_rb(V, T)
    gV, gT = prb.nlin_rb([prb.to_dev(a) for a in V], prb.to_dev(T))
    for i in range(3):
        assert rel(H(gV[i]), nV[i]) < 1e-12
    assert rel(H(gT), nT) < 1e-12


def test_binning_counts_bitexact_and_spectra(ctx):
    """Shell index ceil(K) on the GPU equals the CPU's for every mode: the per-shell mode counts
    (Sk_count) agree exactly, including the RB kfactor (pi, pi/sqrt2, pi/sqrt2)."""
    from tarang_amd import Plan
    for basis, kf, N in [("FOUR", (1, 1, 1), (64, 64, 64)), ("FOUR", (1, 1, 1), (96, 48, 32)),
                         ("SCFT", (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2)), (32, 64, 64))]:
        o = Oracle(N, kfactor=kf, basis=basis)
        p = Plan(ctx, N, kfactor=kf, basis=basis)
        rng = np.random.default_rng(5)
        A = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
        Sk, cnt = o.shell_spectrum(A)
        gSk, gcnt = p.shell_spectrum(p.to_dev(A))
        assert np.array_equal(gcnt, cnt), basis
        assert np.abs(gSk - Sk).max() <= 1e-13 * np.abs(Sk).max()
        assert abs(p.energy(p.to_dev(A)) - o.energy(A)) < 1e-13 * abs(o.energy(A))
        assert abs(p.Sn(p.to_dev(A), 2) - o.Sn(A, 2)) < 1e-13 * abs(o.Sn(A, 2))


def test_ic_parity(ctx):
    from tarang_amd import Plan
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_random(1, 0.5, 4.0)
    gV = p.ic_random(1, 0.5, 4.0)
    for i in range(3):
        assert rel(H(gV[i]), V[i]) < 1e-13
    T = p.ic_taylor_green(1, 1.0)
    To = o.ic_taylor_green(1, 1.0)
    for i in range(3):
        assert np.array_equal(H(T[i]), To[i])
    assert abs(p.total_energy(T) - 0.125) < 1e-16


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_plain_steps_ns(ctx, scheme):
    from tarang_amd import Plan
    N = (16, 16, 16)
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_random(3)
    gV = [p.to_dev(a) for a in V]
    nu, dt = 0.01, 2e-3
    for _ in range(3):
        if scheme == 0:
            o.euler_ns(V, nu, dt)
        elif scheme == 1:
            o.rk2_ns(V, nu, dt)
        else:
            o.rk4_ns(V, nu, dt)
    p.step_plain(0, gV, dt, nu=nu, scheme=scheme, steps=3)
    for i in range(3):
        assert rel(H(gV[i]), V[i]) < 1e-13


def test_plain_steps_mhd_rb(ctx):
    from tarang_amd import Plan
    N = (16, 16, 16)
    o = Oracle(N)
    p = Plan(ctx, N)
    U, B = o.ic_random(2), o.ic_random(3)
    g = [p.to_dev(a) for a in U + B]
    o.rk4_mhd(U, B, 0.01, 0.02, 2e-3, steps=2)
    p.step_plain(1, g, 2e-3, nu=0.01, eta=0.02, steps=2)
    for i, ref in enumerate(U + B):
        assert rel(H(g[i]), ref) < 1e-13
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    orb = Oracle(N, kfactor=kf, basis="SCFT")
    prb = Plan(ctx, N, kfactor=kf, basis="SCFT")
    V = orb.zeros(3)
    T = orb.ic_random_scalar(4, 1e-3)
    g = [prb.to_dev(a) for a in V + [T]]
    orb.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=3)
    prb.step_plain(2, g, 1e-2, Ra=1e5, Pr=1.0, steps=3)
    for i, ref in enumerate(V + [T]):
        assert rel(H(g[i]), ref) < 1e-12
    assert abs(prb.nusselt(g[:3], g[3], 1e5, 1.0) - orb.nusselt(V, T, 1e5, 1.0)) < 1e-13


def test_flux_and_shell_transfer_plain(ctx):
    from tarang_amd import Plan
    N = (16, 16, 16)
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_random(5)
    radii = np.concatenate([np.arange(8), [10000.0]])
    gV = [p.to_dev(a) for a in V]
    f_ref = o.flux_ns(V, radii)
    f = p.flux_ns(gV, radii)
    assert np.abs(f - f_ref).max() < 1e-12 * np.abs(f_ref).max()
    T_ref = o.shell_tr_ns(V, radii)
    T = p.shell_tr_ns(gV, radii)
    assert np.abs(T - T_ref).max() < 1e-12 * np.abs(T_ref).max()


@pytest.mark.parametrize("cyl,axis", [(False, 3), (False, 1), (True, 3), (True, 2)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_ring_tr_plain(ctx, cyl, axis, prec, tol):
    """th_ring_tr (Compute_ring_tr compute_ring_tr.cc:60-291, Compute_cylinder_ring_tr
    compute_cylin_ring_tr.cc:54-133): U to U, the scalar T to T and the MHD U to W matrices (helper, giver,
    receiver and sign of each) match the oracle over shells x sectors (slabs)."""
    from tarang_amd import Plan
    N = (24, 24, 24)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    U, B = o.ic_random(2), o.ic_random(3)
    T = o.dealias(o.ic_random_scalar(4, 0.3))
    radii = np.array([0.0, 2.0, 4.0, 6.0, 10000.0])
    edges = np.array([0.0, 3.0, 5.0, 8.0]) if cyl else np.array([0.0, np.pi / 6, np.pi / 3, np.pi / 2])
    d = lambda X: [p.to_dev(a) for a in X]
    for H, F, R, sign in ((U, U, U, -1.0), (U, [T], [T], -1.0), (B, U, B, 1.0)):
        ref = o.ring_tr(H, F, R, radii, edges, sign=sign, cyl=cyl, axis=axis)
        got = p.ring_tr(d(H), d(F), d(R), radii, edges, sign=sign, cyl=cyl, axis=axis)
        sc = np.abs(ref).max()
        assert sc > 0
        assert np.abs(got - ref).max() <= tol * sc, (len(F), sign)


@pytest.mark.parametrize("N", [(16, 16, 16), (24, 16, 32)])
def test_flux_mhd_plain(ctx, N):
    """th_flux_mhd (IncVF::Compute_flux(IncVF& W), compute_flux.cc:148-262): the nine MHD and
    Elsasser fluxes match the oracle over unit spheres."""
    from tarang_amd import Plan
    o = Oracle(N)
    p = Plan(ctx, N)
    U, B = o.ic_random(2), o.ic_random(3)
    radii = np.concatenate([np.arange(7), [10000.0]])
    ref = o.flux_mhd(U, B, radii, elsasser=True)
    got = p.flux_mhd([p.to_dev(a) for a in U], [p.to_dev(a) for a in B], radii)
    names = ["u2u", "b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in", "u_out_b_out", "zp", "zm"]
    for r, k in enumerate(names):
        sc = max(np.abs(ref[k]).max(), 1e-300)
        assert np.abs(got[r] - ref[k]).max() < 1e-12 * sc, k


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_shell_tr_mhd_plain(ctx, prec, tol):
    """th_shell_tr_mhd (IncVF::Compute_shell_tr(IncVF& W), compute_shell_tr.cc:154-260): U to U, W to W,
    U to W and the Elsasser shell-to-shell matrices match the oracle over unit shells."""
    from tarang_amd import Plan
    N = (24, 24, 24)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    U, B = o.ic_random(2), o.ic_random(3)
    radii = np.concatenate([np.arange(9), [10000.0]])
    ref = o.shell_tr_mhd(U, B, radii)
    got = p.shell_tr_mhd([p.to_dev(a) for a in U], [p.to_dev(a) for a in B], radii)
    for k in ("uu", "ww", "uw", "zp", "zm"):
        sc = np.abs(ref[k]).max()
        assert sc > 0
        assert np.abs(got[k] - ref[k]).max() < tol * sc, k


@pytest.mark.parametrize("basis,prec,tol", [("SCFT", "fp64", 1e-12), ("FOUR", "fp64", 1e-12), ("SCFT", "fp32", 2e-5)])
def test_flux_and_shell_tr_scalar_plain(ctx, basis, prec, tol):
    """th_flux_scalar / th_shell_tr_scalar (Compute_flux_scalar compute_flux.cc:89-109, Compute_flux_RB
    :113-141, Compute_shell_tr_scalar compute_shell_tr.cc:98-117): velocity and scalar fluxes and
    shell-to-shell transfers match the oracle; RB fields on SCFT, a passive scalar on FOUR."""
    import math
    from tarang_amd import Plan
    if basis == "SCFT":
        N = (16, 16, 24)
        kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
        o = Oracle(N, kfactor=kf, basis="SCFT")
        V = o.zeros(3)
        T = o.dealias(o.ic_random_scalar(4, 0.5))
        o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=3)
        p = Plan(ctx, N, kfactor=kf, basis="SCFT", precision=prec)
    else:
        N = (16, 24, 16)
        o = Oracle(N)
        V = o.ic_random(5, E0=0.5, kp=2.0)
        T = o.dealias(o.ic_random_scalar(6, 0.5))
        p = Plan(ctx, N, precision=prec)
    log(f"scalar transfer {basis} {prec} N={N}")
    radii = np.concatenate([np.arange(7), [10000.0]]).astype(float)
    fs, ft = o.flux_scalar(V, T, radii)
    Ts, Tt = o.shell_tr_scalar(V, T, radii)
    dV, dT = [p.to_dev(a) for a in V], p.to_dev(T)
    gfs, gft = p.flux_scalar(dV, dT, radii)
    gTs, gTt = p.shell_tr_scalar(dV, dT, radii)
    for got, ref, name in ((gfs, fs, "flux_self"), (gft, ft, "flux_SF"), (gTs, Ts, "T_self"), (gTt, Tt, "T_SF")):
        sc = np.abs(ref).max()
        assert sc > 0, name
        assert np.abs(got - ref).max() < tol * sc, name


@pytest.mark.parametrize("basis", ["FOUR", "SCFT"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_helicity_and_entropy(ctx, basis, prec, tol):
    """th_helicity / th_shell_spectrum_helicity / th_entropy (four_energy.cc:453-750, scft_energy.cc:448-748)
    match the oracle."""
    from tarang_amd import Plan
    N = (16, 24, 16)
    o = Oracle(N, basis=basis)
    V = o.ic_random(12, E0=0.5, kp=3.0)
    p = Plan(ctx, N, basis=basis, precision=prec)
    dV = [p.to_dev(a) for a in V]
    h, ref = p.helicity(dV), o.helicity(V)
    assert np.abs(h - ref).max() <= tol * np.abs(ref).max()
    Hk, rHk = p.shell_helicity(dV), o.shell_helicity(V)
    assert np.abs(Hk - rHk).max() <= tol * np.abs(rHk).max()
    assert abs(p.entropy(dV) - o.entropy(V)) <= tol * o.entropy(V)
    assert abs(p.entropy_scalar(dV[1]) - o.entropy_scalar(V[1])) <= tol * o.entropy_scalar(V[1])


def test_config1_taylor_green_64_100_steps(ctx):
    """Config 1: TG 64^3 FP64, RK4, dt = 1e-3, 100 steps, nu = 0.01, 2/3 dealias.  The plain GPU path
    matches the oracle to <= 1e-10 relative in E(k) (shells with E >= 1e-10 max E) and in Pi(R)."""
    from tarang_amd import Plan
    N = (64, 64, 64)
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_taylor_green(1, 1.0)
    g = p.ic_taylor_green(1, 1.0)
    log("config1: oracle 100 steps")
    o.rk4_ns(V, 0.01, 1e-3, steps=100)
    log("config1: gpu 100 steps")
    p.step_plain(0, g, 1e-3, nu=0.01, steps=100)
    Ek = o.vector_spectrum(V)
    gEk = p.vector_spectrum(g)
    sel = Ek >= 1e-10 * Ek.max()
    err = np.abs(gEk[sel] - Ek[sel]) / Ek[sel]
    log(f"config1 max rel E(k) err {err.max():.3e} over {sel.sum()} shells")
    assert err.max() <= 1e-10
    radii = np.concatenate([np.arange(1, 22, dtype=float), [10000.0]])
    radii = np.concatenate([[0.0], radii])
    fo = o.flux_ns(V, radii)
    fg = p.flux_ns(g, radii)
    sel = np.abs(fo) >= 1e-10 * np.abs(fo).max()
    ferr = (np.abs(fg - fo)[sel] / np.abs(fo)[sel]).max()
    log(f"config1 max rel flux err {ferr:.3e}")
    assert ferr <= 1e-10
