"""GPU parity of the energy/helicity forcing (force_field_proc 1 and 2, th_compute_force and
th_step_plain_forced) against the FP64 oracle.

Tolerances: FP64 plans match the oracle to 1e-13 of max |F| (force field) and 1e-12 of max |V| (forced
steps); FP32-mixed plans (FP32 storage, FP64 force arithmetic) to 1e-5.  The mode count nf is compared
exactly.  Reference: compute_force_ek_hk_supply.cc:51-157, compute_force_const_ek_hk.cc:51-140,
force_alpha_beta.cc:60-89, single_time_step.cc:351-364."""
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


# (grid, kfactor, inner, outer): band inside N/3; band past N/3 (Dealias_force); band wider than the
# grid in x (every x index in the box); anisotropic kfactor
CASES = [((16, 16, 16), (1.0, 1.0, 1.0), 1.5, 3.5),
         ((16, 16, 16), (1.0, 1.0, 1.0), 3.0, 6.5),
         ((8, 16, 16), (1.0, 1.0, 1.0), 1.0, 4.5),
         ((24, 16, 32), (1.0, 0.5, 2.0), 1.2, 4.1)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("proc,p3,p4", [(1, 0.4, 0.0), (1, 0.4, 0.6), (2, 0.8, 0.3), (2, 0.8, -0.5)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
def test_compute_force_matches_oracle(ctx, case, proc, p3, p4, prec, tol):
    from tarang_amd import Plan
    N, kf, inner, outer = case
    log(f"compute_force N={N} kf={kf} band=({inner},{outer}] proc={proc} p4={p4} {prec}")
    o = Oracle(N, kfactor=kf)
    V = o.ic_random(21, E0=0.5, kp=3.0)
    p = Plan(ctx, N, kfactor=kf, precision=prec)
    dV = [p.to_dev(a) for a in V]
    if prec == "fp32":                        # the oracle sees the same rounded inputs
        V = [H(d) for d in dV]
    ref = o.force(V, proc, inner, outer, p3, p4)
    F, nf = p.compute_force(dV, proc, inner, outer, p3, p4)
    assert nf == o.force_supply_modes(inner, outer)
    m = max(np.abs(r).max() for r in ref)
    assert m > 0
    for c in range(3):
        assert np.abs(H(F[c]) - ref[c]).max() <= tol * m


def test_compute_force_beltrami_modes(ctx):
    """Maximally helical modes take the degenerate branches on the GPU as in the oracle."""
    from tarang_amd import Plan
    log("compute_force beltrami")
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.zeros(3)
    A = 0.6
    V[0][0, 0, 3], V[1][0, 0, 3] = A, 1j * A
    V[1][3, 0, 0], V[2][3, 0, 0] = A, -1j * A
    V[1][-3, 0, 0], V[2][-3, 0, 0] = A, 1j * A
    p = Plan(ctx, N)
    dV = [p.to_dev(a) for a in V]
    for proc, p4 in [(1, 0.0), (2, 0.3)]:
        ref = o.force(V, proc, 2.5, 3.5, 0.2, p4)
        F, _ = p.compute_force(dV, proc, 2.5, 3.5, 0.2, p4)
        for c in range(3):
            assert np.abs(H(F[c]) - ref[c]).max() <= 1e-15


@pytest.mark.parametrize("proc,p3,p4", [(1, 0.5, 0.0), (1, 0.5, 0.4), (2, 0.6, 0.2)])
@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_forced_steps_match_oracle(ctx, proc, p3, p4, scheme):
    from tarang_amd import Plan
    log(f"forced steps proc={proc} p4={p4} scheme={scheme}")
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(3, E0=0.5, kp=3.0)
    p = Plan(ctx, N)
    gV = [p.to_dev(a) for a in V]
    nu, dt, inner, outer = 0.01, 1e-3, 1.5, 3.5
    o.step_ns_forced(V, nu, dt, proc, inner, outer, p3, p4, scheme=scheme, steps=3)
    p.step_plain_forced(gV, dt, nu, proc, inner, outer, p3, p4, scheme=scheme, steps=3)
    for i in range(3):
        assert rel(H(gV[i]), V[i]) < 1e-12


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_forced_step_reports_its_last_stage(ctx, scheme):
    """th_step_plain_forced's last_stage is the field of the step's last stage (the one whose force the
    reference's Force1..3 hold after Time_advance), equal to the oracle's."""
    from tarang_amd import Plan
    log(f"forced last stage scheme={scheme}")
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(6, E0=0.5, kp=3.0)
    p = Plan(ctx, N)
    gV = [p.to_dev(a) for a in V]
    last, glast = o.zeros(3), p.empty(3)
    o.step_ns_forced(V, 0.01, 1e-3, 1, 1.5, 3.5, 0.5, 0.2, scheme=scheme, steps=2, last_stage=last)
    p.step_plain_forced(gV, 1e-3, 0.01, 1, 1.5, 3.5, 0.5, 0.2, scheme=scheme, steps=2, last_stage=glast)
    for c in range(3):
        assert rel(H(glast[c]), last[c]) < 1e-12
        assert rel(H(gV[c]), V[c]) < 1e-12


def test_forced_rk4_fp32(ctx):
    from tarang_amd import Plan
    log("forced rk4 fp32")
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(5, E0=0.5, kp=3.0)
    p = Plan(ctx, N, precision="fp32")
    gV = [p.to_dev(a) for a in V]
    V = [H(d) for d in gV]
    o.step_ns_forced(V, 0.01, 1e-3, 1, 1.5, 3.5, 0.5, 0.0)
    p.step_plain_forced(gV, 1e-3, 0.01, 1, 1.5, 3.5, 0.5, 0.0)
    for i in range(3):
        assert rel(H(gV[i]), V[i]) < 1e-5


def test_force_argument_errors(ctx):
    from tarang_amd import Plan
    log("force errors")
    p = Plan(ctx, (16, 16, 16))
    dV = p.empty(3)
    with pytest.raises(RuntimeError, match="not supported"):
        p.compute_force(dV, 5, 1.5, 3.5, 1.0)
    with pytest.raises(RuntimeError, match="no wavevector"):
        p.compute_force(dV, 1, 0.1, 0.5, 1.0)
    with pytest.raises(RuntimeError, match="h_by_k_E"):
        p.compute_force(dV, 2, 1.5, 3.5, 1.0, 1.5)
    with pytest.raises(RuntimeError, match="FOUR"):
        Plan(ctx, (16, 16, 16), basis="SCFT").compute_force(dV, 1, 1.5, 3.5, 1.0)


@pytest.mark.parametrize("N,kf,inner,outer", [((16, 16, 16), (1.0, 1.0, 1.0), 1.5, 3.5),
                                              ((32, 32, 32), (1.0, 1.0, 1.0), 0.5, 2.5),
                                              ((48, 32, 24), (1.0, 1.0, 1.0), 2.0, 4.0),
                                              ((16, 16, 16), (1.0, 1.0, 1.0), 3.0, 6.5),
                                              ((64, 32, 32), (1.0, 1.0, 1.0), 1.5, 3.5),
                                              ((64, 16, 32), (1.0, 2.0, 1.0), 0.5, 4.5)])
@pytest.mark.parametrize("proc,p3,p4", [(1, 0.5, 0.3), (2, 0.6, 0.2)])
def test_fused_forced_step_matches_oracle(ctx, N, kf, inner, outer, proc, p3, p4):
    """th_state_set_force + th_step: the band kernels around each fused stage give the oracle's forced
    RK4 step (force of each stage field, projected, into the rhs).  The fourth band reaches past N/3
    (Dealias_force); N1 = 48 and 64 step chained (the force's share of each stage output is added to the
    next stage's x-transformed lines), 16 and 32 unchained."""
    from tarang_amd import Plan, State
    log(f"fused forced N={N} band=({inner},{outer}] proc={proc}")
    o = Oracle(N, kfactor=kf)
    p = Plan(ctx, N, kfactor=kf)
    s = State(p)
    V = o.ic_random(8, E0=0.5, kp=3.0)
    s.upload([p.to_dev(v) for v in V])
    s.set_force(proc, inner, outer, p3, p4)
    nu, dt = 0.01, 1e-3
    o.step_ns_forced(V, nu, dt, proc, inner, outer, p3, p4, steps=3)
    s.step(dt, nu=nu, steps=3)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V[i]) < 1e-12


@pytest.mark.parametrize("N", [(32, 32, 32), (64, 32, 32)])
def test_fused_forced_fp32_mixed_and_switch_off(ctx, N):
    """FP32-mixed fused forced step vs the oracle from the same rounded start (<= 1e-5); with the
    force switched off again the state steps exactly as a never-forced state; energy_supply = 0
    leaves the step bit-identical to the unforced one (unchained at N1 = 32, chained at 64)."""
    from tarang_amd import Plan, State
    log(f"fused forced fp32 + off N={N}")
    o = Oracle(N)
    p = Plan(ctx, N, precision="fp32")
    V = o.ic_random(9, E0=0.5, kp=3.0)
    d = [p.to_dev(v) for v in V]
    s = State(p)
    s.upload(d)
    V = [H(x) for x in d]
    s.set_force(1, 1.5, 3.5, 0.5, 0.0)
    o.step_ns_forced(V, 0.01, 1e-3, 1, 1.5, 3.5, 0.5, 0.0, steps=2)
    s.step(1e-3, nu=0.01, steps=2)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V[i]) < 1e-5
    a, b = State(p), State(p)
    a.upload(d)
    b.upload(d)
    a.set_force(1, 1.5, 3.5, 0.5, 0.0)
    a.set_force(0)
    a.step(1e-3, nu=0.01, steps=2)
    b.step(1e-3, nu=0.01, steps=2)
    c = State(p)
    c.upload(d)
    c.set_force(1, 1.5, 3.5, 0.0, 0.0)
    c.step(1e-3, nu=0.01, steps=2)
    oa, ob, oc = a.download(), b.download(), c.download()
    for i in range(3):
        assert np.array_equal(H(oa[i]), H(ob[i]))
        assert np.array_equal(H(oc[i]), H(ob[i]))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_chained_forced_state_keeps_last_stage_field(ctx, prec, tol):
    """A chained forced step keeps the last RK stage's input field (th_state_download_buffer 1: the field the
    reference's Force arrays come from after Time_advance) equal to the oracle's; after a chained step of an
    unforced state the buffer is refused rather than stale."""
    from tarang_amd import Plan, State
    log(f"chained forced stage field {prec}")
    N = (64, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    V = o.ic_random(8, E0=0.5, kp=3.0)
    d = [p.to_dev(v) for v in V]
    V = [H(x) for x in d]
    s = State(p)
    s.upload(d)
    s.set_force(1, 1.5, 3.5, 0.5, 0.3)
    last = o.zeros(3)
    o.step_ns_forced(V, 0.01, 1e-3, 1, 1.5, 3.5, 0.5, 0.3, steps=2, last_stage=last)
    s.step(1e-3, nu=0.01, steps=2)
    got = s.download_buffer(1)
    for i in range(3):
        assert rel(H(got[i]), last[i]) < tol, i
    u = State(p)
    u.upload(d)
    u.step(1e-3, nu=0.01, steps=1)
    with pytest.raises(RuntimeError, match="keeps no stage field"):
        u.download_buffer(1)


def test_fused_force_band_outside_the_stored_modes(ctx):
    """A band whose modes the dealiased state does not store (|kx| = N/3 without Dealias_force) is refused
    by the fused state; the plain stepper runs it."""
    from tarang_amd import Plan, State
    log("fused force refused band")
    p = Plan(ctx, (16, 16, 16))
    s = State(p)
    with pytest.raises(RuntimeError, match="does not store"):
        s.set_force(1, 4.0, 5.0, 0.5, 0.0)
    with pytest.raises(RuntimeError, match="single-box NS"):
        from tarang_amd import abi
        State(p, abi.TH_MHD).set_force(1, 1.5, 3.5, 0.5, 0.0)


@pytest.mark.parametrize("p3,p4,p5,p6", [(0.5, 0.3, 0.4, -0.2), (0.5, 0.0, 0.0, 0.0), (0.0, 0.0, 0.3, 0.5)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_forced_mhd_steps_match_oracle(ctx, p3, p4, p5, p6, prec, tol):
    """th_step_plain_forced_mhd: the supply of U (p3, p4) and of W (p5, p6) recomputed at every RK4 stage
    matches the oracle's forced MHD step; a zero supply leaves its field unforced."""
    from tarang_amd import Plan
    log(f"forced mhd p3={p3} p4={p4} p5={p5} p6={p6} {prec}")
    N = (16, 16, 16)
    o = Oracle(N)
    U, B = o.ic_random(3, E0=0.5, kp=3.0), o.ic_random(4, E0=0.3, kp=3.0)
    p = Plan(ctx, N, precision=prec)
    gU, gB = [p.to_dev(a) for a in U], [p.to_dev(a) for a in B]
    if prec == "fp32":
        U, B = [H(d) for d in gU], [H(d) for d in gB]
    nu, eta, dt = 0.01, 0.02, 1e-3
    o.rk4_mhd_forced(U, B, nu, eta, dt, 1.5, 3.5, p3, p4, p5, p6, steps=3)
    p.step_plain_forced_mhd(gU, gB, dt, nu, eta, 1.5, 3.5, p3, p4, p5, p6, steps=3)
    for i in range(3):
        assert rel(H(gU[i]), U[i]) < tol
        assert rel(H(gB[i]), B[i]) < tol


@pytest.mark.parametrize("proc,p6", [(1, -0.2), (2, -0.5)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-5)])
def test_compute_force_mhd_matches_oracle(ctx, proc, p6, prec, tol):
    """th_compute_force_mhd vs the oracle's Compute_force(IncVF& W), with a band mode of energy 1e-17 in both
    fields (forced in B under proc 2 only)."""
    from tarang_amd import Plan
    log(f"compute_force_mhd proc={proc} {prec}")
    N = (16, 16, 16)
    o = Oracle(N)
    U, B = o.ic_random(3, E0=0.5, kp=3.0), o.ic_random(4, E0=0.3, kp=3.0)
    a = np.sqrt(2e-17)
    for V in (U, B):
        V[0][0, 3, 0] = V[0][0, -3, 0] = a
        V[1][0, 3, 0] = V[1][0, -3, 0] = 0.0
        V[2][0, 3, 0] = V[2][0, -3, 0] = 0.0
    p = Plan(ctx, N, precision=prec)
    dU, dB = [p.to_dev(x) for x in U], [p.to_dev(x) for x in B]
    if prec == "fp32":
        U, B = [H(d) for d in dU], [H(d) for d in dB]
    RU, RW = o.force_mhd(U, B, proc, 1.5, 3.5, 0.5, 0.3, 0.4, p6)
    FU, FW, nf = p.compute_force_mhd(dU, dB, proc, 1.5, 3.5, 0.5, 0.3, 0.4, p6)
    assert nf == o.force_supply_modes(1.5, 3.5)
    m = max(np.abs(r).max() for r in RU + RW)
    for c in range(3):
        assert np.abs(H(FU[c]) - RU[c]).max() <= tol * m
        assert np.abs(H(FW[c]) - RW[c]).max() <= tol * m
    if proc == 2 and prec == "fp64":
        assert np.abs(RW[0][0, 3, 0]) > 0 and RU[0][0, 3, 0] == 0


@pytest.mark.parametrize("p6", [-0.5, 0.0])
def test_forced_mhd_constant_level_steps_match_oracle(ctx, p6):
    from tarang_amd import Plan
    log(f"forced mhd proc 2 p6={p6}")
    N = (16, 16, 16)
    o = Oracle(N)
    U, B = o.ic_random(3, E0=0.5, kp=3.0), o.ic_random(4, E0=0.3, kp=3.0)
    p = Plan(ctx, N)
    gU, gB = [p.to_dev(a) for a in U], [p.to_dev(a) for a in B]
    o.rk4_mhd_forced(U, B, 0.01, 0.02, 1e-3, 1.5, 3.5, 0.6, 0.2, 0.4, p6, steps=3, proc=2)
    p.step_plain_forced_mhd(gU, gB, 1e-3, 0.01, 0.02, 1.5, 3.5, 0.6, 0.2, 0.4, p6, steps=3, proc=2)
    for i in range(3):
        assert rel(H(gU[i]), U[i]) < 1e-12
        assert rel(H(gB[i]), B[i]) < 1e-12


def test_forced_mhd_argument_errors(ctx):
    from tarang_amd import Plan
    log("forced mhd errors")
    p = Plan(ctx, (16, 16, 16))
    U, B = p.empty(3), p.empty(3)
    with pytest.raises(RuntimeError, match="not supported"):
        p.step_plain_forced_mhd(U, B, 1e-3, 0.01, 0.02, 1.5, 3.5, 0.5, 0.2, 0.3, 0.1, proc=5)
    with pytest.raises(RuntimeError, match="must be finite"):
        p.step_plain_forced_mhd(U, B, 1e-3, 0.01, 0.02, 1.5, 3.5, 0.5, 0.0, float("nan"), 0.0)
    with pytest.raises(RuntimeError, match="h_by_k_E_W"):
        p.compute_force_mhd(U, B, 2, 1.5, 3.5, 0.5, 0.0, 0.3, 1.5)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-15), ("fp32", 1e-7)])
def test_taylor_green_force_and_steps_match_oracle(ctx, prec, tol):
    """force_field_proc 3 (compute_force_TG.cc:50-104): the fixed four-mode field, the forced EULER / RK2 / RK4
    NS steps and the MHD step (forceW_amp in p3) against the oracle; the fused state leaves it to the plain
    stepper; k0 outside 1 <= k0 < N/2 or the dealiased modes is refused."""
    from tarang_amd import Plan, State
    log(f"taylor-green force {prec}")
    N = (16, 16, 16)
    o = Oracle(N)
    p = Plan(ctx, N, precision=prec)
    V = o.ic_random(5, E0=0.5, kp=3.0)
    dV = [p.to_dev(a) for a in V]
    if prec == "fp32":
        V = [H(d) for d in dV]
    for k0, amp in [(1, 0.8), (2, -0.5), (4, 0.3)]:
        ref = o.force(V, 3, float(k0), amp, 0.0)
        F, nf = p.compute_force(dV, 3, float(k0), amp, 0.0)
        assert nf == 0
        for c in range(3):
            assert np.abs(H(F[c]) - ref[c]).max() <= tol * abs(amp)
    if prec == "fp64":
        for scheme in (0, 1, 2):
            W = [v.copy() for v in V]
            gW = [p.to_dev(a) for a in W]
            o.step_ns_forced(W, 0.01, 1e-3, 3, 1.0, 0.8, 0.0, 0.0, scheme=scheme, steps=3)
            p.step_plain_forced(gW, 1e-3, 0.01, 3, 1.0, 0.8, 0.0, 0.0, scheme=scheme, steps=3)
            for i in range(3):
                assert rel(H(gW[i]), W[i]) < 1e-12
        U, B = o.ic_random(3, E0=0.5, kp=3.0), o.ic_random(4, E0=0.3, kp=3.0)
        gU, gB = [p.to_dev(a) for a in U], [p.to_dev(a) for a in B]
        o.rk4_mhd_forced(U, B, 0.01, 0.02, 1e-3, 1.0, 0.8, -0.4, 0.0, 0.0, 0.0, steps=3, proc=3)
        p.step_plain_forced_mhd(gU, gB, 1e-3, 0.01, 0.02, 1.0, 0.8, -0.4, 0.0, 0.0, 0.0, steps=3, proc=3)
        for i in range(3):
            assert rel(H(gU[i]), U[i]) < 1e-12
            assert rel(H(gB[i]), B[i]) < 1e-12
        with pytest.raises(RuntimeError, match="plain stepper"):
            State(p).set_force(3, 1.0, 0.8, 0.0, 0.0)
        for k0 in (0.0, 5.0, 6.0, 8.0):
            with pytest.raises(RuntimeError, match="Taylor-Green k0"):
                p.compute_force(dV, 3, k0, 0.8, 0.0)


def test_abc_force_and_steps_match_oracle(ctx):
    """force_field_proc 4 (compute_force_ABC.cc:52-68): k0, A, B, C, amp in inner_radius, outer_radius, p3, p4,
    p5; the field and the forced EULER / RK2 / RK4 NS steps against the oracle; refused for MHD and by the
    fused state."""
    from tarang_amd import Plan, State
    log("abc force")
    N = (16, 16, 16)
    o = Oracle(N)
    p = Plan(ctx, N)
    V = o.ic_random(5, E0=0.5, kp=3.0)
    dV = [p.to_dev(a) for a in V]
    for k0, A, B, C, amp in [(1, 1.0, 0.7, -0.4, 0.5), (4, -0.3, 1.2, 0.9, 2.0)]:
        ref = o.force(V, 4, float(k0), A, B, C, amp)
        F, nf = p.compute_force(dV, 4, float(k0), A, B, C, amp)
        for c in range(3):
            assert np.array_equal(H(F[c]), ref[c])
    for scheme in (0, 1, 2):
        W = [v.copy() for v in V]
        gW = [p.to_dev(a) for a in W]
        o.step_ns_forced(W, 0.01, 1e-3, 4, 1.0, 1.0, 0.7, -0.4, scheme=scheme, steps=3, p5=0.5)
        p.step_plain_forced(gW, 1e-3, 0.01, 4, 1.0, 1.0, 0.7, -0.4, scheme=scheme, steps=3, p5=0.5)
        for i in range(3):
            assert rel(H(gW[i]), W[i]) < 1e-12
    with pytest.raises(RuntimeError, match="plain stepper"):
        State(p).set_force(4, 1.0, 1.0, 0.7, -0.4)
    U, B = p.empty(3), p.empty(3)
    with pytest.raises(RuntimeError, match="not supported for MHD"):
        p.step_plain_forced_mhd(U, B, 1e-3, 0.01, 0.02, 1.0, 1.0, 0.7, -0.4, 0.5, 0.0, proc=4)
    with pytest.raises(RuntimeError, match="ABC k0"):
        p.compute_force(dV, 4, 6.0, 1.0, 0.7, -0.4, 0.5)
