"""Launch-shape variants and the autotuner (th_tune, th_state_set_variants, th_ensemble_tune):
every variant of every fused pass computes the same RK4 step as the oracle, tuning leaves the
state unchanged, and the tune cache carries the choice to a new context."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pyoracle import Oracle  # noqa: E402

NVAR = [4, 4, 4, 4, 3]


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


@pytest.fixture(scope="module", params=[(128, 128, 128), (64, 64, 64)])
def ref_ns(request):
    N = request.param
    o = Oracle(N)
    V0 = o.ic_random(31)
    V = [a.copy() for a in V0]
    o.rk4_ns(V, 2e-3, 1e-3, steps=1)
    return N, V0, V


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_variants_match_oracle_fp64(ctx, ref_ns, v):
    from tarang_amd import Plan, State
    N, V0, V1 = ref_ns
    var = [min(v, n - 1) for n in NVAR]
    log(f"variants {var} fp64 N={N}")
    p = Plan(ctx, N)
    s = State(p)
    s.variants = var
    assert s.variants == var
    s.upload([p.to_dev(a) for a in V0])
    s.step(1e-3, nu=2e-3, steps=1)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V1[i]) < 1e-12, (var, i)


@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_variants_fp32_mixed(ctx, ref_ns, v):
    from tarang_amd import Plan, State
    N, V0, V1 = ref_ns
    var = [min(v, n - 1) for n in NVAR]
    log(f"variants {var} fp32 N={N}")
    p = Plan(ctx, N, precision="fp32")
    s = State(p)
    s.variants = var
    s.upload([p.to_dev(a) for a in V0])
    s.step(1e-3, nu=2e-3, steps=1)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V1[i]) < 2e-5, (var, i)


def test_tune_keeps_state_and_step_matches_oracle(ctx, ref_ns):
    from tarang_amd import Plan, State
    N, V0, V1 = ref_ns
    p = Plan(ctx, N)
    s = State(p)
    s.upload([p.to_dev(a) for a in V0])
    before = [H(a) for a in s.download()]
    var = s.tune(1e-3, nu=2e-3, nrep=2)
    log(f"tuned variants {var}")
    assert all(0 <= x < n for x, n in zip(var, NVAR))
    after = [H(a) for a in s.download()]
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    s.step(1e-3, nu=2e-3, steps=1)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V1[i]) < 1e-12


def test_tune_cache_round_trip(tmp_path):
    import torch
    assert torch.cuda.is_available()
    from tarang_amd import Context, Plan, State
    path = str(tmp_path / "tune.txt")
    c1 = Context(0, tune_cache=path)
    p = Plan(c1, (128, 128, 128), precision="fp32")
    s = State(p)
    p.ic_random(3)
    s.upload(p.ic_random(3))
    var = s.tune(1e-3, nu=1e-3, nrep=2)
    assert os.path.exists(path)
    text = open(path).read()
    assert "fused:gfx950:fp32:ns:N128x128x128:boxes1" in text
    s.close()
    p.close()
    c2 = Context(0, tune_cache=path)
    p2 = Plan(c2, (128, 128, 128), precision="fp32")
    s2 = State(p2)
    assert s2.variants == var
    s2.close()
    p2.close()


def test_ensemble_tune_then_step_matches_single_boxes(ctx):
    from tarang_amd import Ensemble, Plan, State
    shapes = [(128, 128, 128), (64, 64, 64), (128, 128, 128)]
    dt = np.array([1e-3, 2e-3, 1.5e-3])
    nu = np.array([1e-3, 3e-3, 2e-3])
    ens = Ensemble(ctx, shapes, precision="fp64")
    singles = []
    for b, N in enumerate(shapes):
        V = ens.plan(b).ic_random(100 + b)
        ens.upload(b, V)
        p = Plan(ctx, N)
        s = State(p)
        s.upload(V)
        singles.append(s)
    ens.tune(dt, nu, nrep=2)
    ens.step(dt, nu, steps=2)
    for b, s in enumerate(singles):
        s.step(float(dt[b]), nu=float(nu[b]), steps=2)
        got, ref = ens.download(b), s.download()
        for i in range(3):
            assert rel(H(got[i]), H(ref[i])) < 1e-12, (b, i)


@pytest.fixture(scope="module")
def ref_rb128():
    import math
    N = (128, 128, 128)
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    o = Oracle(N, kfactor=kf, basis="SCFT")
    V = o.zeros(3)
    T = o.dealias(o.ic_random_scalar(7, 0.1))
    o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=1)        # non-zero velocity from the buoyancy force
    V0 = [a.copy() for a in V + [T]]
    o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=1)
    return N, kf, V0, V + [T]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("v", [0, 1, 2, 3])
def test_variants_rb_scft(ctx, ref_rb128, v, prec):
    """Every launch variant of the fused RB (SCFT) step gives the oracle's Time_advance(IncSF&)."""
    from tarang_amd import Plan, State, abi
    N, kf, V0, V1 = ref_rb128
    var = [min(v, n - 1) for n in NVAR]
    log(f"rb variants {var} {prec} N={N}")
    p = Plan(ctx, N, kfactor=kf, basis="SCFT", precision=prec)
    s = State(p, abi.TH_RB)
    s.variants = var
    s.upload([p.to_dev(a) for a in V0])
    s.step(1e-2, Ra=1e5, Pr=1.0, steps=1)
    out = s.download()
    tol = 1e-12 if prec == "fp64" else 2e-5
    for i in range(4):
        assert rel(H(out[i]), V1[i]) < tol, (var, i)


def test_flux_shell_tr_with_tuned_y_variants(ctx):
    """th_flux_shell_tr runs its y passes with the state's launch shapes: every y-variant pair gives the
    default-shape result to 1e-12 (128^3, lines long enough for all four y variants)."""
    import numpy as np
    from tarang_amd import Plan, State
    from pyoracle import Oracle
    N = (128, 128, 64)
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p)
    V = o.ic_random(5, kp=3.0)
    s.upload([p.to_dev(v) for v in V])
    radii = np.concatenate([np.arange(8, dtype=float), [10000.0]])
    s.variants = [0, 0, 0, 0, 0]
    Pi0, T0 = s.flux_shell_tr(radii)
    for vy in (1, 2, 3):
        s.variants = [0, vy, 0, vy, 0]
        Pi, T = s.flux_shell_tr(radii)
        assert np.abs(Pi - Pi0).max() <= 1e-12 * np.abs(Pi0).max()
        assert np.abs(T - T0).max() <= 1e-12 * np.abs(T0).max()
