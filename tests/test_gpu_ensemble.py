"""GPU parity of the uneven-box ensemble (th_ensemble_*): every box of a batched, graph-captured
ensemble step equals the same box advanced alone by the single-box fused step (identical kernels
and per-box arithmetic, so bit-exact), the single-box step equals the oracle's RK4, and per-box
dt / viscosity changes between calls take effect without recapturing the step."""
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


def _run_pair(ctx, shapes, prec, dts, nus, steps):
    """Advance an ensemble and, box by box, single-box states from the same ICs."""
    from tarang_amd import Ensemble, Plan, State
    log(f"ensemble {prec} shapes={shapes} steps={steps}")
    ens = Ensemble(ctx, shapes, precision=prec)
    singles = []
    for b, N in enumerate(shapes):
        V = ens.plan(b).ic_random(100 + b)
        ens.upload(b, V)
        p = Plan(ctx, N, precision=prec)
        s = State(p)
        s.upload(V)
        singles.append((p, s))
    for dt, nu in zip(dts, nus):
        ens.step(dt, nu, steps=steps)
        for b, (p, s) in enumerate(singles):
            s.step(float(dt[b]), nu=float(nu[b]), steps=steps)
    return ens, singles


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ensemble_matches_single_boxes(ctx, prec):
    shapes = [(16, 16, 16), (24, 24, 24), (16, 16, 16), (32, 32, 32), (24, 24, 24), (16, 16, 16)]
    rng = np.random.default_rng(5)
    nus = [rng.uniform(5e-4, 5e-3, len(shapes))]
    dts = [rng.uniform(1e-3, 3e-3, len(shapes))]
    ens, singles = _run_pair(ctx, shapes, prec, dts, nus, steps=3)
    for b, (p, s) in enumerate(singles):
        got, ref = ens.download(b), s.download()
        for i in range(3):
            assert np.array_equal(H(got[i]), H(ref[i])), (b, i)


def test_ensemble_fp64_matches_oracle(ctx):
    shapes = [(16, 16, 16), (32, 32, 32), (24, 24, 24)]
    nu = np.array([1e-3, 4e-3, 2.5e-3])
    dt = np.array([2e-3, 1.5e-3, 1e-3])
    from tarang_amd import Ensemble
    ens = Ensemble(ctx, shapes, precision="fp64")
    Vo = []
    for b, N in enumerate(shapes):
        o = Oracle(N)
        V = o.ic_random(100 + b)
        ens.upload(b, [ens.plan(b).to_dev(v) for v in V])
        Vo.append((o, V))
    ens.step(dt, nu, steps=4)
    for b, (o, V) in enumerate(Vo):
        o.rk4_ns(V, nu[b], dt[b], steps=4)
        got = ens.download(b)
        for i in range(3):
            assert rel(H(got[i]), V[i]) < 1e-12, (b, i)


def test_ensemble_dt_and_nu_change_between_calls(ctx):
    shapes = [(16, 16, 16), (32, 32, 32), (16, 16, 16)]
    dts = [np.array([1e-3, 2e-3, 3e-3]), np.array([2.5e-3, 1e-3, 0.5e-3]), np.array([1e-3, 2e-3, 3e-3])]
    nus = [np.array([1e-3, 2e-3, 3e-3]), np.array([4e-3, 1e-3, 2e-3]), np.array([5e-4, 5e-4, 5e-4])]
    ens, singles = _run_pair(ctx, shapes, "fp64", dts, nus, steps=2)
    for b, (p, s) in enumerate(singles):
        got, ref = ens.download(b), s.download()
        for i in range(3):
            assert np.array_equal(H(got[i]), H(ref[i])), (b, i)


def test_single_state_dt_change_matches_oracle(ctx):
    """th_step reuses its captured step across dt / nu changes (coefficients live in device memory)."""
    from tarang_amd import Plan, State
    N = (32, 32, 32)
    o = Oracle(N)
    p = Plan(ctx, N)
    s = State(p)
    V = o.ic_random(7)
    s.upload([p.to_dev(v) for v in V])
    for dt, nu in [(1e-3, 0.01), (2e-3, 0.005), (1e-3, 0.01)]:
        o.rk4_ns(V, nu, dt, steps=2)
        s.step(dt, nu=nu, steps=2)
    out = s.download()
    for i in range(3):
        assert rel(H(out[i]), V[i]) < 1e-12


def test_ensemble_uneven_64_boxes_fp32(ctx):
    """The 64-box uneven ensemble (24 x 64^3, 24 x 96^3, 16 x 128^3), FP32-mixed: sampled boxes are
    bit-identical to single-box runs and every box's energy decays (no forcing)."""
    from tarang_amd import Ensemble, Plan, State
    shapes = [(64,) * 3] * 24 + [(96,) * 3] * 24 + [(128,) * 3] * 16
    rng = np.random.default_rng(100)
    nu = rng.uniform(5e-4, 5e-3, len(shapes))
    dt = np.full(len(shapes), 1e-3)
    log("ensemble 64 boxes fp32")
    ens = Ensemble(ctx, shapes, precision="fp32")
    E0 = []
    for b in range(len(shapes)):
        p = ens.plan(b)
        V = p.ic_random(100 + b)
        ens.upload(b, V)
        E0.append(p.total_energy(V))
    sample = [0, 23, 24, 47, 48, 63]
    singles = {}
    for b in sample:
        p = Plan(ctx, shapes[b], precision="fp32")
        s = State(p)
        s.upload(ens.plan(b).ic_random(100 + b))
        singles[b] = s
    ens.step(dt, nu, steps=2)
    for b in sample:
        singles[b].step(float(dt[b]), nu=float(nu[b]), steps=2)
        got, ref = ens.download(b), singles[b].download()
        for i in range(3):
            assert np.array_equal(H(got[i]), H(ref[i])), (b, i)
    for b in range(len(shapes)):
        p = ens.plan(b)
        E1 = p.total_energy(ens.download(b))
        assert 0 < E1 < E0[b], (b, E0[b], E1)


def test_tuned_ensemble_matches_single_boxes_on_its_variants(ctx):
    """After th_ensemble_tune, th_ensemble_variants names each size class's launch variants; single-box
    states set to those variants (th_state_set_variants) reproduce the tuned, chained ensemble bit for
    bit; a box index outside the ensemble is refused."""
    from tarang_amd import Ensemble, Plan, State
    shapes = [(128,) * 3, (64,) * 3, (128,) * 3, (96,) * 3]
    rng = np.random.default_rng(11)
    nu = rng.uniform(5e-4, 5e-3, len(shapes))
    dt = rng.uniform(1e-3, 2e-3, len(shapes))
    log("tuned ensemble fp32 shapes=" + str(shapes))
    ens = Ensemble(ctx, shapes, precision="fp32")
    for b in range(len(shapes)):
        ens.upload(b, ens.plan(b).ic_random(200 + b))
    ens.tune(dt, nu, nrep=1)
    singles = []
    for b, N in enumerate(shapes):
        var = ens.variants(b)
        assert len(var) == 5 and all(v >= 0 for v in var), var
        assert var == ens.variants([i for i, M in enumerate(shapes) if M == N][0])
        s = State(Plan(ctx, N, precision="fp32"))
        s.upload(ens.plan(b).ic_random(200 + b))
        s.variants = var
        assert s.variants == var
        singles.append(s)
    ens.step(dt, nu, steps=2)
    for b, s in enumerate(singles):
        s.step(float(dt[b]), nu=float(nu[b]), steps=2)
        got, ref = ens.download(b), s.download()
        for i in range(3):
            assert np.array_equal(H(got[i]), H(ref[i])), (b, i, ens.variants(b))
    with pytest.raises(RuntimeError):
        ens.variants(len(shapes))
