"""GPU parity of the anisotropic spectra (th_ring_spectrum, th_cylinder_ring_spectrum) against the FP64
oracle: bins compared exactly in location (the same modes land in the same (row, column)), values to
1e-12 of the largest bin in FP64 and 1e-5 in FP32-mixed.  Reference: four_energy.cc:828-954, 1450-1515."""
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


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,kf", [((16, 16, 16), (1.0, 1.0, 1.0)), ((32, 24, 48), (1.0, 1.0, 1.0)),
                                  ((24, 32, 16), (1.0, 0.5, 2.0))])
@pytest.mark.parametrize("axis", [1, 2, 3])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_ring_and_cylinder_spectra_match_oracle(ctx, N, kf, axis, prec, tol):
    from tarang_amd import Plan
    log(f"aniso N={N} kf={kf} axis={axis} {prec}")
    o = Oracle(N, kfactor=kf)
    V = o.ic_random(4)
    p = Plan(ctx, N, kfactor=kf, precision=prec)
    dV = [p.to_dev(a) for a in V]
    if prec == "fp32":
        V = [H(d) for d in dV]
    for nb in (4, 6):
        for n in (0, 2):
            r1, r2 = o.ring_spectrum(V, nb, n, axis)
            g1, g2 = p.ring_spectrum(dV, nb, n, axis)
            assert g1.shape == r1.shape
            m = max(np.abs(r1).max(), np.abs(r2).max())
            assert np.abs(g1 - r1).max() <= tol * m and np.abs(g2 - r2).max() <= tol * m
            assert np.array_equal(g1 != 0, r1 != 0) and np.array_equal(g2 != 0, r2 != 0)
            c1, c2 = o.cylinder_ring_spectrum(V, nb, n, axis)
            h1, h2 = p.cylinder_ring_spectrum(dV, nb, n, axis)
            assert h1.shape == c1.shape
            m = max(np.abs(c1).max(), np.abs(c2).max())
            assert np.abs(h1 - c1).max() <= tol * m and np.abs(h2 - c2).max() <= tol * m
        # helicity tables on the same bins (four_energy.cc:1357-1412, 1825-1880)
        rh, gh = o.ring_spectrum_helicity(V, nb, axis), p.ring_spectrum_helicity(dV, nb, axis)
        assert gh.shape == rh.shape and np.abs(gh - rh).max() <= tol * np.abs(rh).max()
        ch, gc = o.cylinder_ring_spectrum_helicity(V, nb, axis), p.cylinder_ring_spectrum_helicity(dV, nb, axis)
        assert gc.shape == ch.shape and np.abs(gc - ch).max() <= tol * np.abs(ch).max()


def test_aniso_exact_boundary_modes(ctx):
    """Modes on sector edges (theta = pi/4 from Pythagorean K, K_perp = K_pll) and slab edges fall in the same
    bins as in the oracle, and the past-the-last-slab store wraps to the next row as in the oracle."""
    from tarang_amd import Plan
    log("aniso boundary modes")
    N = (32, 32, 32)
    o = Oracle(N)
    V = o.zeros(3)
    for (l1, l2, l3) in [(3, 4, 5), (4, 3, 5), (0, 5, 5), (5, 0, 5), (6, 8, 10), (2, 2, 8), (0, 0, 12), (1, 0, 14)]:
        V[0][l1, l2, l3] = 0.3 + 0.1j
        V[1][l1, l2, l3] = -0.2
    p = Plan(ctx, N)
    dV = [p.to_dev(a) for a in V]
    for nb in (2, 4, 8):
        r1, r2 = o.ring_spectrum(V, nb, 0, 3)
        g1, g2 = p.ring_spectrum(dV, nb, 0, 3)
        assert np.abs(g1 - r1).max() <= 1e-15 and np.abs(g2 - r2).max() <= 1e-15
        c1, c2 = o.cylinder_ring_spectrum(V, nb, 0, 3)
        h1, h2 = p.cylinder_ring_spectrum(dV, nb, 0, 3)
        assert np.abs(h1 - c1).max() <= 1e-15 and np.abs(h2 - c2).max() <= 1e-15


def test_aniso_errors(ctx):
    from tarang_amd import Plan
    p = Plan(ctx, (16, 16, 16))
    dV = p.empty(3)
    with pytest.raises(RuntimeError, match="axis"):
        p.ring_spectrum(dV, 4, 0, 4)
    with pytest.raises(RuntimeError, match="n in"):
        p.ring_spectrum(dV, 4, 1, 3)
    with pytest.raises(RuntimeError, match="LDS"):
        p.ring_spectrum(dV, 5000, 0, 3)
