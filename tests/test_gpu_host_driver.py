"""The C++ host classes (tarang_host.hpp: CVF/RVF/NLIN/IncVF/IncFluid/EnergyTr over the C-ABI) run the
reference's time loop in the tarang_run driver; its E(k) and Pi(R) must match the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pyoracle import Oracle  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "tarang_amd", "tarang_run")


def run_driver(*args):
    assert os.path.exists(DRV), "tarang_run not built"
    out = subprocess.run([DRV] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    spec, flux, glob = {}, {}, []
    for line in out.stdout.splitlines():
        f = line.split()
        if f[0] == "spectrum":
            spec[int(f[1])] = float(f[2])
        elif f[0] == "flux":
            flux[float(f[1])] = float(f[2])
        elif f[0] == "glob":
            glob.append([float(x) for x in f[1:]])
    return spec, flux, np.array(glob)


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_driver_taylor_green_matches_oracle(path):
    N, steps, dt, nu = 32, 20, 1e-3, 0.01
    spec, flux, glob = run_driver(N, steps, dt, nu, "tg", "FP64", path)
    o = Oracle((N, N, N))
    V = o.ic_taylor_green(1, 1.0)
    o.rk4_ns(V, nu, dt, steps=steps)
    Ek = o.vector_spectrum(V)
    got = np.array([spec[k] for k in range(len(Ek))])
    sel = Ek >= 1e-10 * Ek.max()
    assert (np.abs(got - Ek)[sel] / Ek[sel]).max() < 1e-10
    assert abs(glob[-1][1] - o.total_energy(V)) < 1e-13
    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    fo = o.flux_ns(V, radii)
    for i, R in enumerate(radii[1:], start=1):
        assert abs(flux[R] - fo[i]) <= 1e-10 * np.abs(fo).max()
