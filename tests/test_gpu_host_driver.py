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


def test_driver_restart_from_field_file(tmp_path):
    """Output_field then Init_cond (field_input_proc = 1): 10 steps, save, 10 more from the file equal
    20 steps in one run to the 12 printed digits of the restart file; the file's "%% Time" is the
    restart time."""
    N, dt, nu = 32, 1e-3, 0.01
    f = tmp_path / "field_out.d"
    spec20, flux20, glob20 = run_driver(N, 20, dt, nu, "tg", "FP64", "fused")
    run_driver(N, 10, dt, nu, "tg", "FP64", "fused", f"save={f}")
    assert f.read_text().startswith("%% Time = 0.01\n")
    spec, flux, glob = run_driver(N, 10, dt, nu, f"file={f}", "FP64", "fused")
    assert abs(glob[-1][0] - glob20[-1][0]) < 1e-12
    ks = [k for k in spec20 if spec20[k] > 1e-10 * max(spec20.values())]
    for k in ks:
        assert abs(spec[k] - spec20[k]) <= 1e-10 * spec20[k], k
    for r in flux20:
        assert abs(flux[r] - flux20[r]) <= 1e-10 * max(abs(v) for v in flux20.values()), r
