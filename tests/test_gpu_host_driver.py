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


def test_driver_output_files_match_oracle(tmp_path):
    """Output_global / Output_shell_spectrum / Output_flux (output_energy.cc:43-81, 368-394,
    output_ET.cc:45-74): glob.d, spectrum.d and flux.d in the reference's column layout, every column
    equal to the oracle's value of the same quantity to the 12 printed digits."""
    N, steps, dt, nu = 32, 3, 1e-3, 0.01
    (tmp_path / "out").mkdir()
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", "fused", f"out={tmp_path}")
    o = Oracle((N, N, N))
    V = o.ic_random(1)
    o.rk4_ns(V, nu, dt, steps=steps)

    def close(a, b, tol=1e-10):
        return abs(a - b) <= tol * max(abs(b), 1e-300)

    lines = (tmp_path / "out" / "glob.d").read_text().splitlines()
    assert len(lines) == steps
    t, rest = lines[-1].split("\t")
    g = rest.split("    ")
    assert len(g) == 9, lines[-1]
    E, D, Dn = map(float, g[0].split(" "))
    H1, H2 = map(float, g[1].split(" "))
    DH1, DH2 = map(float, g[2].split(" "))
    S, kmax_eta, nu_out, Rl, tdt = float(g[3]), float(g[4]), float(g[5]), float(g[6]), float(g[7])
    m1, m2, m3 = map(float, g[8].split(" "))
    assert close(float(t), steps * dt) and nu_out == nu and tdt == dt and (m1, m2, m3) == (0.0, 0.0, 0.0)
    assert close(E, o.total_energy(V)) and close(D, o.total_dissipation(V)) and close(Dn, nu * D)
    h = o.helicity(V)
    for a, b in zip((H1, H2, DH1, DH2), h):
        assert abs(a - b) <= 1e-10 * np.abs(h).max()
    assert close(S, o.entropy(V))
    kmax = o.max_radius_inside()
    assert close(kmax_eta, kmax * (nu ** 3 / Dn) ** 0.25) and close(Rl, 2 * E * (15 / (nu * Dn)) ** 0.5)

    blocks = (tmp_path / "out" / "spectrum.d").read_text().split("\n")
    assert blocks[0] == f"%% Time = {steps * dt:.12g}"
    rows = [r for r in blocks[1:] if r.strip()]
    ek = [o.shell_spectrum(V[c])[0] for c in range(3)]
    dk = 2 * sum(o.shell_spectrum(V[c], n=2)[0] for c in range(3))
    hk = o.shell_helicity(V)
    assert len(rows) == len(ek[0])
    for i, r in enumerate(rows):
        f = r.split("\t")
        assert int(f[0]) == i and f[-1] == ""
        e1, e2, e3 = map(float, f[1].split(" "))
        h1, h2, h3 = map(float, f[3].split("  "))
        for a, b, sc in ((e1, ek[0][i], ek[0].max()), (e2, ek[1][i], ek[1].max()), (e3, ek[2][i], ek[2].max()),
                         (float(f[2]), nu * dk[i], nu * dk.max()), (h1, hk[0][i], np.abs(hk).max()),
                         (h2, hk[1][i], np.abs(hk).max()), (h3, hk[2][i], np.abs(hk).max())):
            assert abs(a - b) <= 1e-10 * sc, (i, a, b)
        assert float(f[4]) == 0.0

    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    assert len(frows) == len(flux)
    for r in frows:
        i, R, pi, force = r.split()
        assert close(float(pi), flux[float(R)], 1e-11) and float(force) == 0.0


def test_driver_abc_forced_run_matches_oracle():
    """force=4,K0,A,B,C,AMP,0: Set_force(4, k0, A, B, C, amp) leaves the fused state for the plain stepper and
    each stage subtracts the fixed ABC field (compute_force_ABC.cc:52-68); E(t) follows the oracle."""
    N, steps, dt, nu = 16, 5, 1e-3, 0.01
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", "fused", "force=4,1,1.0,0.7,-0.4,0.5,0")
    o = Oracle((N, N, N))
    V = o.ic_random(1)
    E = []
    for _ in range(steps):
        o.step_ns_forced(V, nu, dt, 4, 1.0, 1.0, 0.7, -0.4, p5=0.5)
        E.append(o.total_energy(V))
    assert np.abs(glob[:, 1] - np.array(E)).max() <= 1e-12 * max(E)


@pytest.mark.parametrize("band", [(0.5, 2.5), (4.0, 5.0)])
def test_driver_forced_run_matches_oracle(band):
    """force=1,...: IncFluid::Set_force then Time_advance with the supply force recomputed at every stage
    (Compute_force_TO_rhs single_time_step.cc:351-364); E(t) and E(k) follow the oracle's forced steps.
    The first band runs on the fused step, the second (|k| = N/3 modes, not stored) on the plain one."""
    N, steps, dt, nu = 16, 5, 1e-3, 0.01
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", "fused", f"force=1,{band[0]},{band[1]},0.3,0.2")
    o = Oracle((N, N, N))
    V = o.ic_random(1)
    E = []
    for _ in range(steps):
        o.step_ns_forced(V, nu, dt, 1, band[0], band[1], 0.3, 0.2)
        E.append(o.total_energy(V))
    assert np.abs(glob[:, 1] - np.array(E)).max() <= 1e-12 * max(E)
    Ek = o.vector_spectrum(V)
    for k, v in spec.items():
        assert abs(v - Ek[k]) <= 1e-12 * max(Ek)


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_driver_forced_output_columns(tmp_path, path):
    """Forced run with out=DIR: the force-spectrum column of spectrum.d (Compute_force_shell_spectrum
    force_spectrum.cc:48-53, 2 sum_c Re V_c.F_c* per shell) and the force-feed column of flux.d
    (Compute_force_feed_shell compute_force_feed.cc:70-76, Shell_mult_all(F, V) over shell_radius).  The
    reference writes them after Time_advance, whose stages recompute the force (time_advance.cc:100-155), so
    F is the force of the last RK stage's field, V the new velocity; the oracle's forced step reports that
    field."""
    N, steps, dt, nu = 32, 3, 1e-3, 0.01
    fp = (1, 1.5, 3.5, 0.3, 0.2)
    (tmp_path / "out").mkdir()
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", path, f"out={tmp_path}",
                                  "force=" + ",".join(str(v) for v in fp))
    o = Oracle((N, N, N))
    V = o.ic_random(1)
    last = o.zeros(3)
    o.step_ns_forced(V, nu, dt, *fp, steps=steps, last_stage=last)
    assert abs(glob[-1][1] - o.total_energy(V)) <= 1e-12 * o.total_energy(V)
    F = o.force(last, *fp)
    fk = 2 * sum(o.shell_spectrum(V[c], F[c])[0] for c in range(3))
    rows = [r for r in (tmp_path / "out" / "spectrum.d").read_text().split("\n")[1:] if r.strip()]
    assert len(rows) == len(fk) and np.abs(fk).max() > 0
    for i, r in enumerate(rows):
        assert abs(float(r.split("\t")[4]) - fk[i]) <= 1e-10 * np.abs(fk).max(), (i, r)
    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    feed = o.shell_mult_all(F, V, radii)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    assert np.abs(feed).max() > 0
    for r in frows:
        i, R, pi, force = r.split()
        assert abs(float(force) - feed[int(i)]) <= 1e-10 * np.abs(feed).max(), r
    # the force of the new velocity differs from the last stage's at O(dt): the columns tell them apart
    F_new = o.force(V, *fp)
    fk_new = 2 * sum(o.shell_spectrum(V[c], F_new[c])[0] for c in range(3))
    assert np.abs(fk_new - fk).max() > 1e-6 * np.abs(fk).max()


@pytest.mark.parametrize("proc,fp", [(1, (1.5, 3.5, 0.5, 0.3, 0.4, -0.2)), (1, (1.5, 3.5, 0.5, 0.3, 0.0, 0.0)),
                                     (2, (1.5, 3.5, 0.6, 0.2, 0.4, -0.5)), (3, (1.0, 0.8, -0.4, 0.0, 0.0, 0.0))])
def test_driver_forced_mhd_matches_oracle(tmp_path, proc, fp):
    """mhd=ETA with force=1,INNER,OUTER,P3,P4,P5,P6: Set_force(W, ...) then Time_advance(W) steps U and B with
    both supplies recomputed at every stage, the energies equal to the oracle's forced MHD steps.  The force
    spectra of spectrum.d (force_spectrum.cc:48-53 for U and W) and the force feeds of flux.d
    (compute_force_feed.cc:130-139) come from the forces of the last RK stage's U and B (Force1..3 and
    W.Force1..3 after Time_advance(W)); a zero W supply leaves W's columns 0."""
    N, steps, dt, nu, eta = 32, 3, 1e-3, 0.01, 0.02
    (tmp_path / "out").mkdir()
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", "fused", f"mhd={eta}", f"out={tmp_path}",
                                  f"force={proc}," + ",".join(str(v) for v in fp))
    o = Oracle((N, N, N))
    U, B = o.ic_random(1), o.ic_random(2)
    last = o.zeros(6)
    o.rk4_mhd_forced(U, B, nu, eta, dt, *fp, steps=steps, last_stage=last, proc=proc)
    Eu, Eb = o.total_energy(U), o.total_energy(B)
    assert abs(glob[-1][1] - Eu) <= 1e-12 * Eu and abs(glob[-1][2] - Eb) <= 1e-12 * Eb
    U0, B0 = o.ic_random(1), o.ic_random(2)
    o.rk4_mhd(U0, B0, nu, eta, dt, steps=steps)
    if proc != 3:                                               # the supply forces act at O(1e-3)
        assert abs(glob[-1][1] - o.total_energy(U0)) > 1e-6
    inner, outer, p3, p4, p5, p6 = fp
    FU, FW = o.force_mhd(last[:3], last[3:], proc, inner, outer, p3, p4, p5, p6)
    fku = 2 * sum(o.shell_spectrum(U[c], FU[c])[0] for c in range(3))
    fkw = 2 * sum(o.shell_spectrum(B[c], FW[c])[0] for c in range(3))
    rows = [r for r in (tmp_path / "out" / "spectrum.d").read_text().split("\n")[1:] if r.strip()]
    assert len(rows) == len(fku) and np.abs(fku).max() > 0
    sc = np.abs(fku).max()
    for i, r in enumerate(rows):
        a, b = (float(x) for x in r.split("\t")[-1].split())
        assert abs(a - fku[i]) <= 1e-10 * sc and abs(b - fkw[i]) <= 1e-10 * sc, (i, r)
    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    feu, few = o.shell_mult_all(FU, U, radii), o.shell_mult_all(FW, B, radii)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    sc = np.abs(feu).max()
    assert sc > 0 and len(frows) > 0
    for r in frows:
        t = r.split()
        assert len(t) == 13, r
        i = int(t[0])
        assert abs(float(t[11]) - feu[i]) <= 1e-10 * sc and abs(float(t[12]) - few[i]) <= 1e-10 * sc, r
    if p5 > 0 or proc == 3:
        assert np.abs(fkw).max() > 0 and np.abs(few).max() > 0


def test_driver_mhd_output_files_match_oracle(tmp_path):
    """mhd=ETA: Get_dt(W), Time_advance(W) (fused MHD step) and Output_global(W) / Output_shell_spectrum(W) /
    Output_flux(W) (output_energy.cc:217-286, 463-500, output_ET.cc:118-165) in the reference's column
    layout, every number equal to the oracle's value of the same quantity to the printed digits.  The
    cross-helicity spectrum, W.flux_VF_out_out and the force columns are printed as 0, as the reference
    leaves them on this path."""
    N, steps, dt, nu, eta = 32, 3, 1e-3, 0.01, 0.02
    (tmp_path / "out").mkdir()
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", "fused", f"out={tmp_path}", f"mhd={eta}")
    o = Oracle((N, N, N))
    U, B = o.ic_random(1), o.ic_random(2)
    o.rk4_mhd(U, B, nu, eta, dt, steps=steps)
    Eu, Eb = o.total_energy(U), o.total_energy(B)
    assert abs(glob[-1][1] - Eu) <= 1e-12 * Eu and abs(glob[-1][2] - Eb) <= 1e-12 * Eb

    def close(a, b, tol=1e-10, sc=None):
        return abs(a - b) <= tol * (sc if sc is not None else max(abs(b), 1e-300))

    line = (tmp_path / "out" / "glob.d").read_text().splitlines()[-1]
    t, mid, mb = line.split("\t")
    assert mid.count("    ") == 9 and mb.split(" ") == ["0", "0", "0"]
    v = [float(x) for x in mid.split()]
    assert len(v) == 26
    Du, Db = o.total_dissipation(U), o.total_dissipation(B)
    Hc = sum(o.energy(U[c], B[c]) for c in range(3))
    hu, hb = o.helicity(U), o.helicity(B)
    kmax = o.max_radius_inside()
    ke1 = kmax * (nu ** 3 / (nu * Du)) ** 0.25
    ref = [Eu, Eb, Du, Db, nu * Du, eta * Db, Hc, hu[0], hu[1], hb[0], hb[1], hu[2], hu[3], hb[2], hb[3],
           o.entropy(U), o.entropy(B), ke1, ke1 * (nu / eta) ** 0.25 / (nu / eta), nu, eta,
           2 * Eu * (15 / (nu * nu * Du)) ** 0.5, dt, 0.0, 0.0, 0.0]
    for i, (a, b) in enumerate(zip(v, ref)):
        assert close(a, b, sc=max(abs(b), 1e-12 * max(abs(x) for x in ref[7:15]))), (i, a, b)
    assert close(float(t), steps * dt)

    rows = [r for r in (tmp_path / "out" / "spectrum.d").read_text().split("\n")[1:] if r.strip()]
    eu = [o.shell_spectrum(U[c])[0] for c in range(3)]
    eb = [o.shell_spectrum(B[c])[0] for c in range(3)]
    du = 2 * sum(o.shell_spectrum(U[c], n=2)[0] for c in range(3))
    db = 2 * sum(o.shell_spectrum(B[c], n=2)[0] for c in range(3))
    hku, hkb = o.shell_helicity(U), o.shell_helicity(B)
    assert len(rows) == len(eu[0])
    for i, r in enumerate(rows):
        f = r.split("\t")
        assert int(f[0]) == i and len(f) == 6 and f[5] == "0 0"
        got = [float(x) for x in " ".join(f[1:5]).split()]
        ref = [eu[0][i], eu[1][i], eu[2][i], eb[0][i], eb[1][i], eb[2][i], nu * du[i], eta * db[i], 0.0,
               hku[0][i], hku[1][i], hku[2][i], hkb[0][i], hkb[1][i], hkb[2][i]]
        scale = [max(np.abs(eu).max(), 1e-300)] * 3 + [np.abs(eb).max()] * 3 + [nu * du.max(), eta * db.max(), 1.0] + \
                [np.abs(hku).max()] * 3 + [np.abs(hkb).max()] * 3
        for j, (a, b, sc) in enumerate(zip(got, ref, scale)):
            assert abs(a - b) <= 1e-10 * sc, (i, j, a, b)

    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    fl = o.flux_mhd([a.copy() for a in U], [a.copy() for a in B], radii, elsasser=True)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    assert len(frows) == len(radii) - 1
    order = ["u2u", "u2b_in_out", "u2b_in_in", "b2b", "b2u_in_out", "b2u_in_in", None, "zp", "zm", None, None]
    for r in frows:
        x = r.split()
        i = int(x[0])
        assert float(x[1]) == radii[i] and len(x) == 13
        for tok, name in zip(x[2:], order):
            if name is None:
                assert float(tok) == 0.0
            else:
                assert abs(float(tok) - fl[name][i]) <= 1e-10 * max(np.abs(fl[name]).max(), 1e-300), (r, name)


def test_driver_rb_glob_matches_oracle(tmp_path):
    """rb=RA,PR: the RB_SLIP program's loop (Get_dt(T), Time_advance(T) on the fused SCFT step) and
    Output_global_RB (output_energy.cc:156-215): every glob.d number (energies, dissipations, Nusselt number,
    SCFT helicities, vector and scalar entropies, Kolmogorov and Taylor-scale figures, means) equals the
    oracle's value to the printed digits."""
    import math
    N, steps, dt, Ra, Pr = 32, 4, 1e-3, 1e6, 1.0
    (tmp_path / "out").mkdir()
    spec, flux, glob = run_driver(N, steps, dt, 0.0, "random", "FP64", "fused", f"out={tmp_path}", f"rb={Ra},{Pr}")
    kf = (math.pi, math.pi / math.sqrt(2), math.pi / math.sqrt(2))
    o = Oracle((N, N, N), kfactor=kf, basis="SCFT")
    V, T = o.zeros(3), o.ic_random_scalar(4, 1e-3)
    last = o.zeros(4)
    o.rk4_rb(V, T, Ra, Pr, dt, steps=steps, last_stage=last)
    nu, kappa = math.sqrt(Pr / Ra), 1 / math.sqrt(Pr * Ra)
    Ev, Et = o.total_energy(V), o.energy(T)
    assert abs(glob[-1][1] - Ev) <= 1e-12 * Ev and abs(glob[-1][2] - Et) <= 1e-12 * Et
    line = (tmp_path / "out" / "glob.d").read_text().splitlines()[-1]
    assert line.count("\t") == 4 and line.count("    ") == 8
    v = [float(x) for x in line.split()]
    assert len(v) == 24
    Dv, Dt = o.total_dissipation(V), 2 * o.Sn(T, 2)
    h = o.helicity(V)
    kmax = o.max_radius_inside()
    ke1 = kmax * (nu ** 3 / (nu * Dv)) ** 0.25
    ref = [steps * dt, Ev, Et, Dv, Dt, nu * Dv, kappa * Dt, o.nusselt(V, T, Ra, Pr), h[0], h[1], h[2], h[3],
           o.entropy(V), o.entropy_scalar(T), ke1, ke1 * (nu / kappa) ** 0.25 / (nu / kappa), nu, kappa,
           2 * Ev * (15 / (nu * nu * Dv)) ** 0.5, dt, 0.0, 0.0, 0.0, 0.0]
    # the RB field carries no helicity beyond rounding: compare H1, H2, DH1, DH2 on their natural scale
    # kmax^2 sqrt(E D) (Cauchy-Schwarz bound) instead of their own roundoff-sized values
    hs = kmax ** 2 * math.sqrt(Ev * Dv)
    for i, (a, b) in enumerate(zip(v, ref)):
        if i >= 20:                                  # mean modes: zero up to rounding
            assert abs(a) <= 1e-15, (i, a)
            continue
        sc = max(abs(b), hs) if 8 <= i <= 11 else max(abs(b), 1e-300)
        assert abs(a - b) <= 1e-10 * sc, (i, a, b)

    # spectrum.d (Output_shell_spectrum_scalar output_energy.cc:410-449) and flux.d (output_ET.cc:81-115);
    # the forces are those of the last RK stage: Force1 = T_s, T.Force = V1_s, dealiased
    Ts, V1s = last[3].copy(), last[0].copy()
    o.dealias(Ts)
    o.dealias(V1s)
    Z = o.zeros()
    ek = [o.shell_spectrum(V[c])[0] for c in range(3)]
    et = o.shell_spectrum(T)[0]
    dk = 2 * sum(o.shell_spectrum(V[c], n=2)[0] for c in range(3))
    cr = [o.shell_spectrum(V[c], T)[0] for c in range(3)]
    hk = o.shell_helicity(V)
    fv = 2 * o.shell_spectrum(V[0], Ts)[0]
    ft = 2 * o.shell_spectrum(T, V1s)[0]
    cols = [ek[0], ek[1], ek[2], et, nu * dk, np.zeros_like(et), cr[0], cr[1], cr[2], hk[0], hk[1], hk[2], fv, ft]
    rows = [r for r in (tmp_path / "out" / "spectrum.d").read_text().split("\n")[1:] if r.strip()]
    assert len(rows) == len(et)
    for i, r in enumerate(rows):
        f = r.split("\t")
        assert int(f[0]) == i and len(f) == 7
        got = [float(x) for x in " ".join(f[1:]).split()]
        assert len(got) == 14
        for j, (a, col) in enumerate(zip(got, cols)):
            sc = np.abs(col).max()
            if j in (9, 10, 11):                     # helicity: natural scale, the field carries none
                sc = max(sc, kmax ** 2 * math.sqrt(Ev * Dv))
            assert abs(a - col[i]) <= 1e-10 * max(sc, 1e-300), (i, j, a, col[i])
    assert np.abs(fv).max() > 0 and np.abs(ft).max() > 0
    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    fs, fsf = o.flux_scalar(V, T, radii)
    feedV = o.shell_mult_all([Ts, Z, Z], V, radii)
    feedT = o.shell_mult_all([V1s, Z, Z], [T, Z, Z], radii)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    assert len(frows) == len(radii) - 1
    for r in frows:
        x = r.split()
        i = int(x[0])
        assert len(x) == 6 and float(x[1]) == radii[i]
        for tok, col in zip(x[2:], (fs, fsf, feedV, feedT)):
            assert abs(float(tok) - col[i]) <= 1e-10 * max(np.abs(col).max(), 1e-300), (r, tok, col[i])
