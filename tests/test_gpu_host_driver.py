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


def g6(a, b, sc=0.0, rel=1e-10):
    """a was printed by the default stream format (6 significant digits: spectrum.d, flux.d, shell.d), b is
    the oracle's value; rel * sc bounds the GPU-vs-oracle difference on the scale sc."""
    return abs(a - b) <= 5.0000001e-6 * abs(b) + rel * sc + 1e-300


def f12(a, b, sc=0.0, rel=1e-10):
    """a was printed fixed with 12 decimals (glob.d, MY_PRECISION)."""
    return abs(a - b) <= 5.0000001e-13 + rel * max(sc, abs(b))


def blitz_2d(lines, i):
    """Parse one Blitz++ array printed as its extents 'R x C [x ...]', then '[ v v ... ]', from lines[i];
    returns (array, next i)."""
    dims = [int(v) for v in lines[i].split(" x ")]
    vals = []
    i += 1
    while True:
        toks = lines[i].replace("[", " ").replace("]", " ").split()
        vals += [float(t) for t in toks]
        if "]" in lines[i]:
            break
        i += 1
    assert len(vals) == int(np.prod(dims)), (dims, len(vals))
    return np.array(vals).reshape(dims), i + 1


def read_shell_d(path):
    """shell.d as Output_prefix, '%% Shellradius = ' (a Blitz++ 1-D array) and, per output time, '%% Time = t'
    and the named Blitz++ 2-D matrices ('%% U to U: ' ...); returns (header lines, radii, [(t, {name: matrix})])."""
    lines = path.read_text().split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("%% Shellradius = "))
    head = lines[:i]
    n = int(lines[i][len("%% Shellradius = "):])
    radii = []
    i += 1
    while len(radii) < n:
        radii += [float(t) for t in lines[i].replace("[", " ").replace("]", " ").split()]
        i += 1
    out = []
    while i < len(lines):
        if lines[i].startswith("%% Time = "):
            out.append((float(lines[i][10:]), {}))
            i += 1
        elif lines[i].startswith("%% ") and lines[i].endswith(": "):
            name = lines[i][3:-2]
            m, i = blitz_2d(lines, i + 1)
            out[-1][1][name] = m
        else:
            i += 1
    return head, radii, out


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
    20 steps in one run to the 12 printed decimals of the restart file; the file's "%% Time" is the
    restart time.  The file rounds every value to 5e-13 (fixed notation), which moves a shell's energy by
    at most 1e-12 sum |u| <= 1e-12 sqrt(n_k E(k)) with n_k ~ 4 pi k^2 modes: E(k) is compared to
    1e-10 E(k) + 1e-11 k sqrt(E(k))."""
    N, dt, nu = 32, 1e-3, 0.01
    f = tmp_path / "field_out.d"
    spec20, flux20, glob20 = run_driver(N, 20, dt, nu, "tg", "FP64", "fused")
    run_driver(N, 10, dt, nu, "tg", "FP64", "fused", f"save={f}")
    assert f.read_text().startswith("%% Time = 0.010000000000\n")
    spec, flux, glob = run_driver(N, 10, dt, nu, f"file={f}", "FP64", "fused")
    assert abs(glob[-1][0] - glob20[-1][0]) < 1e-12
    ks = [k for k in spec20 if spec20[k] > 1e-10 * max(spec20.values())]
    for k in ks:
        assert abs(spec[k] - spec20[k]) <= 1e-10 * spec20[k] + 1e-11 * k * np.sqrt(spec20[k]), k
    for r in flux20:
        assert abs(flux[r] - flux20[r]) <= 1e-10 * max(abs(v) for v in flux20.values()) + 1e-11, r


def test_driver_reduced_field_files(tmp_path):
    """Output_field_reduced (output_field.cc:505-520) and Init_cond_reduced (field_input_proc = 2,
    init_cond_field.cc:292-330) through the driver: the reduced file of a 32^3 run holds that run's modes of the
    16^3 grid (against the full field_out.d of the same run, 6 printed digits); a 16^3 run restarted at 32^3 from
    its field_out.d read as a reduced file keeps its modes: E(k) is unchanged in the shells the coarse grid holds
    whole (k <= 16/3; beyond, the coarse spectrum rescales its partial shells), and no energy appears past the
    coarse corner |K| = 5 sqrt(3)."""
    from tarang_amd import field_read, field_read_reduced
    full, red = tmp_path / "field_out.d", tmp_path / "field_out_reduced.d"
    run_driver(32, 4, 1e-3, 0.01, "random", "FP64", "fused", f"save={full}")
    run_driver(32, 4, 1e-3, 0.01, "random", "FP64", "fused", f"save={red}", "save_reduced=16,16,16")
    assert red.read_text().startswith("%% Time = 0.004\n")
    (u,), _ = field_read(full, (32, 32, 32), 1)
    (r,), t = field_read_reduced(red, (32, 32, 32), (16, 16, 16), 1)
    assert t == 0.004
    for i in list(range(0, 9)) + list(range(25, 32)):        # kx = 0..8, -7..-1 (and -8 = the +8 copy)
        for j in list(range(0, 9)) + list(range(25, 32)):
            a, b = r[i, j, :9], u[i if i != 24 else 8, j if j != 24 else 8, :9]
            assert np.allclose(a, b, rtol=5.0000001e-6, atol=5e-13), (i, j)
    coarse = tmp_path / "coarse.d"
    spec16, _, _ = run_driver(16, 3, 1e-3, 0.01, "random", "FP64", "fused", f"save={coarse}")
    spec32, _, _ = run_driver(32, 0, 1e-3, 0.01, f"file={coarse}", "FP64", "fused", "file_reduced=16,16,16")
    for k in range(0, 6):
        assert abs(spec32[k] - spec16[k]) <= 1e-10 * max(spec16.values()), k
    assert all(spec32[k] == 0.0 for k in spec32 if k >= 10)


def test_driver_output_files_match_oracle(tmp_path):
    """Output_global / Output_shell_spectrum / Output_flux / Output_shell_to_shell (output_energy.cc:43-81,
    368-394, output_ET.cc:45-74, 220-254): glob.d, spectrum.d, flux.d and shell.d in the reference's column
    layout and stream formats (glob.d fixed with 12 decimals, the others 6 significant digits), every number
    equal to the oracle's value of the same quantity to the printed digits."""
    N, steps, dt, nu = 32, 3, 1e-3, 0.01
    (tmp_path / "out").mkdir()
    spec, flux, glob = run_driver(N, steps, dt, nu, "random", "FP64", "fused", f"out={tmp_path}")
    o = Oracle((N, N, N))
    V = o.ic_random(1)
    o.rk4_ns(V, nu, dt, steps=steps)

    lines = (tmp_path / "out" / "glob.d").read_text().splitlines()
    assert len(lines) == steps
    t, rest = lines[-1].split("\t")
    assert t == f"{steps * dt:.12f}"
    g = rest.split("    ")
    assert len(g) == 9, lines[-1]
    assert all(len(x.split(".")[1]) == 12 for x in rest.split()), rest
    E, D, Dn = map(float, g[0].split(" "))
    H1, H2 = map(float, g[1].split(" "))
    DH1, DH2 = map(float, g[2].split(" "))
    S, kmax_eta, nu_out, Rl, tdt = float(g[3]), float(g[4]), float(g[5]), float(g[6]), float(g[7])
    m1, m2, m3 = map(float, g[8].split(" "))
    assert nu_out == nu and tdt == dt and (m1, m2, m3) == (0.0, 0.0, 0.0)
    assert f12(E, o.total_energy(V)) and f12(D, o.total_dissipation(V)) and f12(Dn, nu * o.total_dissipation(V))
    h = o.helicity(V)
    for a, b in zip((H1, H2, DH1, DH2), h):
        assert f12(a, b, np.abs(h).max())
    assert f12(S, o.entropy(V))
    kmax = o.max_radius_inside()
    Dn_ref = nu * o.total_dissipation(V)
    assert f12(kmax_eta, kmax * (nu ** 3 / Dn_ref) ** 0.25) and f12(Rl, 2 * o.total_energy(V) * (15 / (nu * Dn_ref)) ** 0.5)

    blocks = (tmp_path / "out" / "spectrum.d").read_text().split("\n")
    assert blocks[0] == f"%% Time = {steps * dt:g}"
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
            assert g6(a, b, sc), (i, a, b)
        assert float(f[4]) == 0.0

    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    fo = o.flux_ns(V, radii)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    assert len(frows) == len(flux)
    for r in frows:
        i, R, pi, force = r.split()
        assert float(R) == radii[int(i)] and g6(float(pi), fo[int(i)], np.abs(fo).max()) and float(force) == 0.0

    # shell.d: Output_prefix, the shell radii 1..no_shells, then T(n, m) rows 1..no_shells-1, columns 1..no_shells
    head, rad, blocks = read_shell_d(tmp_path / "out" / "shell.d")
    assert head[1] == f"%% Grid size {N} {N} {N}" and head[2] == "%% Basis, Integration scheme:  FOUR RK4"
    assert head[3] == f"%% Tinit, Tfinal, Tdt_fixed:  0    {steps * dt:g}    {dt:g}"
    assert rad == list(radii[1:])
    assert len(blocks) == 1 and abs(blocks[0][0] - steps * dt) < 1e-9 and list(blocks[0][1]) == ["U to U"]
    Tref = o.shell_tr_ns(V, radii)
    nsh = len(radii) - 1
    Tm = blocks[0][1]["U to U"]
    assert Tm.shape == (nsh - 1, nsh)
    sc = np.abs(Tref).max()
    assert sc > 0
    for n in range(1, nsh):
        for m in range(1, nsh + 1):
            assert g6(Tm[n - 1, m - 1], Tref[n, m], sc), (n, m, Tm[n - 1, m - 1], Tref[n, m])


def read_blocks(path):
    """A ring_to_ring.d-like file: '%% Time = t', then '%% NAME: ' lines each followed by a Blitz++ array."""
    lines = path.read_text().split("\n")
    out, i = [], 0
    while i < len(lines):
        if lines[i].startswith("%% Time = "):
            out.append((float(lines[i][10:]), {}))
            i += 1
        elif lines[i].startswith("%% ") and lines[i].endswith(": "):
            name = lines[i][3:-2]
            a, i = blitz_2d(lines, i + 1)
            out[-1][1][name] = a
        else:
            i += 1
    return out


@pytest.mark.parametrize("physics", ["ns", "mhd"])
def test_driver_ring_to_ring_files(tmp_path, physics):
    """ring=NRING,NSECT,NCYL,NSLAB,AXIS,SCHEME with out=DIR: Init_ring_tr (unit radii, sectors over [0, pi/2],
    slabs over [0, Anis_max_Kpll]) and Output_ring_to_ring / Output_cylinder_ring_to_ring (output_ET.cc:427-760):
    every table of ring_to_ring.d and cyl_ring_to_ring.d (U to U; for MHD also W to W, U to W and the Elsasser
    pair) equals the oracle's ring transfer over giver and receiver rings 1..n-1 x 1..m to the printed digits;
    the force feeds of the unforced run are 0."""
    N, steps, dt, nu, eta = 32, 2, 1e-3, 0.01, 0.02
    (tmp_path / "out").mkdir()
    extra = [f"mhd={eta}"] if physics == "mhd" else []
    run_driver(N, steps, dt, nu, "random", "FP64", "fused", f"out={tmp_path}", "ring=5,3,5,3,1,1", *extra)
    o = Oracle((N, N, N))
    U, B = o.ic_random(1), o.ic_random(2)
    if physics == "ns":
        o.rk4_ns(U, nu, dt, steps=steps)
    else:
        o.rk4_mhd(U, B, nu, eta, dt, steps=steps)
    radii = np.array([0.0, 1, 2, 3, 4, 10000.0])
    sect = np.array([0.0, np.pi / 6, np.pi / 3, np.pi / 2])
    kp = (N // 3) * 1.0
    slabs = np.array([0.0, kp / 3, 2 * kp / 3, kp])
    Zp = [u + b for u, b in zip(U, B)]
    Zm = [u - b for u, b in zip(U, B)]
    for fname, cyl, edges in (("ring_to_ring.d", False, sect), ("cyl_ring_to_ring.d", True, slabs)):
        blocks = read_blocks(tmp_path / "out" / fname)
        assert len(blocks) == 1 and abs(blocks[0][0] - steps * dt) < 1e-9
        got = blocks[0][1]
        refs = {"U to U": o.ring_tr(U, U, U, radii, edges, cyl=cyl, axis=1)}
        if physics == "mhd":
            refs["W to W"] = o.ring_tr(U, B, B, radii, edges, cyl=cyl, axis=1)
            refs["U to W"] = o.ring_tr(B, U, B, radii, edges, sign=1.0, cyl=cyl, axis=1)
            refs["Zp to Zp"] = o.ring_tr(Zm, Zp, Zp, radii, edges, cyl=cyl, axis=1)
            refs["Zm to Zm"] = o.ring_tr(Zp, Zm, Zm, radii, edges, cyl=cyl, axis=1)
        for name, ref in refs.items():
            t = got[name]
            sub = ref[1:5, 1:4, 1:5, 1:4]
            assert t.shape == (4, 3, 4, 3) and np.abs(sub).max() > 0
            sc = np.abs(sub).max()
            for a, b in zip(t.ravel(), sub.ravel()):
                assert g6(a, b, sc), (fname, name, a, b)
        feeds = [k for k in got if k.startswith("Force feed")]
        assert feeds and all(np.all(got[k] == 0) and got[k].shape == (4, 3) for k in feeds)


def test_driver_et_scheme0_radii_and_flux():
    """EnergyTr with shell_input_scheme 0 (EnergyTr.cc:136-160): radii 0, 2, 4, 8, log-spaced, kmax/2, kmax,
    INF_RADIUS for no_shells = 8 at 64^3 (kmax_inside 21 > 16), equal to the pinned oracle radii, and
    Compute_flux over those spheres equal to the oracle's."""
    N, steps, dt, nu = 64, 2, 1e-3, 0.01
    import subprocess as sp
    out = sp.run([DRV] + [str(a) for a in (N, steps, dt, nu, "random", "FP64", "fused", "et=8,0")],
                 capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rad = [float(l.split()[2]) for l in out.stdout.splitlines() if l.startswith("radius")]
    flux = [float(l.split()[2]) for l in out.stdout.splitlines() if l.startswith("flux")]
    o = Oracle((N, N, N))
    ref = o.et_radii(0, 8)
    assert len(rad) == 9 and np.abs(np.array(rad) - ref).max() <= 1e-15 * 10000
    V = o.ic_random(1)
    o.rk4_ns(V, nu, dt, steps=steps)
    fo = o.flux_ns(V, np.array(ref))
    assert len(flux) == 8
    assert np.abs(np.array(flux) - fo[1:]).max() <= 1e-10 * np.abs(fo).max()


@pytest.mark.parametrize("physics", ["ns", "mhd"])
def test_driver_get_dt_energy_branch(tmp_path, physics):
    """Get_dt with the energy branch winning (compute_dt.cc:47-59, MHD :75-98): Tdt_fixed = 1 is far above
    2 pi / (kmax_inside 20 sqrt(E)), so every step's dt comes from the energy of the field it starts from.
    The tdt column of glob.d, step by step, and the energies equal the oracle run that takes its dt from
    the same formula (orc_dt, pinned by its closed form; for MHD E = E_u + E_b and B(0,0,0) = 0)."""
    N, steps, dt_fixed, nu, eta = 32, 3, 1.0, 0.01, 0.02
    (tmp_path / "out").mkdir()
    extra = [f"mhd={eta}"] if physics == "mhd" else []
    spec, flux, glob = run_driver(N, steps, dt_fixed, nu, "random", "FP64", "fused", f"out={tmp_path}", *extra)
    o = Oracle((N, N, N))
    V, B = o.ic_random(1), o.ic_random(2)
    kmax = o.max_radius_inside()
    dts, E = [], []
    for _ in range(steps):
        if physics == "ns":
            dts.append(o.dt(V, dt_fixed))
            o.rk4_ns(V, nu, dts[-1])
            E.append(o.total_energy(V))
        else:
            assert all(b[0, 0, 0] == 0 for b in B)
            dts.append(min(2 * np.pi / kmax / (20 * np.sqrt(o.total_energy(V) + o.total_energy(B))), dt_fixed))
            o.rk4_mhd(V, B, nu, eta, dts[-1])
            E.append(o.total_energy(V))
    assert max(dts) < 0.1 * dt_fixed
    lines = (tmp_path / "out" / "glob.d").read_text().splitlines()
    assert len(lines) == steps
    for i, line in enumerate(lines):
        if physics == "ns":
            tdt = float(line.split("\t")[1].split("    ")[7])
        else:
            tdt = float(line.split("\t")[1].split()[22])
        assert f12(tdt, dts[i], rel=1e-11), (i, tdt, dts[i])
        assert f12(float(line.split("\t")[0]), sum(dts[:i + 1]), rel=1e-11)
    assert np.abs(glob[:, 1] - np.array(E)).max() <= 1e-12 * max(E)


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


@pytest.mark.parametrize("N", [32, 64])
@pytest.mark.parametrize("path", ["fused", "plain"])
def test_driver_forced_output_columns(tmp_path, path, N):
    """Forced run with out=DIR: the force-spectrum column of spectrum.d (Compute_force_shell_spectrum
    force_spectrum.cc:48-53, 2 sum_c Re V_c.F_c* per shell) and the force-feed column of flux.d
    (Compute_force_feed_shell compute_force_feed.cc:70-76, Shell_mult_all(F, V) over shell_radius).  The
    reference writes them after Time_advance, whose stages recompute the force (time_advance.cc:100-155), so
    F is the force of the last RK stage's field, V the new velocity; the oracle's forced step reports that
    field.  At N = 64 the fused state steps chained and keeps that field from its stage-3 x-forward."""
    steps, dt, nu = 3, 1e-3, 0.01
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
        assert g6(float(r.split("\t")[4]), fk[i], np.abs(fk).max()), (i, r)
    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    feed = o.shell_mult_all(F, V, radii)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    assert np.abs(feed).max() > 0
    for r in frows:
        i, R, pi, force = r.split()
        assert g6(float(force), feed[int(i)], np.abs(feed).max()), r
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
        assert g6(a, fku[i], sc) and g6(b, fkw[i], sc), (i, r)
    radii = np.concatenate([np.arange(0, N // 3 + 2, dtype=float), [10000.0]])
    feu, few = o.shell_mult_all(FU, U, radii), o.shell_mult_all(FW, B, radii)
    frows = [r for r in (tmp_path / "out" / "flux.d").read_text().split("\n")[1:] if r.strip()]
    sc = np.abs(feu).max()
    assert sc > 0 and len(frows) > 0
    for r in frows:
        t = r.split()
        assert len(t) == 13, r
        i = int(t[0])
        assert g6(float(t[11]), feu[i], sc) and g6(float(t[12]), few[i], sc), r
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

    line = (tmp_path / "out" / "glob.d").read_text().splitlines()[-1]
    t, mid, mb = line.split("\t")
    assert mid.count("    ") == 9 and [float(x) for x in mb.split(" ")] == [0.0, 0.0, 0.0]
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
        assert f12(a, b, sc=max(abs(b), 1e-12 * max(abs(x) for x in ref[7:15]))), (i, a, b)
    assert t == f"{steps * dt:.12f}"

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
            assert g6(a, b, sc), (i, j, a, b)

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
                assert g6(float(tok), fl[name][i], np.abs(fl[name]).max()), (r, name)
    # shell.d (Output_shell_to_shell(W) output_ET.cc:293-352): U to U, W to W, U to W, Zp to Zp, Zm to Zm
    head, rad, blocks = read_shell_d(tmp_path / "out" / "shell.d")
    assert rad == list(radii[1:]) and len(blocks) == 1
    m = o.shell_tr_mhd(U, B, radii)
    names = {"U to U": "uu", "W to W": "ww", "U to W": "uw", "Zp to Zp": "zp", "Zm to Zm": "zm"}
    assert list(blocks[0][1]) == list(names)
    nsh = len(radii) - 1
    for name, key in names.items():
        Tm, ref = blocks[0][1][name], m[key]
        assert Tm.shape == (nsh - 1, nsh) and np.abs(ref).max() > 0
        for n in range(1, nsh):
            for k in range(1, nsh + 1):
                assert g6(Tm[n - 1, k - 1], ref[n, k], np.abs(ref).max()), (name, n, k)


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
        assert f12(a, b, sc), (i, a, b)

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
            assert g6(a, col[i], sc), (i, j, a, col[i])
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
            assert g6(float(tok), col[i], np.abs(col).max()), (r, tok, col[i])

    # shell.d (Output_shell_to_shell(T) output_ET.cc:258-290): U to U and T to T
    head, rad, blocks = read_shell_d(tmp_path / "out" / "shell.d")
    assert rad == list(radii[1:]) and len(blocks) == 1 and list(blocks[0][1]) == ["U to U", "T to T"]
    Tu, Tt = o.shell_tr_scalar(V, T, radii)
    nsh = len(radii) - 1
    for name, ref in (("U to U", Tu), ("T to T", Tt)):
        Tm = blocks[0][1][name]
        assert Tm.shape == (nsh - 1, nsh) and np.abs(ref).max() > 0
        for n in range(1, nsh):
            for k in range(1, nsh + 1):
                assert g6(Tm[n - 1, k - 1], ref[n, k], np.abs(ref).max()), (name, n, k)
