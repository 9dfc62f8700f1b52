"""Pins the oracle's nonlinear terms, projection, IF-RK4, spectra, energy transfer, MHD, RB and
initial conditions (CPU only) with brute-force triad sums, closed forms and invariants."""
import math

import numpy as np
import pytest

from pyoracle import Oracle


def _kept_mask(o):
    N1, N2, N3 = o.N
    m = np.zeros(o.shape, bool)
    for l1 in range(N1):
        for l2 in range(N2):
            for l3 in range(N3 // 2 + 1):
                m[l1, l2, l3] = bool(o.kept(l1, l2, l3))
    return m


def _kvec(o):
    N1, N2, N3 = o.N
    kx = np.array([o.kx(l) for l in range(N1)], float) * o.g.f1
    ky = np.array([o.ky(l) for l in range(N2)], float) * o.g.f2
    kz = np.arange(N3 // 2 + 1, dtype=float) * o.g.f3
    return np.meshgrid(kx, ky, kz, indexing="ij")


def test_nlin_ns_matches_bruteforce_triads(golden_dir):
    """IncVF::Compute_nlin (compute_nlin.cc:58-108) equals the exact truncated convolution
    sum_j i k_j sum_{p+q=k} u_i(p) u_j(q) on kept modes (Orszag 1971) and is 0 elsewhere."""
    G = np.load(f"{golden_dir}/nlin_triad.npz")
    N = tuple(int(v) for v in G["N"])
    o = Oracle(N)
    V = [np.ascontiguousarray(G["V"][i]) for i in range(3)]
    nl = o.nlin_ns(V)
    mask = _kept_mask(o)
    for i in range(3):
        ref = G["nlin"][i]
        assert np.abs(nl[i][mask] - ref[mask]).max() < 1e-12 * np.abs(ref).max()
        assert np.abs(nl[i][~mask]).max() == 0.0


def test_et_nlin_with_helper_equal_giver_is_ns_nlin():
    """EnergyTr_Compute_nlin (ET_compute_nlin.cc:85-120) with G = H = U reproduces the NS
    nonlinear term on kept modes (same products, same derivative placement)."""
    o = Oracle((16, 16, 16))
    V = o.ic_random(3)
    nl = o.nlin_ns([v.copy() for v in V])
    et = o.et_nlin(V, V)
    mask = _kept_mask(o)
    for i in range(3):
        assert np.abs(et[i][mask] - nl[i][mask]).max() < 1e-14 * np.abs(nl[i]).max()


def test_projection_divergence_free_and_pressure_closed_form():
    """Compute_pressure + Compute_rhs (compute_pressure.cc:54-62, compute_rhs.cc:52-60):
    K.rhs = 0 and rhs = -(N - K (K.N)/K^2)."""
    o = Oracle((16, 16, 16), kfactor=(1.0, 2.0, 0.5))
    rng = np.random.default_rng(3)
    nl = [o.dealias(rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)) for _ in range(3)]
    orig = [a.copy() for a in nl]
    p = o.pressure(nl)
    o.rhs(nl, p)
    KX, KY, KZ = _kvec(o)
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K2[0, 0, 0] = 1
    kn = KX * orig[0] + KY * orig[1] + KZ * orig[2]
    for i, Ki in enumerate((KX, KY, KZ)):
        ref = -(orig[i] - Ki * kn / K2)
        ref[0, 0, 0] = -orig[i][0, 0, 0]
        assert np.abs(nl[i] - ref).max() < 1e-13 * np.abs(ref).max()
    div = KX * nl[0] + KY * nl[1] + KZ * nl[2]
    assert np.abs(div).max() < 1e-13 * np.abs(nl[0]).max()


def test_rk4_exact_viscous_decay_when_nonlinearity_vanishes():
    """A single divergence-free shear wave has (u.grad)u = 0, so the IF-RK4 step
    (time_advance.cc:100-155) must give E(t) = E0 exp(-2 nu K^2 t) to roundoff."""
    o = Oracle((16, 16, 16))
    V = o.zeros(3)
    V[0][0, 3, 2] = 0.3 + 0.1j          # k = (0,3,2), u along x: k.u = 0
    nu, dt = 0.05, 0.01
    E0 = o.total_energy(V)
    o.rk4_ns(V, nu, dt, steps=10)
    expect = E0 * math.exp(-2 * nu * 13 * dt * 10)
    assert abs(o.total_energy(V) - expect) < 1e-14 * expect


def test_rk4_fourth_order_convergence_taylor_green():
    """Halving dt reduces the RK4 error by ~16 on Taylor-Green (nonlinear)."""
    o = Oracle((16, 16, 16))
    nu, T = 0.01, 0.4
    def run(dt):
        V = o.ic_taylor_green(1, 4.0)
        o.rk4_ns(V, nu, dt, steps=int(round(T / dt)))
        return V
    ref = run(0.0125)
    e1 = max(np.abs(a - b).max() for a, b in zip(run(0.1), ref))
    e2 = max(np.abs(a - b).max() for a, b in zip(run(0.05), ref))
    assert 12 < e1 / e2 < 20


def test_taylor_green_closed_forms():
    """ic_model.cc:58-107 + four_energy.cc: E = 1/8, E(2) = 1/8, dissipation 2 sum K^2|u|^2/2 = 3/4,
    and Parseval with the real-space field u = (sin x cos y cos z, -cos x sin y cos z, 0)."""
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_taylor_green(1, 1.0)
    assert abs(o.total_energy(V) - 0.125) < 1e-16
    assert abs(o.total_dissipation(V) - 0.75) < 1e-15
    Ek = o.vector_spectrum(V)
    assert abs(Ek[2] - 0.125) < 1e-16 and abs(Ek.sum() - 0.125) < 1e-16
    x = np.arange(16) * 2 * np.pi / 16
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = [np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z)]
    for i in range(2):
        A = V[i].copy()
        o.inverse(A)
        got = A.view(np.float64).reshape(16, 16, -1)[:, :, :16]
        assert np.abs(got - u[i]).max() < 1e-15


def test_spectrum_parseval_random_field():
    """Sum of unit-shell E(k) below kmax_inside plus the unrescaled shells equals the total energy,
    which equals the real-space mean of |u|^2/2 (Parseval)."""
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(7, E0=0.5)
    E = o.total_energy(V)
    assert abs(E - 0.5) < 1e-14
    raw = np.zeros(o.shell_ek_size())
    for a in V:
        Sk, cnt = o.shell_spectrum(a)
        raw += Sk
    # shells above kmax_inside are rescaled by Approx_number_modes_in_shell / count; undoing that gives
    # the plain binned energy, which sums to the total
    kin = o.max_radius_inside()
    assert raw[kin + 1:].sum() > 0
    plain = raw.copy()
    for k in range(kin + 1, len(raw)):
        if cnt[k] >= 1:
            plain[k] = raw[k] * cnt[k] / o.approx_modes(k)
    assert abs(plain.sum() - E) < 1e-14
    usq = 0.0
    for a in V:
        b = a.copy()
        o.inverse(b)
        usq += (b.view(np.float64).reshape(16, 16, -1)[:, :, :16] ** 2).mean()
    assert abs(usq / 2 - E) < 1e-14


def test_random_ic_properties():
    """Seeded IC: divergence free, Hermitian on kz = 0, dealiased, E = E0 and each unit shell
    follows E(k) proportional to k^4 exp(-2 (k/4)^2)."""
    o = Oracle((24, 24, 24))
    V = o.ic_random(1, E0=0.5, kp=4.0)
    KX, KY, KZ = _kvec(o)
    div = KX * V[0] + KY * V[1] + KZ * V[2]
    assert np.abs(div).max() < 1e-15
    mask = _kept_mask(o)
    for a in V:
        assert np.abs(a[~mask]).max() == 0
        b = a.copy()
        o.satisfy_reality(b)
        o.dealias(b)
        assert np.array_equal(a, b)
    assert abs(o.total_energy(V) - 0.5) < 1e-14
    raw = np.zeros(o.shell_ek_size())
    for a in V:
        raw += o.shell_spectrum(a)[0]
    kin = o.max_radius_inside()
    ks = np.arange(1, kin + 1)
    model = ks ** 4 * np.exp(-2 * (ks / 4.0) ** 2)
    ratio = raw[1:kin + 1] / model
    assert np.abs(ratio / ratio[0] - 1).max() < 1e-12


def test_shell_to_shell_antisymmetry_and_flux_identity():
    """Compute_shell_tr / Compute_flux (compute_shell_tr.cc:59-79, compute_flux.cc:52-69):
    T(n,m) = -T(m,n) for a divergence-free U, sum T = 0, Pi(inf) = 0, and with all modes inside
    kmax_inside Pi(R_i) = sum_{n<=i} sum_{m>i} T(n,m)."""
    o = Oracle((16, 16, 16))
    V = o.ic_random(5, E0=0.5, kp=2.0)
    kin = o.max_radius_inside()
    KX, KY, KZ = _kvec(o)
    Kmag = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    for a in V:
        a[Kmag > kin] = 0
    nsh = kin + 2
    radii = np.concatenate([np.arange(nsh), [10000.0]]).astype(float)
    T = o.shell_tr_ns(V, radii)
    Tin = T[1:nsh, 1:nsh]
    assert np.abs(Tin + Tin.T).max() < 1e-13 * np.abs(Tin).max()
    assert abs(T.sum()) < 1e-13 * np.abs(T).max()
    flux = o.flux_ns(V, radii)
    assert abs(flux[nsh]) < 1e-15
    for i in range(1, nsh - 1):
        pi_T = T[1:i + 1, i + 1:].sum()
        assert abs(flux[i] - pi_T) < 1e-12 * np.abs(flux).max()


def test_et_radii_schemes():
    """EnergyTr.cc:136-170: scheme 0 (0,2,4,8, log-spaced, kmax/2, kmax, INF) and unit radii."""
    o = Oracle((64, 64, 64))
    r = o.et_radii(0, 8)
    mri = o.max_radius_inside()
    s = math.log2(mri / 16.0) / 3
    ref = [0, 2, 4, 8, 8 * 2 ** s, 8 * 2 ** (2 * s), mri / 2, mri, 10000]
    assert np.allclose(r, ref, rtol=0, atol=1e-12)
    r = o.et_radii(1, 32)
    assert list(r[:32]) == list(range(32)) and r[32] == 10000


def test_mhd_elsasser_reductions():
    """compute_nlin_VF_MHD.cc: with B = 0 the U-equation term equals the NS nonlinear term and the
    B-term vanishes; with B = U (Z- = 0) both vanish."""
    o = Oracle((16, 16, 16))
    U = o.ic_random(2)
    B = o.zeros(3)
    nU, nB = o.nlin_mhd([u.copy() for u in U], [b.copy() for b in B])
    nl = o.nlin_ns([u.copy() for u in U])
    for i in range(3):
        assert np.abs(nU[i] - nl[i]).max() < 1e-14 * np.abs(nl[i]).max()
        assert np.abs(nB[i]).max() < 1e-16
    nU, nB = o.nlin_mhd([u.copy() for u in U], [u.copy() for u in U])
    assert max(np.abs(a).max() for a in nU + nB) < 1e-16


def test_mhd_flux_reductions_and_role_symmetry():
    """IncVF::Compute_flux(IncVF& W) (compute_flux.cc:148-262): with B = 0 the U<->U> flux and both
    Elsasser fluxes equal the NS flux and every U-B term vanishes; with B = U the self fluxes equal
    the NS flux, U<->B> = B<->U> = -flux and the Elsasser fluxes vanish (Z- = 0); swapping U and B
    exchanges U<->B and B<->U (which fixes helper, giver and receiver of each term)."""
    o = Oracle((16, 16, 16))
    U = o.ic_random(2)
    B = o.ic_random(3)
    radii = np.array([0.0, 1, 2, 3, 4, 5, 7, 10000.0])
    fns = o.flux_ns(U, radii)
    sc = np.abs(fns).max()
    f = o.flux_mhd(U, o.zeros(3), radii, elsasser=True)
    for k in ("u2u", "zp", "zm"):
        assert np.abs(f[k] - fns).max() < 1e-13 * sc, k
    for k in ("b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in", "u_out_b_out"):
        assert np.abs(f[k]).max() < 1e-16, k
    f = o.flux_mhd(U, [u.copy() for u in U], radii, elsasser=True)
    for k in ("u2u", "b2b"):
        assert np.abs(f[k] - fns).max() < 1e-13 * sc, k
    for k in ("u2b_in_out", "b2u_in_out"):
        assert np.abs(f[k] + fns).max() < 1e-13 * sc, k
    assert np.abs(f["u2b_in_in"] - f["b2u_in_in"]).max() < 1e-13 * sc
    for k in ("zp", "zm"):
        assert np.abs(f[k]).max() < 1e-16, k
    a = o.flux_mhd(U, B, radii)
    b = o.flux_mhd(B, U, radii)
    s2 = max(np.abs(a[k]).max() for k in a)
    assert np.abs(a["u2b_in_out"] - b["b2u_in_out"]).max() < 1e-13 * s2
    assert np.abs(a["u2b_in_in"] - b["b2u_in_in"]).max() < 1e-13 * s2
    assert np.abs(a["b2u_in_out"]).max() > 1e-6 * s2


def test_mhd_ideal_energy_conservation():
    """Ideal MHD (nu = eta = 0): one IF-RK4 step conserves E_u + E_b to O(dt^5)."""
    o = Oracle((16, 16, 16))
    U = o.ic_random(2, E0=0.5)
    B = o.ic_random(3, E0=0.5)
    E0 = o.total_energy(U) + o.total_energy(B)
    errs = []
    for dt in (2e-2, 1e-2):
        u = [a.copy() for a in U]
        b = [a.copy() for a in B]
        o.rk4_mhd(u, b, 0.0, 0.0, dt)
        errs.append(abs(o.total_energy(u) + o.total_energy(b) - E0))
    assert errs[1] < 1e-7 and errs[0] / errs[1] > 16


def test_rb_conduction_state_and_nusselt():
    """Get_Nusselt_no (nusselt.cc:53-89): Nu = 1 at V = 0; the conduction state stays at rest."""
    N = (16, 16, 16)
    o = Oracle(N, kfactor=(np.pi, np.pi / np.sqrt(2), np.pi / np.sqrt(2)), basis="SCFT")
    V = o.zeros(3)
    T = o.zeros()
    assert o.nusselt(V, T, 1e5, 1.0) == 1.0
    o.rk4_rb(V, T, 1e5, 1.0, 1e-3, steps=2)
    assert max(np.abs(a).max() for a in V + [T]) == 0.0


def test_rb_linear_growth_rate():
    """Free-slip RB, PRLARGE/ULARGE scaling (RB_slip_main.cc:238-239): a small single mode theta ~
    sin(pi x) e^{i k y} grows at sigma = -(nu+kappa)K^2/2 + sqrt(((nu-kappa)K^2/2)^2 + k_perp^2/K^2)."""
    f = (np.pi, np.pi / np.sqrt(2), np.pi / np.sqrt(2))
    N = (16, 16, 16)
    o = Oracle(N, kfactor=f, basis="SCFT")
    Ra, Pr = 1e4, 1.0
    nu, kap = math.sqrt(Pr / Ra), 1 / math.sqrt(Pr * Ra)
    V = o.zeros(3)
    T = o.zeros()
    T[1, 1, 0] = 1e-12
    T[1, N[1] - 1, 0] = 1e-12           # Hermitian partner on kz = 0
    kperp2 = f[1] ** 2
    K2 = f[0] ** 2 + kperp2
    a = (nu + kap) * K2 / 2
    sigma = -a + math.sqrt(((nu - kap) * K2 / 2) ** 2 + kperp2 / K2)
    dt, steps = 0.02, 1500             # the decaying branch dies out as exp(-1.15 t)
    o.rk4_rb(V, T, Ra, Pr, dt, steps=steps)
    E1 = o.energy(T)
    o.rk4_rb(V, T, Ra, Pr, dt, steps=100)
    E2 = o.energy(T)
    rate = math.log(E2 / E1) / (2 * 100 * dt)
    assert abs(rate - sigma) < 1e-6 * abs(sigma) + 1e-9


def _kmag_grid(o):
    N1, N2, H = o.shape
    K = np.empty((N1, N2, H))
    for l1 in range(N1):
        for l2 in range(N2):
            for l3 in range(H):
                K[l1, l2, l3] = o.kmag(l1, l2, l3)
    return K


@pytest.mark.parametrize("basis", ["SCFT", "FOUR"])
def test_scalar_transfer_antisymmetry_conservation_and_flux_identity(basis):
    """Compute_shell_tr_scalar / Compute_flux_scalar (compute_shell_tr.cc:98-117, compute_flux.cc:89-141,
    the RB drivers for PRLARGE/ULARGE): with U divergence-free, T_SF(n,m) = -T_SF(m,n), sum T_SF = 0, and
    with every mode inside kmax_inside Pi_SF(R_i) = sum_{n<=i} sum_{m>i} T_SF(n,m); the velocity rows are
    Compute_flux / Compute_shell_tr.  On SCFT the scalar transfer of the whole field is the scalar part
    of the RB nonlinear term (compute_nlin_SF.cc:80-177)."""
    N = (16, 16, 16)
    if basis == "SCFT":
        f = (np.pi, np.pi / np.sqrt(2), np.pi / np.sqrt(2))
        o = Oracle(N, kfactor=f, basis="SCFT")
        V = o.zeros(3)
        T = o.dealias(o.ic_random_scalar(4, 0.5))
        o.rk4_rb(V, T, 1e5, 1.0, 1e-2, steps=3)
    else:
        o = Oracle(N)
        V = o.ic_random(5, E0=0.5, kp=2.0)
        T = o.dealias(o.ic_random_scalar(6, 0.5))
    kin = o.max_radius_inside()
    K = _kmag_grid(o)
    for a in V + [T]:
        a[K > kin] = 0
    assert np.abs(V[0]).max() > 0 and np.abs(T).max() > 0
    nsh = kin + 2
    radii = np.concatenate([np.arange(nsh), [10000.0]]).astype(float)
    Ts, Tt = o.shell_tr_scalar(V, T, radii)
    assert np.array_equal(Ts, o.shell_tr_ns(V, radii))
    Tin = Tt[1:nsh, 1:nsh]
    scale = np.abs(Tin).max()
    assert scale > 0
    assert np.abs(Tin + Tin.T).max() < 1e-12 * scale
    assert abs(Tt.sum()) < 1e-12 * scale
    fs, ft = o.flux_scalar(V, T, radii)
    assert np.array_equal(fs, o.flux_ns(V, radii))
    assert ft[nsh] == 0.0
    for i in range(1, nsh - 1):
        assert abs(ft[i] - Tt[1:i + 1, i + 1:].sum()) < 1e-12 * np.abs(ft).max()
    if basis == "SCFT":
        ou = Oracle(N, kfactor=f, basis="SCFT", dealias=False)
        _, nT = ou.nlin_rb(V, T)
        assert np.abs(ou.et_nlin_scalar(V, T) - nT).max() <= 1e-14 * np.abs(nT).max()


@pytest.mark.parametrize("axis,s", [(0, 1), (1, -1), (2, 1), (2, -1)])
def test_helicity_of_a_helical_mode(axis, s):
    """Compute_total_helicity_FOUR (four_energy.cc:453-506): a single helical mode V(k) = A (e1 + i s e2),
    e1 x e2 = k/|k|, has H1 = s |k| E, H2 = H1 / k^2, DH1 = 2 k^2 H1, DH2 = 2 H1 (E in the reference's
    convention), and its shell helicity sits in shell |k| on the k component only."""
    N, m, A = (16, 16, 16), 3, 0.7
    o = Oracle(N)
    V = o.zeros(3)
    e = np.eye(3)
    e1, e2 = e[(axis + 1) % 3], e[(axis + 2) % 3]
    vec = A * (e1 + 1j * s * e2)
    idx = [0, 0, 0]
    idx[axis] = m
    for c in range(3):
        V[c][tuple(idx)] = vec[c]
    if axis != 2:                       # kz = 0: the conjugate partner at -k is stored too
        jdx = [0, 0, 0]
        jdx[axis] = N[axis] - m
        for c in range(3):
            V[c][tuple(jdx)] = np.conj(vec[c])
    E = o.total_energy(V)
    H1, H2, DH1, DH2 = o.helicity(V)
    assert abs(H1 - s * m * E) < 1e-14 and abs(E - 2 * A * A) < 1e-14
    assert abs(H2 - H1 / m ** 2) < 1e-15 and abs(DH1 - 2 * m ** 2 * H1) < 1e-12 and abs(DH2 - 2 * H1) < 1e-14
    Hk = o.shell_helicity(V)
    assert abs(Hk[axis][m] - H1) < 1e-14
    assert abs(np.abs(Hk).sum() - abs(H1)) < 1e-14


def test_shell_helicity_sums_to_total_and_entropy_closed_form():
    """Shell helicity (four_energy.cc:551-671) sums over shells and components to H1 for a field inside
    kmax_inside; the entropy (four_energy.cc:687-732) of n stored kz > 0 modes of equal energy is
    log2(2n) (each mode and its conjugate carry probability 1/(2n)); a random field's entropy lies
    below log2 of twice its stored mode count."""
    o = Oracle((16, 16, 16))
    V = o.ic_random(8, E0=0.5, kp=2.0)
    kin = o.max_radius_inside()
    K = _kmag_grid(o)
    for a in V:
        a[K > kin] = 0
    H1 = o.helicity(V)[0]
    Hk = o.shell_helicity(V)
    assert abs(Hk.sum() - H1) < 1e-13 * max(1.0, abs(H1))
    W = o.zeros(3)
    n = 7
    for j in range(n):
        W[j % 3][j + 1, j, 1 + j] = (0.3 + 0.4j) * (1 if j % 2 else -1)   # off the ky = N2/2 plane (multiplicity 2)
    assert abs(o.entropy(W) - math.log2(2 * n)) < 1e-13
    nmodes = int(np.count_nonzero(np.abs(V[0]) + np.abs(V[1]) + np.abs(V[2])))
    assert 0 < o.entropy(V) < math.log2(2 * nmodes)


def test_scft_helicity_is_the_positive_kx_half_of_a_fourier_field():
    """Compute_total_helicity_SCFT / Compute_shell_spectrum_helicity_SCFT (scft_energy.cc:448-677) evaluate the
    sin/cos field through its e^{+i kx x} half: V1 sin(kx x) contributes -i V1 / 2 there and V2, V3 cos(kx x)
    contribute V / 2.  A FOUR field holding only those +kx modes therefore has H1, H2, DH1, DH2 and the shell
    helicities (shells <= kmax_inside) in the same ratio to the SCFT ones as its energy to the SCFT energy.
    (With the -kx half added the doubled field is mirror-symmetric and its FOUR helicity vanishes.)"""
    N1, N2, N3 = 16, 16, 16
    os_ = Oracle((N1, N2, N3), basis="SCFT")
    of = Oracle((2 * N1, N2, N3))
    rng = np.random.default_rng(1)
    W = os_.zeros(3)
    for c in range(3):
        W[c][...] = rng.standard_normal(os_.shape) + 1j * rng.standard_normal(os_.shape)
        W[c][0] = 0
        os_.satisfy_reality(W[c])
        os_.dealias(W[c])
    G, F = of.zeros(3), of.zeros(3)
    for l1 in range(1, N1):
        for X in (G, F):
            X[0][l1], X[1][l1], X[2][l1] = -0.5j * W[0][l1], 0.5 * W[1][l1], 0.5 * W[2][l1]
        F[0][2 * N1 - l1], F[1][2 * N1 - l1], F[2][2 * N1 - l1] = 0.5j * W[0][l1], 0.5 * W[1][l1], 0.5 * W[2][l1]
    ratio = sum(of.energy(G[c]) for c in range(3)) / sum(os_.energy(W[c]) for c in range(3))
    hs, hg = os_.helicity(W), of.helicity(G)
    assert np.abs(hg - ratio * hs).max() < 1e-13 * np.abs(hs).max()
    kin = os_.max_radius_inside()
    Hs, Hg = np.asarray(os_.shell_helicity(W)), np.asarray(of.shell_helicity(G))
    assert np.abs(Hg[:, :kin + 1] - ratio * Hs[:, :kin + 1]).max() < 1e-13 * np.abs(Hs).max()
    assert abs(of.helicity(F)[0]) < 1e-13 * abs(hs[0])


def test_scft_entropy_closed_form():
    """Get_entropy_SCFT (scft_energy.cc:679-748): n stored modes with kx > 0, kz > 0 and equal |V|^2 have
    Modal_energy_SCFT = |V|^2 and weight mult = 2, so S = -2 n p log2 p with p = |V|^2 / E."""
    o = Oracle((16, 16, 16), basis="SCFT")
    W = o.zeros(3)
    n = 5
    for j in range(n):
        W[j % 3][j + 1, j, 1 + j] = 0.4 - 0.3j
    E = sum(o.energy(W[c]) for c in range(3))
    p = 0.25 / E
    assert abs(o.entropy(W) - (-2 * n * p * math.log2(p))) < 1e-13


def test_entropy_scalar_closed_forms():
    """Get_entropy_scalar_FOUR / _SCFT (four_energy.cc:761-826, scft_energy.cc:750-800): one field with n stored
    kz > 0 modes of equal |A|^2 has entropy log2(2n) on a FOUR box (each mode and its conjugate at 1/(2n)) and
    -2 n p log2 p, p = |A|^2 / E, on a SCFT box (Modal_energy_SCFT = |A|^2 for kx > 0, weight mult = 2)."""
    n = 6
    for basis in ("FOUR", "SCFT"):
        o = Oracle((16, 16, 16), basis=basis)
        A = o.zeros()
        for j in range(n):
            A[j + 1, j, 1 + j] = 0.3 + 0.2j * (1 if j % 2 else -1)
        S = o.entropy_scalar(A)
        if basis == "FOUR":
            assert abs(S - math.log2(2 * n)) < 1e-13
        else:
            p = abs(0.3 + 0.2j) ** 2 / o.energy(A)
            assert abs(S - (-2 * n * p * math.log2(p))) < 1e-13
