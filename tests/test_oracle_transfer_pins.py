"""Pins the oracle's energy fluxes and shell-to-shell transfers in absolute terms (sign and scale) against
explicit triad sums over the full wavenumber lattice (tests/golden/et_triad.npz, written by
tests/golden/make_golden.py without FFTs or multiplicity factors), and against the energy budget of the
pinned nonlinear term: the flux out of a sphere equals the rate at which the nonlinear term removes
energy from inside it.  CPU only."""
import numpy as np
import pytest

from pyoracle import Oracle


@pytest.fixture(scope="module")
def triad(golden_dir):
    G = np.load(f"{golden_dir}/et_triad.npz")
    o = Oracle(tuple(int(v) for v in G["N"]))
    U = [np.ascontiguousarray(G["U"][i]) for i in range(3)]
    B = [np.ascontiguousarray(G["B"][i]) for i in range(3)]
    T = np.ascontiguousarray(G["T"])
    return G, o, U, B, T


def _close(got, ref, rel=1e-12):
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(np.asarray(got) - ref).max() <= rel * scale, (got, ref)


def test_kin_of_fixture(triad):
    G, o, *_ = triad
    assert o.max_radius_inside() == int(G["kin"])


def test_flux_ns_triads(triad):
    """Compute_flux (compute_flux.cc:52-69): flux_self(i) = -sum_{|k|>r_i} Re(N[U, U<] . U*), the factor
    2 x Multiplicity_factor of Shell_mult_single (four_ET.cc:49-80) counting each lattice mode once."""
    G, o, U, B, T = triad
    _close(o.flux_ns(U, G["radii"])[1:], G["flux_ns"][1:])


def test_shell_tr_ns_triads(triad):
    """Compute_shell_tr (compute_shell_tr.cc:59-79): row m = giver shell, column n = receiver shell of
    Shell_mult_all (receivers |k| <= kmax_inside only), with the minus sign."""
    G, o, U, B, T = triad
    Tm = o.shell_tr_ns(U, G["radii"])
    _close(Tm, G["shell_tr_ns"])


def test_flux_mhd_triads(triad):
    """Compute_flux(IncVF& W) (compute_flux.cc:148-262): every one of the nine fluxes with its own helper,
    giver, receiver and sign, the Elsasser pair with Z+- = U +- B (elsasser.cc:53-60)."""
    G, o, U, B, T = triad
    f = o.flux_mhd(U, B, G["radii"], elsasser=True)
    for k in ("u2u", "b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in", "u_out_b_out", "zp", "zm"):
        _close(f[k][1:], G[k][1:])


def test_shell_tr_mhd_triads(triad):
    """Compute_shell_tr(IncVF& W) (compute_shell_tr.cc:154-260): U to U as Compute_shell_tr, W to W with
    helper U and the minus sign, U to W with helper W and the plus sign, and the Elsasser pair with
    Z+- = U +- B, each helper the other variable."""
    G, o, U, B, T = triad
    m = o.shell_tr_mhd(U, B, G["radii"])
    _close(m["uu"], G["shell_tr_ns"])
    for k in ("ww", "uw", "zp", "zm"):
        _close(m[k], G["shell_tr_mhd_" + k])


def test_flux_and_shell_tr_scalar_triads(triad):
    """Compute_flux_scalar / Compute_shell_tr_scalar (compute_flux.cc:89-111, compute_shell_tr.cc:98-117):
    flux_SF(i) = -sum_{|k|>r_i} Re(N[U, T<] T*), T_SF(m, n) = -sum_{k in shell n} Re(N[U, T_m] T*)."""
    G, o, U, B, T = triad
    fs, ft = o.flux_scalar(U, T, G["radii"])
    _close(fs[1:], G["flux_ns"][1:])
    _close(ft[1:], G["flux_sf"][1:])
    Ts, Tt = o.shell_tr_scalar(U, T, G["radii"])
    _close(Ts, G["shell_tr_ns"])
    _close(Tt, G["shell_tr_sf"])


def test_flux_is_the_energy_loss_inside_the_sphere():
    """Energy budget of the pinned nonlinear term (compute_nlin.cc:58-108): with du/dt = -N + ..., the
    energy inside radius R changes by -sum_{|k|<=R} Re(U* . N) through the nonlinear term, and
    Compute_flux's Pi(R) is exactly that loss (triad antisymmetry moves it across R).  A dropped factor
    or a flipped sign of the flux fails this.  The lattice sum over the half spectrum counts kz > 0 modes
    twice (their conjugates) and kz = 0 modes once; that weight is what makes total_energy Parseval's
    real-space mean of |u|^2 / 2 (checked here too)."""
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(21, E0=0.5, kp=3.0)
    nl = o.nlin_ns([v.copy() for v in V])
    w = np.full(o.shape, 2.0)
    w[:, :, 0] = 1.0
    usq = 0.0
    for a in V:
        b = a.copy()
        o.inverse(b)
        usq += (b.view(np.float64).reshape(N[0], N[1], -1)[:, :, :N[2]] ** 2).mean()
    assert abs(0.5 * sum((w * np.abs(a) ** 2).sum() for a in V) - usq / 2) < 1e-14
    K = np.array([[[o.kmag(l1, l2, l3) for l3 in range(o.shape[2])] for l2 in range(N[1])]
                  for l1 in range(N[0])])
    radii = np.array([0.0, 1, 2, 3, 4, 5, 6, 10000.0])
    flux = o.flux_ns(V, radii)
    loss = [sum((w * (np.conj(a) * n).real)[(K > 0) & (K <= r)].sum() for a, n in zip(V, nl)) for r in radii]
    scale = np.abs(flux).max()
    assert scale > 1e-4
    for i in range(1, len(radii)):
        assert abs(flux[i] - loss[i]) < 1e-12 * scale, (i, flux[i], loss[i])
