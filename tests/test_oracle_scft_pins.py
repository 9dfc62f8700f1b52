"""Pins the oracle's sin/cos (SCFT) energy normalisation, Nusselt number and shell mode counts, and its time
step rule, in absolute terms (CPU only):
  * Get_total_energy_SCFT (scft_energy.cc:51-158) equals half the real-space mean of (u - mean u)^2 on the
    DCT grid, with the real-space field built by scipy's DCT-III / DST-III (FFTW REDFT01 / RODFT01 by their
    published definitions) and numpy's irfft2 in (y, z), not by the oracle's own transforms;
  * Get_Nusselt_no (nusselt.cc:53-89, PRLARGE / ULARGE) equals 1 + sqrt(Ra Pr) <u_x theta> with V != 0;
  * the SCFT shell counts (Compute_shell_spectrum_SCFT, weight Multiplicity_factor / 2) equal a quarter of
    the number of lattice points k in Z^3 with ceil|K| = shell, K = (kx f1, ky f2, kz f3);
  * Get_dt (compute_dt.cc:47-59) returns min(2 pi / (kmax_inside 20 sqrt(E)), Tdt_fixed)."""
import math

import numpy as np
import pytest
import scipy.fft as sf

from pyoracle import Oracle

RB_KF = (np.pi, np.pi / np.sqrt(2), np.pi / np.sqrt(2))


def _scft_field(o, rng, parity):
    """A valid sin/cos field: a real random grid function taken to spectral space by scipy's DCT-II / DST-II and
    numpy's rfft2 with Tarang's scaling 1 / (2 N1 N2 N3) (scft_tr.cc:89-219), band-limited to kx < 2 N1 / 3,
    |ky| < N2 / 3, kz < N3 / 3 (inside Dealias_array's kept set and symmetric in ky, so the kz = 0 plane stays
    Hermitian)."""
    N1, N2, N3 = o.N
    u = rng.standard_normal(o.N)
    t = sf.dct(u, type=2, axis=0) if parity == 0 else sf.dst(u, type=2, axis=0)
    A = np.fft.rfftn(t, axes=(1, 2)) / (2 * N1 * N2 * N3)
    if parity == 1:
        A = np.concatenate([np.zeros_like(A[:1]), A[:-1]], axis=0)
    kx = np.arange(N1).reshape(-1, 1, 1)
    ky = np.fft.fftfreq(N2, 1.0 / N2).reshape(1, -1, 1)
    kz = np.arange(N3 // 2 + 1).reshape(1, 1, -1)
    keep = (kx < 2 * N1 / 3) & (np.abs(ky) < N2 / 3) & (kz < N3 / 3)
    return np.ascontiguousarray(A * keep)


def _real_space(A, N, parity):
    """Tarang's inverse SCFT transform through scipy / numpy: rows of a sine field hold sin(l1 x), so the
    DST-III input is the array shifted down by one row."""
    N1, N2, N3 = N
    D = A if parity == 0 else np.concatenate([A[1:], np.zeros_like(A[:1])], axis=0)
    r = np.fft.irfftn(D, s=(N2, N3), axes=(1, 2)) * (N2 * N3)
    return sf.dct(r, type=3, axis=0) if parity == 0 else sf.dst(r, type=3, axis=0)


@pytest.mark.parametrize("N,kf", [((16, 16, 16), (1.0, 1.0, 1.0)), ((12, 8, 16), RB_KF)])
@pytest.mark.parametrize("parity", [0, 1])
def test_scft_energy_is_half_the_real_space_variance(N, kf, parity):
    o = Oracle(N, kfactor=kf, basis="SCFT")
    rng = np.random.default_rng(31 + parity)
    A = _scft_field(o, rng, parity)
    u = _real_space(A, N, parity)
    # the cosine field's mean is its k = 0 mode, which Get_total_energy_SCFT leaves out; a sine field has
    # no k = 0 mode (its grid mean is not zero: sin(k x) of odd k does not average out on [0, pi])
    ref = 0.5 * ((u - (u.mean() if parity == 0 else 0.0)) ** 2).mean()
    assert abs(o.energy(A) - ref) < 1e-13 * ref
    if parity == 0:
        assert abs(u.mean() - A[0, 0, 0].real) < 1e-13 * np.abs(u).max()


@pytest.mark.parametrize("N", [(16, 16, 16), (12, 8, 16)])
def test_scft_cross_energy_and_nusselt_with_flow(N):
    """Nu = 1 + 2 sqrt(Ra Pr) Get_total_energy(V1, T) = 1 + sqrt(Ra Pr) <u_x theta>, V1 and theta both sine
    fields (free-slip, fixed temperature walls), and Get_total_energy(A, B) = <a b> / 2 for cosine fields
    with their k = 0 modes left out."""
    o = Oracle(N, kfactor=RB_KF, basis="SCFT")
    rng = np.random.default_rng(41)
    V = [_scft_field(o, rng, 1), _scft_field(o, rng, 0), _scft_field(o, rng, 0)]
    T = _scft_field(o, rng, 1)
    ux, th = _real_space(V[0], N, 1), _real_space(T, N, 1)
    Ra, Pr = 1e7, 1.0
    ref = 1 + math.sqrt(Ra * Pr) * (ux * th).mean()
    assert abs(ref - 1) > 0.1
    assert abs(o.nusselt(V, T, Ra, Pr) - ref) < 1e-12 * abs(ref)
    a, b = _real_space(V[1], N, 0), _real_space(V[2], N, 0)
    cross = 0.5 * ((a - a.mean()) * (b - b.mean())).mean()
    assert abs(o.energy(V[1], V[2]) - cross) < 1e-13 * abs(o.energy(V[1]))


@pytest.mark.parametrize("N,kf", [((16, 16, 16), (1.0, 1.0, 1.0)), ((24, 16, 20), RB_KF),
                                  ((12, 12, 12), (1.0, 2.0, 0.5))])
def test_scft_shell_counts_are_a_quarter_of_the_lattice(N, kf):
    N1, N2, N3 = N
    o = Oracle(N, kfactor=kf, basis="SCFT")
    A = o.zeros() + 1.0
    _, cnt = o.shell_spectrum(A)
    kx = np.arange(-(N1 - 1), N1)
    ky = np.arange(-(N2 // 2), N2 // 2 + 1)
    kz = np.arange(-(N3 // 2), N3 // 2 + 1)
    X, Y, Z = np.meshgrid(kx * kf[0], ky * kf[1], kz * kf[2], indexing="ij")
    idx = np.ceil(np.sqrt(X * X + Y * Y + Z * Z)).astype(int)
    ref = np.bincount(idx.ravel(), minlength=len(cnt))[:len(cnt)] / 4.0
    assert np.array_equal(cnt, ref)


def test_four_shell_counts_are_half_the_lattice():
    """The Fourier counterpart: Multiplicity_factor_FOUR sums to half the lattice points of the shell
    (kz = +-N3/2 and the Nyquist planes included once each; the corner modes lie beyond the last shell)."""
    N = (24, 16, 20)
    N1, N2, N3 = N
    o = Oracle(N, kfactor=(1.0, 1.5, 0.75))
    _, cnt = o.shell_spectrum(o.zeros() + 1.0)
    kx = np.arange(-(N1 // 2), N1 // 2 + 1)
    ky = np.arange(-(N2 // 2), N2 // 2 + 1)
    kz = np.arange(-(N3 // 2), N3 // 2 + 1)
    X, Y, Z = np.meshgrid(kx * 1.0, ky * 1.5, kz * 0.75, indexing="ij")
    idx = np.ceil(np.sqrt(X * X + Y * Y + Z * Z)).astype(int)
    ref = np.bincount(idx.ravel(), minlength=len(cnt))[:len(cnt)] / 2.0
    assert np.array_equal(cnt, ref)


@pytest.mark.parametrize("basis", ["FOUR", "SCFT"])
def test_get_dt_energy_branch_closed_form(basis):
    """Taylor-Green (E = 1/8 on a 2 pi box, four_energy / scft_energy normalisation) at 24^3: kmax_inside =
    8, dt = (2 pi / 8) / (20 sqrt(1/8)); a larger Tdt_fixed leaves it, a smaller one wins."""
    o = Oracle((24, 24, 24))
    V = o.ic_taylor_green(1, 1.0)
    assert o.max_radius_inside() == 8
    dt = (2 * math.pi / 8) / (20 * math.sqrt(0.125))
    assert abs(o.dt(V, 1.0) - dt) < 1e-15
    assert o.dt(V, 1e-3) == 1e-3
    if basis == "SCFT":
        os_ = Oracle((24, 24, 24), basis="SCFT")
        rng = np.random.default_rng(5)
        W = [_scft_field(os_, rng, 1), _scft_field(os_, rng, 0), _scft_field(os_, rng, 0)]
        E = 0.5 * (_real_space(W[0], (24, 24, 24), 1) ** 2).mean()
        E += sum(0.5 * ((u - u.mean()) ** 2).mean() for u in
                 (_real_space(W[1], (24, 24, 24), 0), _real_space(W[2], (24, 24, 24), 0)))
        kin = os_.max_radius_inside()
        assert abs(os_.dt(W, 10.0) - (2 * math.pi / kin) / (20 * math.sqrt(E))) < 1e-13 * os_.dt(W, 10.0)
