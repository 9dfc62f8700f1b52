"""Pins the oracle's transforms, spectral operators and binning (CPU only).

Each test checks the oracle against something that does not come from the oracle's own code:
numpy/scipy golden vectors (tests/golden/), the reference's recorded FFTW output
(trunk/test/test-fftw/out1.d), closed forms, or brute-force integer arithmetic.
"""
import math

import numpy as np
import pytest

from pyoracle import Oracle, lib


def _real_view(A, N3):
    return A.view(np.float64).reshape(A.shape[0], A.shape[1], -1)[:, :, :N3]


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 16, 24, 32, 48, 64, 96, 100, 128, 256, 512])
def test_fft1d_matches_numpy(n):
    L = lib()
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = x.copy()
    L.orc_fft1d(n, -1, y.ctypes.data)
    z = x.copy()
    L.orc_fft1d(n, +1, z.ctypes.data)
    ref = np.fft.fft(x)
    assert np.abs(y - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.abs(z - np.fft.ifft(x) * n).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_four_transforms_vs_numpy_golden(golden_dir, tag):
    """ArrayFFT_FOUR / ArrayIFFT_FOUR (four_tr.cc:199-295) == numpy rfftn/N^3 and irfftn*N^3."""
    G = np.load(f"{golden_dir}/four_fft.npz")
    N = tuple(int(v) for v in G[f"{tag}_N"])
    o = Oracle(N, dealias=False)
    A = o.zeros()
    Af = A.view(np.float64).reshape(N[0], N[1], -1)
    Af[:, :, :N[2]] = G[f"{tag}_u"]
    Af[:, :, N[2]:] = 123.0                   # padding content must not matter
    o.forward(A)
    assert np.abs(A - G[f"{tag}_fwd"]).max() < 1e-15 * max(1, np.abs(G[f"{tag}_fwd"]).max()) * 10
    B = G[f"{tag}_spec"].copy()               # non-Hermitian: c2r semantics of FFTW/numpy
    o.inverse(B)
    ref = G[f"{tag}_inv"]
    assert np.abs(_real_view(B, N[2]) - ref).max() < 1e-13 * np.abs(ref).max()
    assert np.all(B[:, :, N[2] // 2] == 0)    # Zero_pad_last_plane_FOUR (four_tr.cc:293)


def test_kat_out1(golden_dir):
    """trunk/test/test-fftw/out1.d: unit Nyquist-z mode at 4^3.  After c2r the data carry
    sum|u|^2 = 80 - 16 (padding share) = 64; after r2c FFTW gives (64,0), i.e. 1 after Norm_FOUR."""
    K = np.load(f"{golden_dir}/kat_out1.npz")
    N = tuple(int(v) for v in K["N"])
    o = Oracle(N, dealias=False)
    A = o.zeros()
    A[tuple(K["mode"])] = 1.0
    o.inverse(A)
    data_sq = float((_real_view(A, N[2]) ** 2).sum())
    assert data_sq == float(K["tot_after_c2r_incl_padding"]) - float(K["padding_share"])
    o.forward(A)
    expect = complex(*K["r2c_value"]) / (N[0] * N[1] * N[2])
    assert abs(A[tuple(K["mode"])] - expect) < 1e-15
    A[tuple(K["mode"])] = 0
    assert np.abs(A).max() < 1e-15


@pytest.mark.parametrize("tag", ["a", "b"])
@pytest.mark.parametrize("parity", [0, 1])
def test_scft_transforms_vs_scipy_golden(golden_dir, tag, parity):
    """ArraySFT/CFT_SCFT and inverses (scft_tr.cc:230-639) == scipy dst/dct types 2/3."""
    G = np.load(f"{golden_dir}/scft_fft.npz")
    N = tuple(int(v) for v in G[f"{tag}_N"])
    o = Oracle(N, basis="SCFT", dealias=False)
    A = o.zeros()
    _ = A.view(np.float64).reshape(N[0], N[1], -1)
    A.view(np.float64).reshape(N[0], N[1], -1)[:, :, :N[2]] = G[f"{tag}_u"]
    o.forward(A, parity)
    ref = G[f"{tag}_fwd{parity}"]
    assert np.abs(A - ref).max() < 1e-14 * np.abs(ref).max()
    B = G[f"{tag}_spec{parity}"].copy()
    o.inverse(B, parity)
    ref = G[f"{tag}_inv{parity}"]
    assert np.abs(_real_view(B, N[2]) - ref).max() < 1e-13 * np.abs(ref).max()


def test_four_derivative_closed_form():
    """Xderiv/Yderiv/Zderiv_FOUR (four_tr.cc:314-353): d/dx e^{ik.x} = i k f e^{ik.x}."""
    N = (8, 12, 10)
    f = (1.0, 0.5, 2.0)
    o = Oracle(N, kfactor=f, dealias=False)
    x = np.arange(N[0]) * 2 * np.pi / N[0] / f[0]
    y = np.arange(N[1]) * 2 * np.pi / N[1] / f[1]
    z = np.arange(N[2]) * 2 * np.pi / N[2] / f[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    k = (2, -3, 4)
    u = np.cos(k[0] * f[0] * X + k[1] * f[1] * Y + k[2] * f[2] * Z)
    A = o.zeros()
    A.view(np.float64).reshape(N[0], N[1], -1)[:, :, :N[2]] = u
    o.forward(A)
    for axis in (1, 2, 3):
        B = o.deriv(axis, A)
        o.inverse(B)
        du = -k[axis - 1] * f[axis - 1] * np.sin(k[0] * f[0] * X + k[1] * f[1] * Y + k[2] * f[2] * Z)
        assert np.abs(_real_view(B, N[2]) - du).max() < 1e-12


@pytest.mark.parametrize("parity", [0, 1])
def test_scft_xderiv_closed_form(parity):
    """Xderiv_Sin/Cos_SCFT (scft_tr.cc:649-672): d/dx of a sine (cosine) series evaluated at the
    DST-III / DCT-III sample points x_j = (j+1/2) pi / (N1 f1)."""
    N = (12, 4, 4)
    f1 = np.pi
    o = Oracle(N, kfactor=(f1, 1, 1), basis="SCFT", dealias=False)
    rng = np.random.default_rng(5)
    A = o.zeros()
    coef = rng.standard_normal(N[0])
    if parity == 1:
        coef[0] = 0
    A[:, 0, 0] = coef
    xj = (np.arange(N[0]) + 0.5) * np.pi / N[0]            # f1 * x
    m = np.arange(N[0])
    B = o.deriv(1, A, parity)
    Bi = B.copy()
    o.inverse(Bi, 1 - parity)
    got = _real_view(Bi, N[2])[:, 0, 0]
    if parity == 1:    # sine series -> cosine series with +k f
        expect = (coef[None, :] * m[None, :] * f1 * np.cos(np.outer(xj, m))).sum(1)
    else:              # cosine series -> sine series with -k f
        expect = -(coef[None, :] * m[None, :] * f1 * np.sin(np.outer(xj, m))).sum(1)
    # the reference inverse uses weight 1 on the m=0 cosine term and 2 elsewhere (REDFT01)
    w = np.where(m == 0, 1.0, 2.0) if parity == 1 else np.where(m == N[0], 1.0, 2.0)
    if parity == 1:
        expect = (w[None, :] * coef[None, :] * m[None, :] * f1 * np.cos(np.outer(xj, m))).sum(1)
    else:
        expect = -(w[None, :] * coef[None, :] * m[None, :] * f1 * np.sin(np.outer(xj, m))).sum(1)
    assert np.abs(got - expect).max() < 1e-11 * max(1, np.abs(expect).max())


@pytest.mark.parametrize("N", [(64, 64, 64), (96, 96, 96), (128, 96, 64), (512, 512, 512)])
def test_dealias_mask_formula(N):
    """Dealias_array universal_misc.cc:118-159: the kept counts per axis follow the integer formulas
    (|kx| < N1/3; l2 outside [N2/3, 2N2/3+1]; l3 < N3/3)."""
    o = Oracle(N)
    N1, N2, N3 = N
    kx_kept = sum(1 for l1 in range(N1) if o.kept(l1, 0, 0))
    ky_kept = sum(1 for l2 in range(N2) if o.kept(0, l2, 0))
    kz_kept = sum(1 for l3 in range(N3 // 2 + 1) if o.kept(0, 0, l3))
    assert kx_kept == 2 * (N1 // 3) - 1
    assert ky_kept == N2 - (2 * N2 // 3 + 1 - N2 // 3 + 1)
    assert kz_kept == N3 // 3


@pytest.mark.parametrize("N", [(32, 32, 32), (24, 16, 20)])
def test_shell_binning_bruteforce_integer(N):
    """index = ceil(Kmagnitude) (four_energy.cc:276) equals the integer ceil(sqrt(k^2)) for
    kfactor = 1, for every stored mode (bit-exact binning target)."""
    o = Oracle(N)
    N1, N2, N3 = N
    for l1 in range(N1):
        kx = l1 if l1 <= N1 // 2 else l1 - N1
        for l2 in range(N2):
            ky = l2 if l2 <= N2 // 2 else l2 - N2
            for l3 in range(N3 // 2 + 1):
                k2 = kx * kx + ky * ky + l3 * l3
                r = math.isqrt(k2)
                ic = r if r * r == k2 else r + 1
                assert math.ceil(o.kmag(l1, l2, l3)) == ic


def test_shell_counts_and_rescale_closed_form():
    """Compute_shell_spectrum_FOUR (four_energy.cc:291-338): shell counts equal brute-force sums
    of Multiplicity_factor_FOUR; partial shells above kmax_inside are rescaled by
    2 pi k^2 / count (Approx_number_modes_in_shell_FOUR four_inline.h:276-306)."""
    N = (24, 24, 24)
    o = Oracle(N)
    A = o.zeros() + 1.0                                     # |A|^2 = 1 everywhere
    Sk, cnt = o.shell_spectrum(A)
    size = o.shell_ek_size()
    assert size == o.min_radius_outside() + 1 == int(math.ceil(math.sqrt(3 * 8 * 8))) + 1
    ref = np.zeros(size)
    for l1 in range(N[0]):
        kx = l1 if l1 <= 12 else l1 - 24
        for l2 in range(N[1]):
            ky = l2 if l2 <= 12 else l2 - 24
            for l3 in range(13):
                k2 = kx * kx + ky * ky + l3 * l3
                r = math.isqrt(k2)
                idx = r if r * r == k2 else r + 1
                if idx < size:
                    f = 1.0 if l3 > 0 else 0.5
                    if kx == 12 or ky == 12:
                        f *= 2
                    ref[idx] += f
    assert np.array_equal(cnt, ref)
    kin = o.max_radius_inside()
    assert kin == 8
    for k in range(size):
        if k > kin and cnt[k] >= 1:
            assert abs(Sk[k] - 2 * np.pi * k * k) < 1e-9 * Sk[k]
        else:
            assert Sk[k] == ref[k]
