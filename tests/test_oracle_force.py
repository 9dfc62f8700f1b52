"""Pins the oracle's energy/helicity forcing (force_field_proc 1 and 2 on a FOUR box) with a brute-force
mode count, the energy-injection identity, Beltrami modes (i K x V = s |K| V) and the modal energy and
helicity of the forced field (CPU only).

Reference: Compute_force_const_energy_helicity_supply compute_force_ek_hk_supply.cc:51-157,
Compute_force_const_energy_helicity compute_force_const_ek_hk.cc:51-140, Force_alpha_beta
force_alpha_beta.cc:60-89, Get_local_number_modes_in_shell_FOUR four_basic.cc:48-90."""
import itertools
import math

import numpy as np
import pytest

from pyoracle import Oracle


def _kgrid(o):
    N1, N2, N3 = o.N
    kx = np.array([o.kx(l) for l in range(N1)], float) * o.g.f1
    ky = np.array([o.ky(l) for l in range(N2)], float) * o.g.f2
    kz = np.arange(N3 // 2 + 1, dtype=float) * o.g.f3
    return np.meshgrid(kx, ky, kz, indexing="ij")


def _weights(o):
    """Full-sphere weight of each stored mode: kz = 0 modes are stored with their conjugates, kz > 0
    modes stand for two."""
    w = np.full(o.shape, 2.0)
    w[:, :, 0] = 1.0
    return w


def _project(o, V):
    KX, KY, KZ = _kgrid(o)
    K2 = KX * KX + KY * KY + KZ * KZ
    K2[0, 0, 0] = 1.0
    d = (KX * V[0] + KY * V[1] + KZ * V[2]) / K2
    return [V[0] - KX * d, V[1] - KY * d, V[2] - KZ * d]


def _real_coefficient_field(o, seed):
    """Solenoidal field with real Fourier coefficients: Vr x Vi = 0, so every mode has zero helicity."""
    rng = np.random.default_rng(seed)
    V = [rng.standard_normal(o.shape) for _ in range(3)]
    V = _project(o, V)
    out = o.zeros(3)
    for c in range(3):
        out[c][...] = V[c]
        o.satisfy_reality(out[c])
        o.dealias(out[c])
    return out


def _curl(o, V):
    KX, KY, KZ = _kgrid(o)
    return [1j * (KY * V[2] - KZ * V[1]), 1j * (KZ * V[0] - KX * V[2]), 1j * (KX * V[1] - KY * V[0])]


@pytest.mark.parametrize("kf,inner,outer", [((1, 1, 1), 0.0, 1.5), ((1, 1, 1), 1.0, 2.0), ((1, 1, 1), 2.5, 3.5),
                                             ((1.0, 0.5, 2.0), 1.2, 2.7), ((0.7, 1.3, 1.0), 0.0, 3.0)])
def test_mode_count_is_the_full_sphere_count(kf, inner, outer):
    """nf counts every wavevector (conjugates included) with inner <= |K| < outer; brute force over Z^3."""
    o = Oracle((32, 32, 32), kfactor=kf)
    m = int(math.ceil(outer / min(kf))) + 1
    n = 0
    for a, b, c in itertools.product(range(-m, m + 1), repeat=3):
        k = math.sqrt((a * kf[0]) ** 2 + (b * kf[1]) ** 2 + (c * kf[2]) ** 2)
        n += inner <= k < outer
    assert o.force_supply_modes(inner, outer) == n
    if kf == (1, 1, 1) and (inner, outer) == (0.0, 1.5):
        assert n == 19                    # |K|^2 in {0, 1, 2}: 1 + 6 + 12
    if kf == (1, 1, 1) and (inner, outer) == (1.0, 2.0):
        assert n == 26                    # |K|^2 in {1, 2, 3}: 6 + 12 + 8


@pytest.mark.parametrize("r", [0.0, 0.5])
def test_supply_injects_the_energy_supply(r):
    """Proc 1 on a field whose modes carry no helicity: alpha = eps / (2 nf E_k), so sum_k Re F.V* = eps;
    the vorticity term (beta = r alpha) adds helicity injection sum_k Re F.(iKxV)* = sum_k beta |K x V|^2
    and no energy.  F vanishes outside inner < |K| <= outer."""
    o = Oracle((16, 16, 16))
    V = _real_coefficient_field(o, 5)
    inner, outer, eps = 2.5, 3.5, 0.37
    F = o.force(V, 1, inner, outer, eps, r)
    KX, KY, KZ = _kgrid(o)
    K = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    band = (K > inner) & (K <= outer)
    w = _weights(o)
    inj = sum((w * (F[c] * np.conj(V[c])).real).sum() for c in range(3))
    assert abs(inj - eps) < 1e-13 * eps
    for c in range(3):
        assert np.abs(F[c][~band]).max() == 0.0
    nf = o.force_supply_modes(inner, outer)
    E = 0.5 * sum(np.abs(V[c]) ** 2 for c in range(3))
    W = _curl(o, V)
    for c in range(3):
        ref = np.where(band, eps / nf / (2 * np.where(band, E, 1)) * (V[c] + r * W[c]), 0)
        assert np.abs(F[c] - ref).max() < 1e-13 * np.abs(ref).max()
    hinj = sum((w * (F[c] * np.conj(W[c])).real).sum() for c in range(3))
    href = (w * np.where(band, eps / nf / (2 * np.where(band, E, 1)) * r, 0) *
            sum(np.abs(W[c]) ** 2 for c in range(3))).sum()
    assert abs(hinj - href) < 1e-12 * max(1.0, abs(href))


def _helical_modes(o, A=0.6):
    """A positive-helicity mode along z at kz = 3 and a negative-helicity mode along x at kx = +-3:
    V = A (e1 + i s e2), e1 x e2 = K/|K|, so i K x V = s |K| V."""
    V = o.zeros(3)
    V[0][0, 0, 3], V[1][0, 0, 3] = A, 1j * A                     # s = +1, K = 3 z
    V[1][3, 0, 0], V[2][3, 0, 0] = A, -1j * A                    # s = -1, K = 3 x
    V[1][-3, 0, 0], V[2][-3, 0, 0] = A, 1j * A                   # its conjugate partner at -K
    return V


def test_beltrami_modes_take_the_degenerate_branches():
    """|sk| = 1.  Proc 1: alpha = eps_k / (4 E), beta = alpha / (sk |K|); for s = +1 the vorticity term
    doubles alpha V, for s = -1 beta < 0 and the vorticity is dropped (beta > MYEPS only).  Proc 2:
    alpha = sqrt(e_k / E), beta = 0 for either sign."""
    o = Oracle((16, 16, 16))
    V = _helical_modes(o)
    inner, outer, eps = 2.5, 3.5, 0.2
    nf = o.force_supply_modes(inner, outer)
    E = 0.6 ** 2                                                  # |V|^2 / 2
    F = o.force(V, 1, inner, outer, eps, 0.0)
    t = eps / nf / (2 * E)
    for c in range(3):
        assert np.allclose(F[c][0, 0, 3], t * V[c][0, 0, 3], rtol=1e-14, atol=0)
        assert np.allclose(F[c][3, 0, 0], 0.5 * t * V[c][3, 0, 0], rtol=1e-14, atol=0)
        assert np.allclose(F[c][-3, 0, 0], 0.5 * t * V[c][-3, 0, 0], rtol=1e-14, atol=0)
    F2 = o.force(V, 2, inner, outer, eps, 0.3)
    a = math.sqrt(eps / nf / E)
    for c in range(3):
        for idx in [(0, 0, 3), (3, 0, 0), (-3, 0, 0)]:
            assert abs(F2[c][idx] - a * V[c][idx]) <= 1e-14 * abs(a * 0.6)
    nz = sum(np.count_nonzero(F2[c]) for c in range(3))
    assert nz == 6


@pytest.mark.parametrize("h", [0.0, 0.4, 0.9])
def test_constant_level_sets_modal_energy_and_relative_helicity(h):
    """Proc 2 on modes with zero helicity: |F|^2 / 2 = e_k and K.(Fr x Fi) / (|K| |F|^2 / 2) = h on every
    forced mode (alpha^2 + beta^2 K^2 = e_k / E, 2 alpha beta K^2 E = h |K| e_k)."""
    o = Oracle((16, 16, 16))
    V = _real_coefficient_field(o, 9)
    inner, outer, e = 1.5, 3.2, 0.8
    F = o.force(V, 2, inner, outer, e, h)
    nf = o.force_supply_modes(inner, outer)
    KX, KY, KZ = _kgrid(o)
    K = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    band = (K > inner) & (K <= outer)
    EF = 0.5 * sum(np.abs(F[c]) ** 2 for c in range(3))
    assert np.abs(EF[band] - e / nf).max() < 1e-13 * e / nf
    Fr, Fi = [F[c].real for c in range(3)], [F[c].imag for c in range(3)]
    HF = (KX * (Fr[1] * Fi[2] - Fr[2] * Fi[1]) + KY * (Fr[2] * Fi[0] - Fr[0] * Fi[2]) +
          KZ * (Fr[0] * Fi[1] - Fr[1] * Fi[0]))
    rel = HF[band] / (K[band] * EF[band])
    assert np.abs(rel - h).max() < 1e-12


def test_constant_level_negative_helicity_drops_the_vorticity():
    """Proc 2 with h < 0 gives beta < 0, which Force_alpha_beta ignores: F = alpha V with
    alpha = sqrt(e_k / E) (sqrt(1 + h) + sqrt(1 - h)) / 2, so F carries no helicity."""
    o = Oracle((16, 16, 16))
    V = _real_coefficient_field(o, 11)
    inner, outer, e, h = 1.5, 3.2, 0.8, -0.3
    F = o.force(V, 2, inner, outer, e, h)
    nf = o.force_supply_modes(inner, outer)
    KX, KY, KZ = _kgrid(o)
    K = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    band = (K > inner) & (K <= outer)
    E = 0.5 * sum(np.abs(V[c]) ** 2 for c in range(3))
    alpha = np.sqrt(e / nf / np.where(band, E, 1)) * (math.sqrt(1 + h) + math.sqrt(1 - h)) / 2
    for c in range(3):
        ref = np.where(band, alpha * V[c], 0)
        assert np.abs(F[c] - ref).max() < 1e-13 * np.abs(ref).max()


def test_zero_field_gets_no_force():
    """Modes with modal energy <= MYEPS are not forced."""
    o = Oracle((16, 16, 16))
    F = o.force(o.zeros(3), 1, 1.5, 3.5, 1.0, 0.2)
    assert all(np.abs(f).max() == 0.0 for f in F)
    with pytest.raises(ValueError):
        o.force(o.zeros(3), 1, 0.1, 0.5, 1.0, 0.0)                 # no wavevector in [0.1, 0.5)


def test_forced_step_injects_energy_at_the_supply_rate():
    """One IF-RK4 step with the supply force from a field whose modes carry no helicity (injection
    exactly eps at t): E_forced(t+dt) - E_free(t+dt) = eps dt + O(dt^2), the remainder halving with dt."""
    o = Oracle((16, 16, 16))
    nu, eps = 0.01, 0.5
    V0 = _real_coefficient_field(o, 4)
    s = math.sqrt(0.5 / o.total_energy(V0))
    V0 = [v * s for v in V0]
    res = []
    for dt in (2e-3, 1e-3):
        Vf = o.step_ns_forced([v.copy() for v in V0], nu, dt, 1, 1.5, 3.5, eps, 0.0)
        Vu = o.rk4_ns([v.copy() for v in V0], nu, dt)
        res.append(abs((o.total_energy(Vf) - o.total_energy(Vu)) / dt - eps))
    assert res[1] < 2e-4 * eps
    assert 0.4 < res[1] / res[0] < 0.6


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_forced_step_with_zero_supply_is_the_free_step(scheme):
    """energy_supply = 0 gives F = 0, so the forced EULER / RK2 / RK4 steps equal the free ones bit for bit."""
    o = Oracle((16, 16, 16))
    V0 = o.ic_random(6, E0=0.5, kp=3.0)
    Vf = o.step_ns_forced([v.copy() for v in V0], 0.01, 1e-3, 1, 1.5, 3.5, 0.0, 0.0, scheme=scheme)
    Vu = [v.copy() for v in V0]
    [o.euler_ns, o.rk2_ns, o.rk4_ns][scheme](Vu, 0.01, 1e-3)
    for c in range(3):
        assert np.array_equal(Vf[c], Vu[c])


@pytest.mark.parametrize("epsU,epsW", [(0.5, 0.0), (0.0, 0.4), (0.3, 0.2)])
def test_forced_mhd_step_injects_each_supply_into_its_field(epsU, epsW):
    """MHD step with the supply of both fields (compute_force_ek_hk_supply.cc:294-420) from U and B with
    real coefficients: E_U grows by epsU dt and E_B by epsW dt over the free step, up to O(dt^2) (the
    remainder of a forced field halving with dt); a supply sent to the wrong field fails the split."""
    o = Oracle((16, 16, 16))
    nu, eta = 0.01, 0.02
    U0, B0 = _real_coefficient_field(o, 4), _real_coefficient_field(o, 9)
    U0 = [u * math.sqrt(0.5 / o.total_energy(U0)) for u in U0]
    B0 = [b * math.sqrt(0.3 / o.total_energy(B0)) for b in B0]
    res = []
    for dt in (2e-3, 1e-3):
        Uf, Bf = o.rk4_mhd_forced([u.copy() for u in U0], [b.copy() for b in B0], nu, eta, dt, 1.5, 3.5,
                                  epsU, 0.0, epsW, 0.0)
        Uu, Bu = o.rk4_mhd([u.copy() for u in U0], [b.copy() for b in B0], nu, eta, dt)
        res.append((abs((o.total_energy(Uf) - o.total_energy(Uu)) / dt - epsU),
                    abs((o.total_energy(Bf) - o.total_energy(Bu)) / dt - epsW)))
    for f, eps in enumerate((epsU, epsW)):
        assert res[1][f] < 2e-4 * 0.5
        if eps > 0:                       # the unforced field's change is of higher order
            assert 0.4 < res[1][f] / res[0][f] < 0.6


def test_forced_mhd_step_with_zero_supply_is_the_free_step():
    """Both supplies zero (below MYEPS per mode): no force is added and the step equals the free MHD step
    bit for bit."""
    o = Oracle((16, 16, 16))
    U0, B0 = o.ic_random(6, E0=0.5, kp=3.0), o.ic_random(7, E0=0.3, kp=3.0)
    Uf, Bf = o.rk4_mhd_forced([u.copy() for u in U0], [b.copy() for b in B0], 0.01, 0.02, 1e-3, 1.5, 3.5,
                              0.0, 0.3, 0.0, 0.2)
    Uu, Bu = o.rk4_mhd([u.copy() for u in U0], [b.copy() for b in B0], 0.01, 0.02, 1e-3)
    for c in range(3):
        assert np.array_equal(Uf[c], Uu[c])
        assert np.array_equal(Bf[c], Bu[c])


@pytest.mark.parametrize("proc,p4,p6", [(1, 0.3, -0.2), (2, 0.3, -0.5)])
def test_mhd_force_is_the_field_force_of_each_part(proc, p4, p6):
    """Compute_force(IncVF& W) (compute_force_ek_hk_supply.cc:294-420, compute_force_const_ek_hk.cc:296-437):
    on fields without sub-MYEPS modes FU is the pinned force of U with (p3, p4) and FW that of B with (p5, p6);
    a zero level gives a zero part."""
    o = Oracle((16, 16, 16))
    U, B = o.ic_random(3, E0=0.5, kp=3.0), o.ic_random(4, E0=0.3, kp=3.0)
    FU, FW = o.force_mhd(U, B, proc, 1.5, 3.5, 0.5, p4, 0.4, p6)
    RU, RW = o.force(U, proc, 1.5, 3.5, 0.5, p4), o.force(B, proc, 1.5, 3.5, 0.4, p6)
    for c in range(3):
        assert np.array_equal(FU[c], RU[c]) and np.array_equal(FW[c], RW[c])
    assert max(np.abs(f).max() for f in FU) > 0 and max(np.abs(f).max() for f in FW) > 0
    FU, FW = o.force_mhd(U, B, proc, 1.5, 3.5, 0.0, p4, 0.4, p6)
    assert max(np.abs(f).max() for f in FU) == 0 and max(np.abs(f).max() for f in FW) > 0


def test_constant_level_forces_weak_magnetic_modes():
    """proc 2 forces a mode of B whose modal energy is above 0 (compute_force_const_ek_hk.cc:401), a mode of U
    only above MYEPS (:377): a band mode with energy 1e-17 in both fields is set to |FW|^2/2 = energy_level_W
    / nf (the constant-level closed form, no helicity) and FU stays 0 there; under proc 1 neither is forced."""
    o = Oracle((16, 16, 16))
    U, B = o.zeros(3), o.zeros(3)
    a = math.sqrt(2e-17)                  # |V|^2 / 2 = 1e-17, K = (3, 0, 0), V along y (solenoidal, real)
    for V in (U, B):
        V[1][3, 0, 0] = a
        V[1][-3, 0, 0] = a
    nf = o.force_supply_modes(2.5, 3.5)
    FU, FW = o.force_mhd(U, B, 2, 2.5, 3.5, 0.5, 0.0, 0.4, 0.0)
    assert max(np.abs(f).max() for f in FU) == 0
    e = 0.5 * sum(abs(FW[c][3, 0, 0]) ** 2 for c in range(3))
    assert abs(e - 0.4 / nf) < 1e-13 * (0.4 / nf)
    FU, FW = o.force_mhd(U, B, 1, 2.5, 3.5, 0.5, 0.0, 0.4, 0.0)
    assert max(np.abs(f).max() for f in FU) == 0 and max(np.abs(f).max() for f in FW) == 0


@pytest.mark.parametrize("k0,amp", [(1, 0.8), (2, -0.5)])
def test_taylor_green_force_is_the_closed_form_field(k0, amp):
    """force_field_proc 3 (Setup_Taylor_Green_force_field force_model.cc:57-88): the inverse transform of the
    force is amp (sin k0x cos k0y cos k0z, -cos k0x sin k0y cos k0z, 0) on the grid, whatever the velocity;
    the MHD W part takes forceW_amp (compute_force_TG.cc:90-104)."""
    N = (16, 16, 16)
    o = Oracle(N)
    V = o.ic_random(3, E0=0.5, kp=3.0)
    F = o.force(V, 3, float(k0) + 0.7, amp, 0.0)           # k0 is truncated to int, as the reference does
    x = 2 * np.pi * np.arange(16) / 16
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ref = [amp * np.sin(k0 * X) * np.cos(k0 * Y) * np.cos(k0 * Z),
           -amp * np.cos(k0 * X) * np.sin(k0 * Y) * np.cos(k0 * Z), np.zeros_like(X)]
    for c in range(3):
        A = F[c].copy()
        o.inverse(A)
        got = A.view(np.float64).reshape(16, 16, -1)[:, :, :16]
        assert np.abs(got - ref[c]).max() < 1e-13
    FU, FW = o.force_mhd(V, V, 3, float(k0), amp, 0.3, 0.0, 0.0, 0.0)
    for c in range(3):
        assert np.array_equal(FU[c], F[c])
        assert np.allclose(FW[c], F[c] * (0.3 / amp), rtol=0, atol=1e-15)


@pytest.mark.parametrize("k0,A,B,C,amp", [(1, 1.0, 0.7, -0.4, 0.5), (3, -0.3, 1.2, 0.9, 2.0)])
def test_abc_force_is_the_closed_form_field(k0, A, B, C, amp):
    """force_field_proc 4 (Setup_ABC_force_field force_model.cc:122-160): the inverse transform of the force is
    amp (B cos k0y + C sin k0z, A sin k0x + C cos k0z, A cos k0x + B sin k0y); the field is solenoidal and
    satisfies the reality condition."""
    N = (16, 16, 16)
    o = Oracle(N)
    F = o.force(o.zeros(3), 4, float(k0), A, B, C, amp)
    x = 2 * np.pi * np.arange(16) / 16
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ref = [amp * (B * np.cos(k0 * Y) + C * np.sin(k0 * Z)), amp * (A * np.sin(k0 * X) + C * np.cos(k0 * Z)),
           amp * (A * np.cos(k0 * X) + B * np.sin(k0 * Y))]
    for c in range(3):
        R = F[c].copy()
        o.satisfy_reality(R)
        assert np.array_equal(R, F[c])
        o.inverse(R)
        got = R.view(np.float64).reshape(16, 16, -1)[:, :, :16]
        assert np.abs(got - ref[c]).max() < 1e-13
    KX, KY, KZ = _kgrid(o)
    assert np.abs(KX * F[0] + KY * F[1] + KZ * F[2]).max() == 0
