"""Pins the oracle's anisotropic (ring and cylindrical-ring) spectra (CPU only) against the already
pinned shell spectrum, the toroidal / poloidal split of single modes and closed-form bin locations.

Reference: Compute_local_ring_spectrum_FOUR four_energy.cc:828-954, Compute_local_cylinder_ring_spectrum_FOUR
four_energy.cc:1450-1515, Get_sector_index basis_basicfn.cc:304-322, Get_slab_index :373-390, the sector and
slab arrays of Cvf.cc:85-106, 166-190."""
import math

import numpy as np
import pytest

from pyoracle import Oracle


def _shell_sum(o, V):
    tot = 0
    for c in range(3):
        r = o.shell_spectrum(V[c])
        tot = tot + np.asarray(r[0] if isinstance(r, tuple) else r)
    return tot


def _zero_on_axis(o, V, axis, comps=(0, 1, 2)):
    W = [v.copy() for v in V]
    N1, N2, N3 = o.N
    for l1 in range(N1):
        for l2 in range(N2):
            for l3 in range(N3 // 2 + 1):
                k = [o.kx(l1), o.ky(l2), l3]
                off = [k[j] for j in range(3) if j != axis - 1]
                if off[0] == 0 and off[1] == 0:
                    for c in comps:
                        W[c][l1, l2, l3] = 0
    return W


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_ring_sectors_sum_to_the_shell_spectrum(axis):
    """Summed over sectors, S1 + S2 of a solenoidal field is the shell spectrum on shells <= Max_radius_inside,
    except for the on-axis modes: axis 1 counts them fully, axis 2 counts |V_z|^2 + |V_y|^2 (so |V_x|^2 is
    missing), axis 3 never assigns them (four_energy.cc:883-897)."""
    o = Oracle((16, 16, 16))
    V = o.ic_random(3)
    S1, S2 = o.ring_spectrum(V, 4, 0, axis)
    kk = o.max_radius_inside()
    ref = {1: V, 2: _zero_on_axis(o, V, 2, comps=(0,)), 3: _zero_on_axis(o, V, 3)}[axis]
    sp = _shell_sum(o, ref)[:kk + 1]
    assert np.abs((S1 + S2).sum(axis=1) - sp).max() < 1e-14 * sp.max()
    assert np.all(S1 >= 0) and np.all(S2 >= -1e-15)
    assert np.all(S1[:, 0] == 0) and np.all(S2[:, 0] == 0)          # sector 0 is never filled


@pytest.mark.parametrize("nsect", [3, 4])
def test_ring_single_modes_toroidal_poloidal_and_sector(nsect):
    """Axis 3.  K = (0, 2, 2) has theta = pi/4; V along (1, 0, 0) is toroidal (perpendicular to the axis and
    to K): all of it in S1.  K = (3, 0, 4) has theta = atan(3/4); V = (4, 0, -3) is poloidal: all in S2.
    Shells ceil|K| = 3 and 5; sector = first i with theta <= i (pi/2) / nsect; n = 2 weights by |K|^2."""
    o = Oracle((16, 16, 16))
    V = o.zeros(3)
    A, B = 0.7, 0.2
    V[0][0, 2, 2] = A
    V[0][3, 0, 4], V[2][3, 0, 4] = 4 * B, -3 * B
    dth = (math.pi / 2) / nsect
    sec = lambda t: next(i for i in range(nsect + 1) if t <= i * dth + (1e-15 if i == nsect else 0))
    for n in (0, 2):
        S1, S2 = o.ring_spectrum(V, nsect, n, 3)
        t1, t2 = math.pi / 4, math.atan(3 / 4)
        e1 = A * A * (8.0 if n == 2 else 1.0)
        e2 = 25 * B * B * (25.0 if n == 2 else 1.0)
        R1, R2 = np.zeros_like(S1), np.zeros_like(S2)
        R1[3, sec(t1)] += e1
        R2[5, sec(t2)] += e2
        assert np.abs(S1 - R1).max() < 1e-14 and np.abs(S2 - R2).max() < 1e-13


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_cylinder_rings_sum_to_the_field_energy(axis):
    """Every mode with ceil(K_perp) <= Anis_max_Krho_radius_inside is binned: S1 + S2 sums to sum mult |V|^2 over
    them, S1 to sum mult |V_a|^2 (zero for axis 3, whose V_a is never assigned, four_energy.cc:1493-1500)."""
    o = Oracle((16, 24, 32))
    V = o.ic_random(5)
    S1, S2 = o.cylinder_ring_spectrum(V, 5, 0, axis)
    N1, N2, N3 = o.N
    kpm = S1.shape[0] - 1
    tot = va = 0.0
    for l1 in range(N1):
        for l2 in range(N2):
            for l3 in range(N3 // 2 + 1):
                k = [o.kx(l1), o.ky(l2), l3]
                kp = math.sqrt(sum(k[j] ** 2 for j in range(3) if j != axis - 1))
                if math.ceil(kp) > kpm:
                    continue
                m = o.multiplicity(l1, l2, l3)
                tot += m * sum(abs(V[c][l1, l2, l3]) ** 2 for c in range(3))
                if axis != 3:
                    va += m * abs(V[axis - 1][l1, l2, l3]) ** 2
    assert abs((S1 + S2).sum() - tot) < 1e-13 * tot
    assert abs(S1.sum() - va) < 1e-13 * max(tot, 1e-300)


def test_cylinder_slab_location_and_overflow():
    """Axis 3, N = 16, dealiased: slabs i (N3/3)/nslab, |K_pll| = 0 in slab 1.  A mode with K_perp = 2,
    K_pll = 3 sits in ring 2, slab ceil(3 / (5/5)) = 3; a (non-dealiased) mode with K_pll = 7 > 5 runs past
    the last column and lands on the next ring's column 0, as the reference's unchecked store does."""
    o = Oracle((16, 16, 16))
    V = o.zeros(3)
    V[0][2, 0, 3] = 0.5           # K = (2, 0, 3)
    V[1][0, 0, 0] = 0.25          # K = 0: ring 0, slab 1
    V[0][0, 2, 7] = 0.1           # K = (0, 2, 7): K_pll beyond N3/3
    S1, S2 = o.cylinder_ring_spectrum(V, 5, 0, 3)
    R = np.zeros_like(S2)
    R[2, 3] += 0.25
    R[0, 1] += 0.5 * 0.0625       # kz = 0: multiplicity 1/2
    R[3, 0] += 0.01
    assert np.abs(S1).max() == 0.0
    assert np.abs(S2 - R).max() < 1e-15


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_ring_and_cylinder_helicity_sum_to_the_pinned_helicities(axis):
    """Compute_ring_spectrum_helicity_FOUR / Compute_cylinder_ring_spectrum_helicity_FOUR (four_energy.cc:1357-1412,
    1825-1880) bin the modal helicity 2 mult K.(Vr x Vi): summed over sectors the ring table is the shell
    helicity spectrum (shells <= kmax_inside), and for a field inside every cylindrical ring the cylinder table
    sums to the total helicity H1."""
    o = Oracle((16, 16, 16))
    V = o.ic_random(7)
    R = o.ring_spectrum_helicity(V, 5, axis)
    kin = o.max_radius_inside()
    Hk = np.asarray(o.shell_helicity(V)).sum(axis=0)
    assert np.abs(R.sum(axis=1) - Hk[:kin + 1]).max() < 1e-13 * np.abs(Hk).max()
    KX, KY, KZ = np.meshgrid(np.array([o.kx(l) for l in range(16)], float), np.array([o.ky(l) for l in range(16)], float),
                             np.arange(9, dtype=float), indexing="ij")
    W = [np.where(np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2) <= 3.0, v, 0) for v in V]
    C = o.cylinder_ring_spectrum_helicity(W, 4, axis)
    H1 = o.helicity(W)[0]
    assert abs(C.sum() - H1) < 1e-13 * max(abs(H1), 1e-300) and abs(H1) > 0
