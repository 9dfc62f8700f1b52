"""Pins the oracle's ring-to-ring and cylinder-ring-to-ring transfers (compute_ring_tr.cc:60-173,
compute_cylin_ring_tr.cc:54-133) against explicit triad sums over the full wavenumber lattice
(tests/golden/ring_triad.npz, written by tests/golden/make_golden.py), and against the pinned shell-to-shell
transfer: one sector spanning [0, pi/2] about the z axis holds every stored mode, so the ring transfer is the
shell transfer, and with more sectors its sums over both sector indices are.  CPU only."""
import numpy as np
import pytest

from pyoracle import Oracle


@pytest.fixture(scope="module")
def rt(golden_dir):
    G = np.load(f"{golden_dir}/ring_triad.npz")
    o = Oracle(tuple(int(v) for v in G["N"]))
    U = [np.ascontiguousarray(G["U"][i]) for i in range(3)]
    T = np.ascontiguousarray(G["T"])
    return G, o, U, T


@pytest.mark.parametrize("key", ["ring_ns_axis3", "ring_ns_axis1", "ring_sf_axis3", "ring_sf_axis1",
                                 "cyl_ns_axis3", "cyl_ns_axis2", "cyl_sf_axis3", "cyl_sf_axis2"])
def test_ring_tr_triads(rt, key):
    G, o, U, T = rt
    kind, name, ax = key.split("_")
    cyl, a = kind == "cyl", int(ax[-1])
    F = U if name == "ns" else [T]
    edges = G["slabs"] if cyl else G["sectors"]
    got = o.ring_tr(U, F, F, G["radii"], edges, cyl=cyl, axis=a)
    ref = G[key]
    assert np.abs(ref).max() > 1e-6
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), key


def test_ring_tr_one_sector_is_the_shell_transfer():
    o = Oracle((16, 16, 16))
    V = o.ic_random(3)
    radii = np.array([0.0, 2, 4, 6, 10000.0])
    T = o.shell_tr_ns(V, radii)
    R = o.ring_tr(V, V, V, radii, np.array([0.0, np.pi / 2]))
    assert np.abs(R[:, 1, :, 1] - T).max() <= 1e-15 * np.abs(T).max()
    R3 = o.ring_tr(V, V, V, radii, np.array([0.0, 0.3, 1.0, np.pi / 2]))
    assert np.abs(R3[:, 1:, :, 1:].sum(axis=(1, 3)) - T).max() <= 1e-13 * np.abs(T).max()
    # cylinder rings about z: one slab spanning [0, kpll_max] and K_perp shells; summed over slabs and
    # with every mode of a K_perp shell inside [0, kpll_max] the giver is the K_perp shell
    C = o.ring_tr(V, V, V, radii, np.array([0.0, 2.0, 5.0]), cyl=True)
    C1 = o.ring_tr(V, V, V, radii, np.array([0.0, 5.0]), cyl=True)
    assert np.abs(C[:, 1:, :, 1:].sum(axis=(1, 3)) - C1[:, 1, :, 1]).max() <= 1e-13 * np.abs(C1).max()
