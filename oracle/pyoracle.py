"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg
of bench.py.  Fields are numpy complex128 arrays of shape (N1, N2, N3//2+1), C-order, the
reference's layout (trunk/lib/fields/Cvf.cc:112-114).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


class Grid(C.Structure):
    _fields_ = [("N1", C.c_int), ("N2", C.c_int), ("N3", C.c_int),
                ("f1", C.c_double), ("f2", C.c_double), ("f3", C.c_double),
                ("basis", C.c_int), ("dealias", C.c_int)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        d, i, p, vp = C.c_double, C.c_int, C.POINTER(Grid), C.c_void_p
        pp = C.POINTER(C.c_void_p)
        sig = {
            "orc_fft1d": (None, [i, i, vp]),
            "orc_set_threads": (None, [i]),
            "orc_get_threads": (i, []),
            "orc_kx": (i, [p, i]), "orc_ky": (i, [p, i]),
            "orc_kmag": (d, [p, i, i, i]),
            "orc_multiplicity": (d, [p, i, i, i]),
            "orc_min_radius_outside": (i, [p]),
            "orc_max_radius_inside": (i, [p]),
            "orc_approx_modes_in_shell": (d, [p, i]),
            "orc_shell_ek_size": (i, [p]),
            "orc_kept": (i, [p, i, i, i]),
            "orc_forward": (None, [p, vp, i]),
            "orc_inverse": (None, [p, vp, i]),
            "orc_deriv": (None, [p, i, vp, vp, i]),
            "orc_dealias": (None, [p, vp]),
            "orc_real_mult": (None, [p, vp, vp, vp]),
            "orc_divide_ksqr": (None, [p, vp]),
            "orc_exp_ksqr": (None, [p, vp, d]),
            "orc_satisfy_reality": (None, [p, vp]),
            "orc_nlin_ns": (None, [p, pp, pp]),
            "orc_pressure": (None, [p, pp, vp]),
            "orc_rhs": (None, [p, pp, vp]),
            "orc_rk4_step_ns": (None, [p, pp, d, d]),
            "orc_euler_step_ns": (None, [p, pp, d, d]),
            "orc_rk2_step_ns": (None, [p, pp, d, d]),
            "orc_force_supply_modes": (i, [p, d, d]),
            "orc_force_ek_hk": (i, [p, i, pp, pp, d, d, d, d]),
            "orc_step_ns_forced": (None, [p, pp, d, d, i, i, d, d, d, d, d, pp]),
            "orc_force_abc": (None, [p, pp, i, d, d, d, d]),
            "orc_rk4_stage1_ns": (None, [p, pp, d, d]),
            "orc_nlin_mhd": (None, [p, pp, pp, pp, pp]),
            "orc_rk4_step_mhd": (None, [p, pp, pp, d, d, d]),
            "orc_rk4_step_mhd_forced": (None, [p, pp, pp, d, d, d, i, d, d, d, d, d, d, pp]),
            "orc_force_mhd": (i, [p, i, pp, pp, pp, pp, d, d, d, d, d, d]),
            "orc_energy": (d, [p, vp, vp]),
            "orc_Sn": (d, [p, vp, d]),
            "orc_shell_spectrum": (None, [p, vp, vp, d, vp, vp]),
            "orc_dt": (d, [p, pp, d]),
            "orc_fill_array_shell": (None, [p, vp, vp, d, d]),
            "orc_shell_mult_single": (d, [p, vp, vp, d, d]),
            "orc_shell_mult_all": (None, [p, pp, pp, vp, i, vp]),
            "orc_et_nlin": (None, [p, pp, pp, pp]),
            "orc_flux_ns": (None, [p, pp, vp, i, vp]),
            "orc_shell_tr_ns": (None, [p, pp, vp, i, vp]),
            "orc_flux_mhd": (None, [p, pp, pp, vp, i] + [vp] * 9),
            "orc_shell_tr_mhd": (None, [p, pp, pp, vp, i] + [vp] * 5),
            "orc_ring_tr": (None, [p, i, i, pp, pp, pp, i, vp, i, vp, i, i, d, vp]),
            "orc_et_radii": (None, [p, i, i, vp]),
            "orc_et_nlin_scalar": (None, [p, pp, vp, vp]),
            "orc_helicity": (None, [p, pp, vp]),
            "orc_ring_spectrum": (None, [p, pp, i, i, i, vp, vp]),
            "orc_cylinder_shell_ek_size": (i, [p, i]),
            "orc_aniso_helicity": (None, [p, pp, i, i, i, vp]),
            "orc_cylinder_ring_spectrum": (None, [p, pp, i, i, i, vp, vp]),
            "orc_shell_helicity": (None, [p, pp, vp, vp, vp]),
            "orc_entropy": (d, [p, pp]),
            "orc_entropy_scalar": (d, [p, vp]),
            "orc_flux_scalar": (None, [p, pp, vp, vp, i, vp, vp]),
            "orc_shell_tr_scalar": (None, [p, pp, vp, vp, i, vp, vp]),
            "orc_nlin_rb": (None, [p, pp, vp, pp, vp]),
            "orc_rk4_step_rb": (None, [p, pp, vp, d, d, d, pp]),
            "orc_nusselt": (d, [p, pp, vp, d, d]),
            "orc_ic_taylor_green": (None, [p, pp, i, d]),
            "orc_ic_random": (None, [p, pp, C.c_ulonglong, d, d, i]),
            "orc_ic_random_scalar": (None, [p, vp, C.c_ulonglong, d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _ptr(a):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _pp(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class Oracle:
    """FP64 CPU restatement bound to one grid."""

    def __init__(self, N, kfactor=(1.0, 1.0, 1.0), basis="FOUR", dealias=True, threads=0):
        self.N = tuple(int(n) for n in N)
        self.g = Grid(self.N[0], self.N[1], self.N[2], float(kfactor[0]), float(kfactor[1]),
                      float(kfactor[2]), 0 if basis == "FOUR" else 1, 1 if dealias else 0)
        self.gp = C.byref(self.g)
        self.L = lib()
        if threads:
            self.L.orc_set_threads(int(threads))

    @property
    def shape(self):
        return (self.N[0], self.N[1], self.N[2] // 2 + 1)

    def zeros(self, n=None):
        if n is None:
            return np.zeros(self.shape, np.complex128)
        return [np.zeros(self.shape, np.complex128) for _ in range(n)]

    # transforms and spectral ops
    def forward(self, A, parity=0):
        self.L.orc_forward(self.gp, _ptr(A), parity)
        return A

    def inverse(self, A, parity=0):
        self.L.orc_inverse(self.gp, _ptr(A), parity)
        return A

    def deriv(self, axis, A, parity=0):
        B = np.empty_like(A)
        self.L.orc_deriv(self.gp, axis, _ptr(A), _ptr(B), parity)
        return B

    def dealias(self, A):
        self.L.orc_dealias(self.gp, _ptr(A))
        return A

    def real_mult(self, A, B):
        Cc = np.empty_like(A)
        self.L.orc_real_mult(self.gp, _ptr(A), _ptr(B), _ptr(Cc))
        return Cc

    def divide_ksqr(self, A):
        self.L.orc_divide_ksqr(self.gp, _ptr(A))
        return A

    def exp_ksqr(self, A, factor):
        self.L.orc_exp_ksqr(self.gp, _ptr(A), factor)
        return A

    def satisfy_reality(self, A):
        self.L.orc_satisfy_reality(self.gp, _ptr(A))
        return A

    # wavenumbers / binning
    def kx(self, l1):
        return self.L.orc_kx(self.gp, l1)

    def ky(self, l2):
        return self.L.orc_ky(self.gp, l2)

    def kmag(self, l1, l2, l3):
        return self.L.orc_kmag(self.gp, l1, l2, l3)

    def multiplicity(self, l1, l2, l3):
        return self.L.orc_multiplicity(self.gp, l1, l2, l3)

    def min_radius_outside(self):
        return self.L.orc_min_radius_outside(self.gp)

    def max_radius_inside(self):
        return self.L.orc_max_radius_inside(self.gp)

    def shell_ek_size(self):
        return self.L.orc_shell_ek_size(self.gp)

    def kept(self, l1, l2, l3):
        return self.L.orc_kept(self.gp, l1, l2, l3)

    def approx_modes(self, r):
        return self.L.orc_approx_modes_in_shell(self.gp, r)

    # physics
    def nlin_ns(self, V):
        N = self.zeros(3)
        self.L.orc_nlin_ns(self.gp, _pp(V), _pp(N))
        return N

    def pressure(self, nlin):
        p = self.zeros()
        self.L.orc_pressure(self.gp, _pp(nlin), _ptr(p))
        return p

    def rhs(self, nlin, p):
        self.L.orc_rhs(self.gp, _pp(nlin), _ptr(p))
        return nlin

    def rk4_ns(self, V, nu, dt, steps=1):
        for _ in range(steps):
            self.L.orc_rk4_step_ns(self.gp, _pp(V), nu, dt)
        return V

    def rk4_stage1_ns(self, V, nu, dt):
        self.L.orc_rk4_stage1_ns(self.gp, _pp(V), nu, dt)
        return V

    def euler_ns(self, V, nu, dt):
        self.L.orc_euler_step_ns(self.gp, _pp(V), nu, dt)
        return V

    def rk2_ns(self, V, nu, dt):
        self.L.orc_rk2_step_ns(self.gp, _pp(V), nu, dt)
        return V

    def force_supply_modes(self, inner, outer):
        """nf of Compute_force_const_energy_helicity_supply (modes with inner <= |K| < outer)."""
        return self.L.orc_force_supply_modes(self.gp, inner, outer)

    def force(self, V, proc, inner, outer, p3, p4=0.0, p5=0.0):
        """Force field F[3] of force_field_proc 1 (energy/helicity supply: p3 = energy_supply,
        p4 = epsh_by_k_epse), 2 (constant energy/helicity: p3 = energy_level, p4 = h_by_k_E), 3 (Taylor-Green:
        inner = k0, outer = amp) or 4 (ABC: inner = k0, outer = A, p3 = B, p4 = C, p5 = amp)."""
        F = self.zeros(3)
        if proc == 4:
            self.L.orc_force_abc(self.gp, _pp(F), int(inner), outer, p3, p4, p5)
            return F
        nf = self.L.orc_force_ek_hk(self.gp, proc, _pp(V), _pp(F), inner, outer, p3, p4)
        if nf <= 0 and proc != 3:             # proc 3 (Taylor-Green: inner = k0, outer = amp) counts no modes
            raise ValueError("no mode in the forcing band")
        return F

    def step_ns_forced(self, V, nu, dt, proc, inner, outer, p3, p4=0.0, scheme=2, steps=1, last_stage=None, p5=0.0):
        """NS steps (scheme 0 EULER, 1 RK2, 2 RK4) with the force recomputed at every stage; last_stage
        (3 arrays, optional) receives the field of the last stage of the last step."""
        for _ in range(steps):
            self.L.orc_step_ns_forced(self.gp, _pp(V), nu, dt, scheme, proc, inner, outer, p3, p4, p5,
                                      _pp(last_stage) if last_stage is not None else None)
        return V

    def nlin_mhd(self, U, B):
        nU, nB = self.zeros(3), self.zeros(3)
        self.L.orc_nlin_mhd(self.gp, _pp(U), _pp(B), _pp(nU), _pp(nB))
        return nU, nB

    def rk4_mhd(self, U, B, nu, eta, dt, steps=1):
        for _ in range(steps):
            self.L.orc_rk4_step_mhd(self.gp, _pp(U), _pp(B), nu, eta, dt)
        return U, B

    def rk4_mhd_forced(self, U, B, nu, eta, dt, inner, outer, p3, p4, p5, p6, steps=1, last_stage=None, proc=1):
        """MHD RK4 steps with force_field_proc `proc` forcing of U (p3, p4) and W (p5, p6); last_stage (6
        arrays, optional) receives U1..U3, B1..B3 of the last stage."""
        for _ in range(steps):
            self.L.orc_rk4_step_mhd_forced(self.gp, _pp(U), _pp(B), nu, eta, dt, proc, inner, outer, p3, p4, p5, p6,
                                           _pp(last_stage) if last_stage is not None else None)
        return U, B

    def force_mhd(self, U, B, proc, inner, outer, p3, p4, p5, p6):
        """Compute_force(IncVF& W): (FU, FW)."""
        FU, FW = self.zeros(3), self.zeros(3)
        self.L.orc_force_mhd(self.gp, proc, _pp(U), _pp(B), _pp(FU), _pp(FW), inner, outer, p3, p4, p5, p6)
        return FU, FW

    def nlin_rb(self, V, T):
        nV, nT = self.zeros(3), self.zeros()
        self.L.orc_nlin_rb(self.gp, _pp(V), _ptr(T), _pp(nV), _ptr(nT))
        return nV, nT

    def rk4_rb(self, V, T, Ra, Pr, dt, steps=1, last_stage=None):
        """RB RK4 steps; last_stage (4 arrays, optional) receives the last stage's V1..V3, T."""
        for _ in range(steps):
            self.L.orc_rk4_step_rb(self.gp, _pp(V), _ptr(T), Ra, Pr, dt,
                                   _pp(last_stage) if last_stage is not None else None)
        return V, T

    def nusselt(self, V, T, Ra, Pr):
        return self.L.orc_nusselt(self.gp, _pp(V), _ptr(T), Ra, Pr)

    def energy(self, A, B=None):
        return self.L.orc_energy(self.gp, _ptr(A), _ptr(B) if B is not None else None)

    def Sn(self, A, n):
        return self.L.orc_Sn(self.gp, _ptr(A), float(n))

    def total_energy(self, V):
        return sum(self.energy(a) for a in V)

    def total_dissipation(self, V):
        return 2 * sum(self.Sn(a, 2) for a in V)

    def shell_spectrum(self, A, B=None, n=0):
        size = self.shell_ek_size()
        Sk = np.zeros(size)
        cnt = np.zeros(size)
        self.L.orc_shell_spectrum(self.gp, _ptr(A), _ptr(B) if B is not None else None, float(n),
                                  Sk.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p))
        return Sk, cnt

    def vector_spectrum(self, V, n=0):
        return sum(self.shell_spectrum(a, None, n)[0] for a in V)

    def dt(self, V, dt_fixed):
        return self.L.orc_dt(self.gp, _pp(V), dt_fixed)

    # energy transfer
    def et_radii(self, scheme, no_shells):
        r = np.zeros(no_shells + 1)
        self.L.orc_et_radii(self.gp, scheme, no_shells, r.ctypes.data_as(C.c_void_p))
        return r

    def fill_shell(self, A, inner, outer):
        B = np.empty_like(A)
        self.L.orc_fill_array_shell(self.gp, _ptr(A), _ptr(B), inner, outer)
        return B

    def shell_mult_single(self, A, B, inner, outer):
        return self.L.orc_shell_mult_single(self.gp, _ptr(A), _ptr(B), inner, outer)

    def shell_mult_all(self, A, B, radii):
        res = np.zeros(len(radii))
        self.L.orc_shell_mult_all(self.gp, _pp(A), _pp(B), radii.ctypes.data_as(C.c_void_p),
                                  len(radii), res.ctypes.data_as(C.c_void_p))
        return res

    def et_nlin(self, H, G):
        N = self.zeros(3)
        self.L.orc_et_nlin(self.gp, _pp(H), _pp(G), _pp(N))
        return N

    def helicity(self, V):
        """(H1, H2, DH1, DH2) of Compute_total_helicity_FOUR."""
        out = np.zeros(4)
        self.L.orc_helicity(self.gp, _pp(V), out.ctypes.data_as(C.c_void_p))
        return out

    def ring_spectrum(self, V, nsect, n=0, axis=3):
        """(S1k, S2k) of Compute_ring_spectrum_FOUR, shape (Max_radius_inside+1, nsect+1)."""
        shape = (self.max_radius_inside() + 1, nsect + 1)
        S1, S2 = np.zeros(shape), np.zeros(shape)
        self.L.orc_ring_spectrum(self.gp, _pp(V), nsect, n, axis, S1.ctypes.data_as(C.c_void_p),
                                 S2.ctypes.data_as(C.c_void_p))
        return S1, S2

    def cylinder_ring_spectrum(self, V, nslab, n=0, axis=3):
        """(S1k, S2k) of Compute_cylinder_ring_spectrum_FOUR, shape (Anis_max_Krho_radius_inside+1, nslab+1)."""
        shape = (self.L.orc_cylinder_shell_ek_size(self.gp, axis), nslab + 1)
        S1, S2 = np.zeros(shape), np.zeros(shape)
        self.L.orc_cylinder_ring_spectrum(self.gp, _pp(V), nslab, n, axis, S1.ctypes.data_as(C.c_void_p),
                                          S2.ctypes.data_as(C.c_void_p))
        return S1, S2

    def ring_spectrum_helicity(self, V, nsect, axis=3):
        """H1k of Compute_ring_spectrum_helicity_FOUR, shape (Max_radius_inside+1, nsect+1)."""
        H = np.zeros((self.max_radius_inside() + 1, nsect + 1))
        self.L.orc_aniso_helicity(self.gp, _pp(V), 0, nsect, axis, H.ctypes.data_as(C.c_void_p))
        return H

    def cylinder_ring_spectrum_helicity(self, V, nslab, axis=3):
        """H1k of Compute_cylinder_ring_spectrum_helicity_FOUR, shape (Anis_max_Krho_radius_inside+1, nslab+1)."""
        H = np.zeros((self.L.orc_cylinder_shell_ek_size(self.gp, axis), nslab + 1))
        self.L.orc_aniso_helicity(self.gp, _pp(V), 1, nslab, axis, H.ctypes.data_as(C.c_void_p))
        return H

    def shell_helicity(self, V):
        n = self.shell_ek_size()
        H = np.zeros((3, n))
        self.L.orc_shell_helicity(self.gp, _pp(V), *[H[c].ctypes.data_as(C.c_void_p) for c in range(3)])
        return H

    def entropy(self, V):
        return self.L.orc_entropy(self.gp, _pp(V))

    def entropy_scalar(self, A):
        """Get_entropy_scalar of one field (bits)."""
        return self.L.orc_entropy_scalar(self.gp, _ptr(A))

    def et_nlin_scalar(self, H, G):
        N = self.zeros()
        self.L.orc_et_nlin_scalar(self.gp, _pp(H), _ptr(G), _ptr(N))
        return N

    def flux_scalar(self, V, T, radii):
        """(flux_self, flux_SF) of Compute_flux_scalar / Compute_flux_RB for spheres radii[1..]."""
        nsph = len(radii) - 1
        fs, ft = np.zeros(nsph + 1), np.zeros(nsph + 1)
        self.L.orc_flux_scalar(self.gp, _pp(V), _ptr(T), radii.ctypes.data_as(C.c_void_p), nsph,
                               fs.ctypes.data_as(C.c_void_p), ft.ctypes.data_as(C.c_void_p))
        return fs, ft

    def shell_tr_scalar(self, V, T, radii):
        """(T_self, T_SF), each (nsh+1)^2, of Compute_shell_tr_scalar / Compute_shell_tr_RB."""
        nsh = len(radii) - 1
        Ts, Tt = np.zeros((nsh + 1, nsh + 1)), np.zeros((nsh + 1, nsh + 1))
        self.L.orc_shell_tr_scalar(self.gp, _pp(V), _ptr(T), radii.ctypes.data_as(C.c_void_p), nsh,
                                   Ts.ctypes.data_as(C.c_void_p), Tt.ctypes.data_as(C.c_void_p))
        return Ts, Tt

    def flux_ns(self, V, radii):
        nsph = len(radii) - 1
        flux = np.zeros(nsph + 1)
        self.L.orc_flux_ns(self.gp, _pp(V), radii.ctypes.data_as(C.c_void_p), nsph,
                           flux.ctypes.data_as(C.c_void_p))
        return flux

    def shell_tr_ns(self, V, radii):
        nsh = len(radii) - 1
        T = np.zeros((nsh + 1, nsh + 1))
        self.L.orc_shell_tr_ns(self.gp, _pp(V), radii.ctypes.data_as(C.c_void_p), nsh,
                               T.ctypes.data_as(C.c_void_p))
        return T

    def shell_tr_mhd(self, U, B, radii):
        """The five (nsh+1)^2 matrices of Compute_shell_tr(IncVF& W): uu, ww, uw, zp, zm."""
        nsh = len(radii) - 1
        outs = [np.zeros((nsh + 1, nsh + 1)) for _ in range(5)]
        self.L.orc_shell_tr_mhd(self.gp, _pp(U), _pp(B), radii.ctypes.data_as(C.c_void_p), nsh,
                                *[o.ctypes.data_as(C.c_void_p) for o in outs])
        return dict(zip(["uu", "ww", "uw", "zp", "zm"], outs))

    def ring_tr(self, H, F, R, radii, edges, ngiv=None, sign=-1.0, cyl=False, axis=3):
        """Ring-to-ring (cyl=False: shells x polar sectors) or cylinder-ring-to-ring (cyl=True: K_perp shells x
        K_pll slabs) transfer of Compute_ring_tr / Compute_cylinder_ring_tr: helper H (3 comps), giver source F
        and receiver R (3 comps, or 1 for the scalar form); returns a 4-D array
        [giver shell, giver sector, receiver shell, receiver sector]."""
        radii = np.ascontiguousarray(radii, dtype=float)
        edges = np.ascontiguousarray(edges, dtype=float)
        nr, nsec = len(radii) - 1, len(edges) - 1
        ngiv = nr - 1 if ngiv is None else ngiv
        M = (nr + 1) * (nsec + 1)
        out = np.zeros((nr + 1, nsec + 1, nr + 1, nsec + 1))
        nc = len(F)
        pad = lambda X: list(X) + [X[0]] * (3 - len(X))
        self.L.orc_ring_tr(self.gp, int(cyl), axis, _pp(H), _pp(pad(F)), _pp(pad(R)), nc,
                           radii.ctypes.data_as(C.c_void_p), nr, edges.ctypes.data_as(C.c_void_p), nsec, ngiv,
                           float(sign), out.ctypes.data_as(C.c_void_p))
        return out

    def flux_mhd(self, U, B, radii, elsasser=False):
        nsph = len(radii) - 1
        outs = [np.zeros(nsph + 1) for _ in range(9)]
        ptrs = [o.ctypes.data_as(C.c_void_p) for o in outs]
        if not elsasser:
            ptrs[7] = None
            ptrs[8] = None
        self.L.orc_flux_mhd(self.gp, _pp(U), _pp(B), radii.ctypes.data_as(C.c_void_p), nsph, *ptrs)
        names = ["u2u", "b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in",
                 "u_out_b_out", "zp", "zm"]
        return dict(zip(names, outs))

    # initial conditions
    def ic_taylor_green(self, k0=1, amp=1.0):
        V = self.zeros(3)
        self.L.orc_ic_taylor_green(self.gp, _pp(V), k0, amp)
        return V

    def ic_random(self, seed, E0=0.5, kp=4.0, project=True):
        V = self.zeros(3)
        self.L.orc_ic_random(self.gp, _pp(V), seed, E0, kp, 1 if project else 0)
        return V

    def ic_random_scalar(self, seed, amp=1e-3):
        T = self.zeros()
        self.L.orc_ic_random_scalar(self.gp, _ptr(T), seed, amp)
        return T
