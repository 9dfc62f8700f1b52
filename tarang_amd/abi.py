"""ctypes binding of libtarang_hip.so (the C-ABI in include/tarang_hip.h).

Device arrays are torch tensors on cuda (torch is used only for device memory): complex128 for
FP64 plans, complex64 for FP32-mixed plans, shape (N1, N2, N3//2+1) = the reference layout.
The library runs on its own HIP stream; every wrapper synchronises torch's stream before the call
and the library stream after it, so callers can mix the two freely.
"""
import atexit
import ctypes as C
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtarang_hip.so")

TH_FOUR, TH_SCFT = 0, 1
TH_FP64, TH_FP32_MIXED = 0, 1
TH_NS, TH_MHD, TH_RB = 0, 1, 2

_lib = None
_live_plans = weakref.WeakSet()
_live_ctx = weakref.WeakSet()
_live_ens = weakref.WeakSet()
_shutting_down = False


@atexit.register
def _release_all():
    """Destroy ensembles, plans, then contexts, before the HIP runtime is torn down at process exit."""
    global _shutting_down
    for obj in list(_live_ens) + list(_live_plans) + list(_live_states if "_live_states" in globals() else []):
        try:
            obj.close()
        except Exception:
            pass
    for c in list(_live_ctx):
        try:
            c.close()
        except Exception:
            pass
    _shutting_down = True


class Phys(C.Structure):
    _fields_ = [("nu", C.c_double), ("eta", C.c_double), ("kappa", C.c_double),
                ("Ra", C.c_double), ("Pr", C.c_double)]


class Force(C.Structure):
    """th_force: force_field_proc 1 (p3 = energy_supply, p4 = epsh_by_k_epse) or 2 (p3 = energy_level,
    p4 = h_by_k_E) on the band inner_radius < |K| <= outer_radius; p5, p6 are the magnetic field's
    energy_supply_W and epsh_by_k_epse_W of the MHD step."""
    _fields_ = [("proc", C.c_int), ("inner_radius", C.c_double), ("outer_radius", C.c_double),
                ("p3", C.c_double), ("p4", C.c_double), ("p5", C.c_double), ("p6", C.c_double)]


def lib():
    """Load the gfx950 library.  Raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tarang_amd.build` "
                           "(the HIP path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    pp = C.POINTER(C.c_void_p)
    dp = C.POINTER(C.c_double)
    sig = {
        "th_ctx_create": (i, [C.POINTER(vp), i, C.c_char_p]),
        "th_ctx_destroy": (i, [vp]),
        "th_field_write": (i, [C.c_char_p, C.POINTER(C.c_int), i, i, pp, d, i]),
        "th_field_read": (i, [C.c_char_p, C.POINTER(C.c_int), i, i, pp, i, dp]),
        "th_field_write_reduced": (i, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), i, i, i, pp, d, i]),
        "th_field_read_reduced": (i, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), i, i, i, pp, i, dp]),
        "th_last_error": (C.c_char_p, []),
        "th_version": (C.c_char_p, []),
        "th_plan_create": (i, [vp, C.POINTER(vp), i, i, i, i, dp, i, i]),
        "th_plan_destroy": (i, [vp]),
        "th_plan_field_elems": (sz, [vp]),
        "th_plan_elem_bytes": (sz, [vp]),
        "th_shell_ek_size": (i, [vp]),
        "th_max_radius_inside": (i, [vp]),
        "th_sync": (i, [vp]),
        "th_forward": (i, [vp, vp, i]),
        "th_inverse": (i, [vp, vp, i]),
        "th_deriv": (i, [vp, i, vp, vp, i]),
        "th_dealias": (i, [vp, vp]),
        "th_real_mult": (i, [vp, vp, vp, vp]),
        "th_divide_ksqr": (i, [vp, vp]),
        "th_exp_ksqr": (i, [vp, vp, d]),
        "th_nlin_ns": (i, [vp, pp, pp]),
        "th_pressure": (i, [vp, pp, vp]),
        "th_rhs": (i, [vp, pp, vp]),
        "th_nlin_mhd": (i, [vp, pp, pp, pp, pp]),
        "th_nlin_rb": (i, [vp, pp, vp, pp, vp]),
        "th_et_nlin": (i, [vp, pp, pp, pp]),
        "th_step_plain": (i, [vp, i, i, pp, d, C.POINTER(Phys)]),
        "th_step_plain_forced": (i, [vp, i, pp, d, C.POINTER(Phys), C.POINTER(Force), pp]),
        "th_step_plain_forced_mhd": (i, [vp, i, pp, pp, d, C.POINTER(Phys), C.POINTER(Force), pp]),
        "th_compute_force": (i, [vp, pp, pp, C.POINTER(Force), C.POINTER(C.c_long)]),
        "th_compute_force_mhd": (i, [vp, pp, pp, pp, pp, C.POINTER(Force), C.POINTER(C.c_long)]),
        "th_energy": (i, [vp, vp, vp, dp]),
        "th_Sn": (i, [vp, vp, d, dp]),
        "th_shell_spectrum": (i, [vp, vp, vp, i, dp, dp]),
        "th_fill_array_shell": (i, [vp, vp, vp, d, d]),
        "th_shell_mult_single": (i, [vp, pp, pp, d, d, dp]),
        "th_shell_mult_all": (i, [vp, pp, pp, dp, i, dp]),
        "th_flux_ns": (i, [vp, pp, dp, i, dp]),
        "th_shell_tr_ns": (i, [vp, pp, dp, i, dp]),
        "th_flux_mhd": (i, [vp, pp, pp, dp, i, dp]),
        "th_flux_scalar": (i, [vp, pp, vp, dp, i, dp]),
        "th_helicity": (i, [vp, pp, dp]),
        "th_shell_spectrum_helicity": (i, [vp, pp, dp, dp, dp]),
        "th_entropy": (i, [vp, pp, dp]),
        "th_shell_tr_scalar": (i, [vp, pp, vp, dp, i, dp, dp]),
        "th_shell_tr_mhd": (i, [vp, pp, pp, dp, i, dp]),
        "th_shell_spectrum_noalloc": (i, [vp, vp, vp, i, dp, dp]),
        "th_ring_tr": (i, [vp, i, i, pp, pp, pp, i, dp, i, dp, i, i, d, dp]),
        "th_ring_mult_all": (i, [vp, i, i, pp, pp, i, dp, i, dp, i, dp]),
        "th_ub_to_elsasser": (i, [vp, pp, pp, i]),
        "th_flux_mhd_fused": (i, [vp, vp, dp, i, dp]),
        "th_nusselt": (i, [vp, pp, vp, d, d, dp]),
        "th_ic_taylor_green": (i, [vp, pp, i, d]),
        "th_ic_random": (i, [vp, pp, C.c_ulonglong, d, d]),
        "th_ic_random_scalar": (i, [vp, vp, C.c_ulonglong, d]),
        "th_state_create": (i, [vp, C.POINTER(vp), i]),
        "th_state_destroy": (i, [vp]),
        "th_state_upload": (i, [vp, i, vp, i]),
        "th_state_download": (i, [vp, i, vp, i]),
        "th_step": (i, [vp, vp, d, C.POINTER(Phys), i]),
        "th_state_set_force": (i, [vp, C.POINTER(Force)]),
        "th_entropy_scalar": (i, [vp, vp, dp]),
        "th_ring_spectrum": (i, [vp, pp, i, i, i, dp, dp]),
        "th_cylinder_ring_spectrum": (i, [vp, pp, i, i, i, dp, dp]),
        "th_cylinder_shell_ek_size": (i, [vp, i]),
        "th_ring_spectrum_helicity": (i, [vp, pp, i, i, dp]),
        "th_cylinder_ring_spectrum_helicity": (i, [vp, pp, i, i, dp]),
        "th_state_download_buffer": (i, [vp, i, i, vp, i]),
        "th_flux_shell_tr": (i, [vp, vp, dp, i, dp, dp]),
        "th_profile_passes": (i, [vp, vp, d, C.POINTER(Phys), i, i, dp, dp]),
        "th_tune": (i, [vp, vp, d, C.POINTER(Phys), i, C.POINTER(C.c_int)]),
        "th_dft16_probe": (i, [vp, i, i, vp, vp, C.c_long, i, i, dp]),
        "th_state_set_variants": (i, [vp, C.POINTER(C.c_int)]),
        "th_state_variants": (i, [vp, C.POINTER(C.c_int)]),
        "th_ensemble_tune": (i, [vp, dp, dp, i]),
        "th_ensemble_variants": (i, [vp, i, C.POINTER(C.c_int)]),
        "th_ensemble_create": (i, [vp, C.POINTER(vp), i, C.POINTER(C.c_int), i]),
        "th_ensemble_destroy": (i, [vp]),
        "th_ensemble_plan": (i, [vp, i, C.POINTER(vp)]),
        "th_ensemble_upload": (i, [vp, i, i, vp, i]),
        "th_ensemble_download": (i, [vp, i, i, vp, i]),
        "th_ensemble_step": (i, [vp, dp, dp, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().th_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def _torch():
    import torch
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _pp(ts):
    return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def _dbl(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _host_arrays(fields):
    import numpy as np
    out = []
    for f in fields:
        if hasattr(f, "detach"):
            f = f.detach().cpu().numpy()
        out.append(np.ascontiguousarray(f))
    return out


def field_write(path, N, fields, time=0.0, append=False):
    """th_field_write: one snapshot of the reference's field file (field_out.d of Output_field).
    fields: complex128 (FP64) or complex64 arrays of shape (N1, N2, N3/2+1), host or device."""
    import numpy as np
    arrs = _host_arrays(fields)
    dt = arrs[0].dtype
    if dt not in (np.complex128, np.complex64) or any(a.dtype != dt for a in arrs):
        raise TypeError("field_write: all fields complex128 or all complex64")
    shape = (N[0], N[1], N[2] // 2 + 1)
    if any(a.shape != shape for a in arrs):
        raise ValueError(f"field_write: fields must have shape {shape}")
    prec = TH_FP64 if dt == np.complex128 else TH_FP32_MIXED
    n = (C.c_int * 3)(*N)
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    _check(lib().th_field_write(str(path).encode(), n, prec, len(arrs), ptrs, float(time), 1 if append else 0),
           "th_field_write")


def field_read(path, N, ncomp, snapshot=0, precision="fp64"):
    """th_field_read: snapshot `snapshot` of a field file; returns (list of ncomp host arrays, time)."""
    import numpy as np
    dt = np.complex128 if precision == "fp64" else np.complex64
    arrs = [np.zeros((N[0], N[1], N[2] // 2 + 1), dt) for _ in range(ncomp)]
    n = (C.c_int * 3)(*N)
    ptrs = (C.c_void_p * ncomp)(*[a.ctypes.data for a in arrs])
    t = C.c_double()
    _check(lib().th_field_read(str(path).encode(), n, TH_FP64 if precision == "fp64" else TH_FP32_MIXED, ncomp, ptrs,
                               int(snapshot), C.byref(t)), "th_field_read")
    return arrs, t.value


def field_write_reduced(path, N, Nr, fields, time=0.0, append=False, basis="FOUR"):
    """th_field_write_reduced: one snapshot of the reference's reduced field file (field_out_reduced.d of
    Output_field_reduced): the modes of the Nr grid out of fields of shape (N1, N2, N3/2+1)."""
    import numpy as np
    arrs = _host_arrays(fields)
    dt = arrs[0].dtype
    if dt not in (np.complex128, np.complex64) or any(a.dtype != dt for a in arrs):
        raise TypeError("field_write_reduced: all fields complex128 or all complex64")
    shape = (N[0], N[1], N[2] // 2 + 1)
    if any(a.shape != shape for a in arrs):
        raise ValueError(f"field_write_reduced: fields must have shape {shape}")
    prec = TH_FP64 if dt == np.complex128 else TH_FP32_MIXED
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    _check(lib().th_field_write_reduced(str(path).encode(), (C.c_int * 3)(*N), (C.c_int * 3)(*Nr),
                                        TH_FOUR if basis == "FOUR" else TH_SCFT, prec, len(arrs), ptrs, float(time),
                                        1 if append else 0), "th_field_write_reduced")


def field_read_reduced(path, N, Nr, ncomp, snapshot=0, precision="fp64", basis="FOUR"):
    """th_field_read_reduced: snapshot `snapshot` of a reduced field file of grid Nr into arrays of grid N;
    returns (list of ncomp host arrays, time)."""
    import numpy as np
    dt = np.complex128 if precision == "fp64" else np.complex64
    arrs = [np.zeros((N[0], N[1], N[2] // 2 + 1), dt) for _ in range(ncomp)]
    ptrs = (C.c_void_p * ncomp)(*[a.ctypes.data for a in arrs])
    t = C.c_double()
    _check(lib().th_field_read_reduced(str(path).encode(), (C.c_int * 3)(*N), (C.c_int * 3)(*Nr),
                                       TH_FOUR if basis == "FOUR" else TH_SCFT,
                                       TH_FP64 if precision == "fp64" else TH_FP32_MIXED, ncomp, ptrs, int(snapshot),
                                       C.byref(t)), "th_field_read_reduced")
    return arrs, t.value


class Context:
    def __init__(self, device=0, tune_cache=None):
        self.h = C.c_void_p()
        _check(lib().th_ctx_create(C.byref(self.h), device,
                                   tune_cache.encode() if tune_cache else None), "th_ctx_create")
        self.device = device
        _live_ctx.add(self)

    def close(self):
        if self.h and not _shutting_down:
            if lib().th_ctx_destroy(self.h) == 0:
                self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dft16_probe(self, x, y, method, sign=-1, reps=1, ntimed=0):
        """th_dft16_probe on complex64 device tensors of shape (16, nvec): y = (W / 4)^reps x by the VALU
        codelet (method 0) or the MFMA DFT-as-GEMM (method 1); returns the mean ms of ntimed launches."""
        assert x.shape == y.shape and x.shape[0] == 16 and x.is_contiguous() and y.is_contiguous()
        t = _torch()
        t.cuda.synchronize(self.device)
        ms = C.c_double(0.0)
        _check(lib().th_dft16_probe(self.h, int(method), int(sign), _ptr(x), _ptr(y), int(x.shape[1]), int(reps),
                                    int(ntimed), C.byref(ms)), "th_dft16_probe")
        t.cuda.synchronize(self.device)
        return ms.value if ntimed > 0 else None


class Plan:
    """One box: basis, grid, kfactor, precision, dealias switch."""

    def __init__(self, ctx, N, kfactor=(1.0, 1.0, 1.0), basis="FOUR", precision="fp64", dealias=True):
        self.ctx = ctx
        self.N = tuple(int(n) for n in N)
        self.basis = TH_FOUR if basis == "FOUR" else TH_SCFT
        self.prec = TH_FP64 if precision == "fp64" else TH_FP32_MIXED
        kf = (C.c_double * 3)(*[float(v) for v in kfactor])
        self.h = C.c_void_p()
        _check(lib().th_plan_create(ctx.h, C.byref(self.h), self.basis, *self.N, kf, self.prec,
                                    1 if dealias else 0), "th_plan_create")
        self._owned = True
        _live_plans.add(self)

    @classmethod
    def _borrowed(cls, ctx, handle, N, prec):
        """A plan owned by the library (an ensemble's size class); close() leaves it alone."""
        p = cls.__new__(cls)
        p.ctx, p.N, p.basis, p.prec, p.h, p._owned = ctx, tuple(N), TH_FOUR, prec, handle, False
        return p

    def close(self):
        if self.h and self._owned and not _shutting_down:
            for st in list(_live_states):
                if st.plan is self:
                    st.close()
            lib().th_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.N[0], self.N[1], self.N[2] // 2 + 1)

    @property
    def dtype(self):
        torch = _torch()
        return torch.complex128 if self.prec == TH_FP64 else torch.complex64

    def empty(self, n=None):
        torch = _torch()
        dev = f"cuda:{self.ctx.device}"
        if n is None:
            return torch.zeros(self.shape, dtype=self.dtype, device=dev)
        return [torch.zeros(self.shape, dtype=self.dtype, device=dev) for _ in range(n)]

    def to_dev(self, arr):
        torch = _torch()
        t = torch.from_numpy(arr).to(device=f"cuda:{self.ctx.device}", dtype=self.dtype).contiguous()
        return t

    def _pre(self):
        _torch().cuda.synchronize(self.ctx.device)

    def _post(self, rc, what):
        _check(rc, what)
        _check(lib().th_sync(self.h), "th_sync")

    # transforms and ops
    def forward(self, A, parity=0):
        self._pre()
        self._post(lib().th_forward(self.h, _ptr(A), parity), "th_forward")
        return A

    def inverse(self, A, parity=0):
        self._pre()
        self._post(lib().th_inverse(self.h, _ptr(A), parity), "th_inverse")
        return A

    def deriv(self, axis, A, parity=0, out=None):
        B = out if out is not None else _torch().empty_like(A)
        self._pre()
        self._post(lib().th_deriv(self.h, axis, _ptr(A), _ptr(B), parity), "th_deriv")
        return B

    def dealias(self, A):
        self._pre()
        self._post(lib().th_dealias(self.h, _ptr(A)), "th_dealias")
        return A

    def real_mult(self, A, B):
        Cc = _torch().empty_like(A)
        self._pre()
        self._post(lib().th_real_mult(self.h, _ptr(A), _ptr(B), _ptr(Cc)), "th_real_mult")
        return Cc

    def divide_ksqr(self, A):
        self._pre()
        self._post(lib().th_divide_ksqr(self.h, _ptr(A)), "th_divide_ksqr")
        return A

    def exp_ksqr(self, A, factor):
        self._pre()
        self._post(lib().th_exp_ksqr(self.h, _ptr(A), factor), "th_exp_ksqr")
        return A

    def nlin_ns(self, V):
        N = self.empty(3)
        self._pre()
        self._post(lib().th_nlin_ns(self.h, _pp(V), _pp(N)), "th_nlin_ns")
        return N

    def pressure(self, nl):
        p = self.empty()
        self._pre()
        self._post(lib().th_pressure(self.h, _pp(nl), _ptr(p)), "th_pressure")
        return p

    def rhs(self, nl, p):
        self._pre()
        self._post(lib().th_rhs(self.h, _pp(nl), _ptr(p)), "th_rhs")
        return nl

    def nlin_mhd(self, U, B):
        nU, nB = self.empty(3), self.empty(3)
        self._pre()
        self._post(lib().th_nlin_mhd(self.h, _pp(U), _pp(B), _pp(nU), _pp(nB)), "th_nlin_mhd")
        return nU, nB

    def nlin_rb(self, V, T):
        nV, nT = self.empty(3), self.empty()
        self._pre()
        self._post(lib().th_nlin_rb(self.h, _pp(V), _ptr(T), _pp(nV), _ptr(nT)), "th_nlin_rb")
        return nV, nT

    def et_nlin(self, H, G):
        N = self.empty(3)
        self._pre()
        self._post(lib().th_et_nlin(self.h, _pp(H), _pp(G), _pp(N)), "th_et_nlin")
        return N

    def step_plain(self, physics, fields, dt, nu=0.0, eta=0.0, kappa=0.0, Ra=0.0, Pr=0.0, scheme=2,
                   steps=1):
        ph = Phys(nu, eta, kappa, Ra, Pr)
        self._pre()
        for _ in range(steps):
            _check(lib().th_step_plain(self.h, physics, scheme, _pp(fields), dt, C.byref(ph)),
                   "th_step_plain")
        self._post(0, "")
        return fields

    def step_plain_forced(self, V, dt, nu, proc, inner, outer, p3, p4=0.0, scheme=2, steps=1, last_stage=None,
                          p5=0.0):
        """NS steps with the force_field_proc `proc` forcing recomputed at every stage; last_stage (3 device
        arrays, optional) receives the field of the last stage."""
        ph = Phys(nu, 0.0, 0.0, 0.0, 0.0)
        fc = Force(proc, inner, outer, p3, p4, p5)
        self._pre()
        for _ in range(steps):
            _check(lib().th_step_plain_forced(self.h, scheme, _pp(V), dt, C.byref(ph), C.byref(fc),
                                              _pp(last_stage) if last_stage is not None else None),
                   "th_step_plain_forced")
        self._post(0, "")
        return V

    def step_plain_forced_mhd(self, U, B, dt, nu, eta, inner, outer, p3, p4=0.0, p5=0.0, p6=0.0, scheme=2,
                              steps=1, last_stage=None, proc=1):
        """MHD steps with the force_field_proc `proc` forcing of U (p3, p4) and B (p5, p6) recomputed at every
        stage; last_stage (6 device arrays U1..U3, B1..B3, optional) receives the fields of the last stage."""
        ph = Phys(nu, eta, 0.0, 0.0, 0.0)
        fc = Force(proc, inner, outer, p3, p4, p5, p6)
        self._pre()
        for _ in range(steps):
            _check(lib().th_step_plain_forced_mhd(self.h, scheme, _pp(U), _pp(B), dt, C.byref(ph), C.byref(fc),
                                                  _pp(last_stage) if last_stage is not None else None),
                   "th_step_plain_forced_mhd")
        self._post(0, "")
        return U, B

    def entropy_scalar(self, A):
        out = C.c_double()
        self._pre()
        _check(lib().th_entropy_scalar(self.h, _ptr(A), C.byref(out)), "th_entropy_scalar")
        return out.value

    def ring_spectrum(self, V, nsect, n=0, axis=3):
        """(S1k, S2k) of Compute_ring_spectrum_FOUR, shape (Max_radius_inside+1, nsect+1)."""
        import numpy as np
        shape = (lib().th_max_radius_inside(self.h) + 1, nsect + 1)
        S1, S2 = np.zeros(shape), np.zeros(shape)
        self._pre()
        _check(lib().th_ring_spectrum(self.h, _pp(V), nsect, n, axis, S1.ctypes.data_as(C.POINTER(C.c_double)),
                                      S2.ctypes.data_as(C.POINTER(C.c_double))), "th_ring_spectrum")
        return S1, S2

    def cylinder_ring_spectrum(self, V, nslab, n=0, axis=3):
        """(S1k, S2k) of Compute_cylinder_ring_spectrum_FOUR, shape (Anis_max_Krho_radius_inside+1, nslab+1)."""
        import numpy as np
        rows = lib().th_cylinder_shell_ek_size(self.h, axis)
        _check(0 if rows > 0 else rows, "th_cylinder_shell_ek_size")
        S1, S2 = np.zeros((rows, nslab + 1)), np.zeros((rows, nslab + 1))
        self._pre()
        _check(lib().th_cylinder_ring_spectrum(self.h, _pp(V), nslab, n, axis, S1.ctypes.data_as(C.POINTER(C.c_double)),
                                               S2.ctypes.data_as(C.POINTER(C.c_double))), "th_cylinder_ring_spectrum")
        return S1, S2

    def ring_spectrum_helicity(self, V, nsect, axis=3):
        """H1k of Compute_ring_spectrum_helicity_FOUR, shape (Max_radius_inside+1, nsect+1)."""
        import numpy as np
        H = np.zeros((lib().th_max_radius_inside(self.h) + 1, nsect + 1))
        self._pre()
        _check(lib().th_ring_spectrum_helicity(self.h, _pp(V), nsect, axis, H.ctypes.data_as(C.POINTER(C.c_double))),
               "th_ring_spectrum_helicity")
        return H

    def cylinder_ring_spectrum_helicity(self, V, nslab, axis=3):
        """H1k of Compute_cylinder_ring_spectrum_helicity_FOUR, shape (Anis_max_Krho_radius_inside+1, nslab+1)."""
        import numpy as np
        rows = lib().th_cylinder_shell_ek_size(self.h, axis)
        _check(0 if rows > 0 else rows, "th_cylinder_shell_ek_size")
        H = np.zeros((rows, nslab + 1))
        self._pre()
        _check(lib().th_cylinder_ring_spectrum_helicity(self.h, _pp(V), nslab, axis,
                                                        H.ctypes.data_as(C.POINTER(C.c_double))),
               "th_cylinder_ring_spectrum_helicity")
        return H

    def compute_force(self, V, proc, inner, outer, p3, p4=0.0, p5=0.0):
        """Force field F[3] of force_field_proc `proc` for the velocity V, and nf."""
        F = self.empty(3)
        fc = Force(proc, inner, outer, p3, p4, p5)
        nf = C.c_long()
        self._pre()
        self._post(lib().th_compute_force(self.h, _pp(V), _pp(F), C.byref(fc), C.byref(nf)), "th_compute_force")
        return F, nf.value

    def compute_force_mhd(self, U, B, proc, inner, outer, p3, p4, p5, p6):
        """Forces (FU[3], FW[3]) of an MHD run (force_field_proc `proc`), and nf."""
        FU, FW = self.empty(3), self.empty(3)
        fc = Force(proc, inner, outer, p3, p4, p5, p6)
        nf = C.c_long()
        self._pre()
        self._post(lib().th_compute_force_mhd(self.h, _pp(U), _pp(B), _pp(FU), _pp(FW), C.byref(fc), C.byref(nf)),
                   "th_compute_force_mhd")
        return FU, FW, nf.value

    # diagnostics
    def energy(self, A, B=None):
        out = C.c_double()
        self._pre()
        _check(lib().th_energy(self.h, _ptr(A), _ptr(B) if B is not None else None, C.byref(out)),
               "th_energy")
        return out.value

    def Sn(self, A, n):
        out = C.c_double()
        self._pre()
        _check(lib().th_Sn(self.h, _ptr(A), float(n), C.byref(out)), "th_Sn")
        return out.value

    def total_energy(self, V):
        return sum(self.energy(a) for a in V)

    def total_dissipation(self, V):
        return 2 * sum(self.Sn(a, 2) for a in V)

    def shell_spectrum(self, A, B=None, n=0):
        import numpy as np
        size = lib().th_shell_ek_size(self.h)
        Sk = np.zeros(size)
        cnt = np.zeros(size)
        self._pre()
        _check(lib().th_shell_spectrum(self.h, _ptr(A), _ptr(B) if B is not None else None, int(n),
                                       Sk.ctypes.data_as(C.POINTER(C.c_double)),
                                       cnt.ctypes.data_as(C.POINTER(C.c_double))), "th_shell_spectrum")
        return Sk, cnt

    def vector_spectrum(self, V, n=0):
        return sum(self.shell_spectrum(a, None, n)[0] for a in V)

    def fill_shell(self, A, inner, outer):
        B = _torch().empty_like(A)
        self._pre()
        self._post(lib().th_fill_array_shell(self.h, _ptr(A), _ptr(B), inner, outer), "th_fill_array_shell")
        return B

    def shell_mult_single(self, A, B, inner, outer):
        out = C.c_double()
        self._pre()
        _check(lib().th_shell_mult_single(self.h, _pp(A), _pp(B), inner, outer, C.byref(out)),
               "th_shell_mult_single")
        return out.value

    def shell_mult_all(self, A, B, radii):
        import numpy as np
        r, rp = _dbl(radii)
        out = np.zeros(len(r))
        self._pre()
        _check(lib().th_shell_mult_all(self.h, _pp(A), _pp(B), rp, len(r),
                                       out.ctypes.data_as(C.POINTER(C.c_double))), "th_shell_mult_all")
        return out

    def flux_ns(self, V, radii):
        import numpy as np
        r, rp = _dbl(radii)
        flux = np.zeros(len(r))
        self._pre()
        _check(lib().th_flux_ns(self.h, _pp(V), rp, len(r) - 1, flux.ctypes.data_as(C.POINTER(C.c_double))),
               "th_flux_ns")
        return flux

    def flux_mhd(self, U, B, radii):
        """Rows: u2u, b2b, u2b_io, u2b_ii, b2u_io, b2u_ii, uo_bo, zp, zm (each over spheres 0..nsph)."""
        import numpy as np
        r, rp = _dbl(radii)
        out = np.zeros((9, len(r)))
        self._pre()
        _check(lib().th_flux_mhd(self.h, _pp(U), _pp(B), rp, len(r) - 1,
                                 out.ctypes.data_as(C.POINTER(C.c_double))), "th_flux_mhd")
        return out

    def helicity(self, V):
        """(H1, H2, DH1, DH2): Compute_total_helicity_FOUR."""
        import numpy as np
        out = np.zeros(4)
        self._pre()
        _check(lib().th_helicity(self.h, _pp(V), out.ctypes.data_as(C.POINTER(C.c_double))), "th_helicity")
        return out

    def shell_helicity(self, V):
        """(3, shell_ek_size) helicity spectra H1k1..3: Compute_shell_spectrum_helicity_FOUR."""
        import numpy as np
        H = np.zeros((3, lib().th_shell_ek_size(self.h)))
        self._pre()
        _check(lib().th_shell_spectrum_helicity(self.h, _pp(V), *[H[c].ctypes.data_as(C.POINTER(C.c_double))
                                                                 for c in range(3)]), "th_shell_spectrum_helicity")
        return H

    def entropy(self, V):
        out = C.c_double()
        self._pre()
        _check(lib().th_entropy(self.h, _pp(V), C.byref(out)), "th_entropy")
        return out.value

    def flux_scalar(self, V, T, radii):
        """(flux_self, flux_SF) over spheres 0..nsph (Compute_flux_scalar / Compute_flux_RB)."""
        import numpy as np
        r, rp = _dbl(radii)
        out = np.zeros((2, len(r)))
        self._pre()
        _check(lib().th_flux_scalar(self.h, _pp(V), _ptr(T), rp, len(r) - 1,
                                    out.ctypes.data_as(C.POINTER(C.c_double))), "th_flux_scalar")
        return out[0], out[1]

    def shell_tr_scalar(self, V, T, radii):
        """(T_self, T_SF), each (nsh+1)^2 (Compute_shell_tr_scalar / Compute_shell_tr_RB)."""
        import numpy as np
        r, rp = _dbl(radii)
        nsh = len(r) - 1
        Ts, Tt = np.zeros((nsh + 1, nsh + 1)), np.zeros((nsh + 1, nsh + 1))
        self._pre()
        _check(lib().th_shell_tr_scalar(self.h, _pp(V), _ptr(T), rp, nsh, Ts.ctypes.data_as(C.POINTER(C.c_double)),
                                        Tt.ctypes.data_as(C.POINTER(C.c_double))), "th_shell_tr_scalar")
        return Ts, Tt

    def shell_tr_mhd(self, U, B, radii):
        """The five (nsh+1)^2 matrices of Compute_shell_tr(IncVF& W): uu, ww, uw, zp, zm."""
        import numpy as np
        r, rp = _dbl(radii)
        nsh = len(r) - 1
        out = np.zeros((5, nsh + 1, nsh + 1))
        self._pre()
        _check(lib().th_shell_tr_mhd(self.h, _pp(U), _pp(B), rp, nsh, out.ctypes.data_as(C.POINTER(C.c_double))),
               "th_shell_tr_mhd")
        return dict(zip(["uu", "ww", "uw", "zp", "zm"], out))

    def ring_tr(self, H, F, R, radii, edges, ngiv=None, sign=-1.0, cyl=False, axis=3):
        """th_ring_tr: [giver shell, giver sector, receiver shell, receiver sector] (see include/tarang_hip.h)."""
        import numpy as np
        r, rp = _dbl(radii)
        e, ep = _dbl(edges)
        nr, nsec = len(r) - 1, len(e) - 1
        ngiv = nr - 1 if ngiv is None else ngiv
        out = np.zeros((nr + 1, nsec + 1, nr + 1, nsec + 1))
        pad = lambda X: list(X) + [X[0]] * (3 - len(X))
        self._pre()
        _check(lib().th_ring_tr(self.h, int(cyl), axis, _pp(H), _pp(pad(F)), _pp(pad(R)), len(F), rp, nr, ep, nsec,
                                ngiv, float(sign), out.ctypes.data_as(C.POINTER(C.c_double))), "th_ring_tr")
        return out

    def shell_tr_ns(self, V, radii):
        import numpy as np
        r, rp = _dbl(radii)
        nsh = len(r) - 1
        T = np.zeros((nsh + 1, nsh + 1))
        self._pre()
        _check(lib().th_shell_tr_ns(self.h, _pp(V), rp, nsh, T.ctypes.data_as(C.POINTER(C.c_double))),
               "th_shell_tr_ns")
        return T

    def nusselt(self, V, T, Ra, Pr):
        out = C.c_double()
        self._pre()
        _check(lib().th_nusselt(self.h, _pp(V), _ptr(T), Ra, Pr, C.byref(out)), "th_nusselt")
        return out.value

    # initial conditions
    def ic_taylor_green(self, k0=1, amp=1.0):
        V = self.empty(3)
        self._pre()
        self._post(lib().th_ic_taylor_green(self.h, _pp(V), k0, amp), "th_ic_taylor_green")
        return V

    def ic_random(self, seed, E0=0.5, kp=4.0):
        V = self.empty(3)
        self._pre()
        self._post(lib().th_ic_random(self.h, _pp(V), seed, E0, kp), "th_ic_random")
        return V

    def ic_random_scalar(self, seed, amp=1e-3):
        T = self.empty()
        self._pre()
        self._post(lib().th_ic_random_scalar(self.h, _ptr(T), seed, amp), "th_ic_random_scalar")
        return T


_live_states = weakref.WeakSet()


class State:
    """Device-resident fused-layout fields of one box (th_state)."""

    def __init__(self, plan, physics=TH_NS):
        self.plan = plan
        self.physics = physics
        self.h = C.c_void_p()
        _check(lib().th_state_create(plan.h, C.byref(self.h), physics), "th_state_create")
        self.ncomp = {TH_NS: 3, TH_MHD: 6, TH_RB: 4}[physics]
        _live_states.add(self)

    def close(self):
        if self.h and not _shutting_down:
            lib().th_state_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, fields):
        self.plan._pre()
        for c, f in enumerate(fields):
            _check(lib().th_state_upload(self.h, c, _ptr(f), 0), "th_state_upload")
        self.plan._post(0, "")

    def download(self):
        out = self.plan.empty(self.ncomp)
        self.plan._pre()
        for c, f in enumerate(out):
            _check(lib().th_state_download(self.h, c, _ptr(f), 0), "th_state_download")
        self.plan._post(0, "")
        return out

    def download_buffer(self, which):
        """Diagnostics: 1 = RK stage field, 2 = RK4 accumulator (reference layout)."""
        out = self.plan.empty(self.ncomp)
        self.plan._pre()
        for c, f in enumerate(out):
            _check(lib().th_state_download_buffer(self.h, which, c, _ptr(f), 0), "th_state_download_buffer")
        self.plan._post(0, "")
        return out

    def step(self, dt, nu=0.0, eta=0.0, kappa=0.0, Ra=0.0, Pr=0.0, steps=1):
        ph = Phys(nu, eta, kappa, Ra, Pr)
        self.plan._pre()
        self.plan._post(lib().th_step(self.plan.h, self.h, dt, C.byref(ph), steps), "th_step")

    def set_force(self, proc=0, inner=0.0, outer=0.0, p3=0.0, p4=0.0):
        """Forcing (force_field_proc 1 / 2) of later steps; proc 0 switches it off."""
        fc = Force(proc, inner, outer, p3, p4)
        _check(lib().th_state_set_force(self.h, C.byref(fc)), "th_state_set_force")

    def profile_passes(self, dt, nu=0.0, eta=0.0, stage=2, nrep=5, kappa=0.0, Ra=0.0, Pr=0.0):
        import numpy as np
        ph = Phys(nu, eta, kappa, Ra, Pr)
        ms = np.zeros(5)
        b = np.zeros(5)
        self.plan._pre()
        _check(lib().th_profile_passes(self.plan.h, self.h, dt, C.byref(ph), stage, nrep,
                                       ms.ctypes.data_as(C.POINTER(C.c_double)),
                                       b.ctypes.data_as(C.POINTER(C.c_double))), "th_profile_passes")
        return ms, b

    def tune(self, dt, nu=0.0, eta=0.0, nrep=3, kappa=0.0, Ra=0.0, Pr=0.0):
        """Autotune the launch shapes of the five fused passes; returns the chosen variants."""
        ph = Phys(nu, eta, kappa, Ra, Pr)
        v = (C.c_int * 5)()
        self.plan._pre()
        self.plan._post(lib().th_tune(self.plan.h, self.h, dt, C.byref(ph), nrep, v), "th_tune")
        return list(v)

    @property
    def variants(self):
        v = (C.c_int * 5)()
        _check(lib().th_state_variants(self.h, v), "th_state_variants")
        return list(v)

    @variants.setter
    def variants(self, var):
        v = (C.c_int * 5)(*[int(x) for x in var])
        self.plan._pre()
        _check(lib().th_state_set_variants(self.h, v), "th_state_set_variants")

    def flux_mhd(self, radii):
        """th_flux_mhd_fused: the nine MHD flux rows (th_flux_mhd's order) of the state's U and B."""
        import numpy as np
        r, rp = _dbl(radii)
        out = np.zeros((9, len(r)))
        self.plan._pre()
        _check(lib().th_flux_mhd_fused(self.plan.h, self.h, rp, len(r) - 1, out.ctypes.data_as(C.POINTER(C.c_double))),
               "th_flux_mhd_fused")
        return out

    def flux_shell_tr(self, radii):
        import numpy as np
        r, rp = _dbl(radii)
        nsh = len(r) - 1
        Pi = np.zeros(nsh + 1)
        T = np.zeros((nsh + 1, nsh + 1))
        self.plan._pre()
        _check(lib().th_flux_shell_tr(self.plan.h, self.h, rp, nsh,
                                      Pi.ctypes.data_as(C.POINTER(C.c_double)),
                                      T.ctypes.data_as(C.POINTER(C.c_double))), "th_flux_shell_tr")
        return Pi, T


class Ensemble:
    """Independent NS boxes of uneven sizes stepped together (th_ensemble)."""

    def __init__(self, ctx, shapes, precision="fp32"):
        self.ctx = ctx
        self.shapes = [tuple(int(n) for n in N) for N in shapes]
        self.prec = TH_FP64 if precision == "fp64" else TH_FP32_MIXED
        flat = (C.c_int * (3 * len(self.shapes)))(*[n for N in self.shapes for n in N])
        self.h = C.c_void_p()
        _check(lib().th_ensemble_create(ctx.h, C.byref(self.h), len(self.shapes), flat, self.prec),
               "th_ensemble_create")
        self._plans = {}
        _live_ens.add(self)

    def close(self):
        if self.h and not _shutting_down:
            lib().th_ensemble_destroy(self.h)
            self.h = C.c_void_p()
            self._plans = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, box):
        """The library-owned plan of box `box`'s size (ICs, spectra on reference-layout arrays)."""
        N = self.shapes[box]
        if N not in self._plans:
            h = C.c_void_p()
            _check(lib().th_ensemble_plan(self.h, box, C.byref(h)), "th_ensemble_plan")
            self._plans[N] = Plan._borrowed(self.ctx, h, N, self.prec)
        return self._plans[N]

    def upload(self, box, fields):
        p = self.plan(box)
        p._pre()
        for c, f in enumerate(fields):
            _check(lib().th_ensemble_upload(self.h, box, c, _ptr(f), 0), "th_ensemble_upload")
        p._post(0, "")

    def download(self, box):
        p = self.plan(box)
        out = p.empty(3)
        p._pre()
        for c, f in enumerate(out):
            _check(lib().th_ensemble_download(self.h, box, c, _ptr(f), 0), "th_ensemble_download")
        p._post(0, "")
        return out

    def tune(self, dt, nu, nrep=3):
        dt_, dtp = _dbl(dt)
        nu_, nup = _dbl(nu)
        p = self.plan(0)
        p._pre()
        p._post(lib().th_ensemble_tune(self.h, dtp, nup, nrep), "th_ensemble_tune")

    def variants(self, box):
        """Launch variants of box `box`'s size class (x-inverse, y-inverse, z, y-forward, x-forward)."""
        v = (C.c_int * 5)()
        _check(lib().th_ensemble_variants(self.h, int(box), v), "th_ensemble_variants")
        return list(v)

    def step(self, dt, nu, steps=1, sync=True):
        dt_, dtp = _dbl(dt)
        nu_, nup = _dbl(nu)
        if len(dt_) != len(self.shapes) or len(nu_) != len(self.shapes):
            raise ValueError("dt and nu need one value per box")
        p = self.plan(0)
        p._pre()
        _check(lib().th_ensemble_step(self.h, dtp, nup, steps), "th_ensemble_step")
        if sync:
            _check(lib().th_sync(p.h), "th_sync")
