"""Generate the small golden fixtures under tests/golden/ with numpy / scipy (run once, committed).

These vectors pin the oracle independently of the oracle's own code:
  * four_fft.npz   -- numpy rfftn / irfftn with Tarang's scaling (forward / N1N2N3, inverse
                      unnormalised), the convention of ArrayFFT_FOUR / ArrayIFFT_FOUR
                      (trunk/lib/four/four_tr.cc:199-295).
  * scft_fft.npz   -- scipy.fft dct/dst types 2 and 3 (= FFTW REDFT10/RODFT10/REDFT01/RODFT01 by
                      FFTW's documented definitions) with Tarang's 1/(2 N1 N2 N3) norm and sine
                      shift (trunk/lib/scft/scft_tr.cc:89-219), combined with numpy rfft2 over (y,z).
  * kat_out1.npz   -- the only recorded known answer of the reference: trunk/test/test-fftw/out1.d
                      (4^3, unit Nyquist-z mode: sum|u|^2 = 80 incl. 16 from the padding slot,
                      so 64 from the data; r2c gives (64,0)).
  * nlin_triad.npz -- exact truncated convolution N_i(k) = sum_j i k_j sum_{p+q=k} u_i(p) u_j(q)
                      for a sparse seeded field, computed by explicit double loops over modes
                      (brute force, no FFT), for the divergence form of compute_nlin.cc:58-108.
  * et_triad.npz   -- energy fluxes and shell-to-shell transfers (NS, MHD with Elsasser fluxes,
                      passive scalar) of band-limited seeded fields at 12^3 as explicit triad sums over
                      the full wavenumber lattice (et_triad below).
Usage: python tests/golden/make_golden.py  (rewrites every fixture; each generator is deterministic)
"""
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def four_fft():
    out = {}
    rng = np.random.default_rng(11)
    for tag, N in [("a", (8, 8, 8)), ("b", (6, 10, 8)), ("c", (4, 4, 4))]:
        N1, N2, N3 = N
        u = rng.standard_normal(N)
        out[f"{tag}_N"] = np.array(N)
        out[f"{tag}_u"] = u
        out[f"{tag}_fwd"] = np.fft.rfftn(u) / (N1 * N2 * N3)
        C = rng.standard_normal((N1, N2, N3 // 2 + 1)) + 1j * rng.standard_normal((N1, N2, N3 // 2 + 1))
        out[f"{tag}_spec"] = C
        out[f"{tag}_inv"] = np.fft.irfftn(C, s=N, axes=(0, 1, 2)) * (N1 * N2 * N3)
    np.savez_compressed(os.path.join(HERE, "four_fft.npz"), **out)


def scft_fft():
    out = {}
    rng = np.random.default_rng(12)
    for tag, N in [("a", (8, 8, 8)), ("b", (12, 6, 8))]:
        N1, N2, N3 = N
        u = rng.standard_normal(N)
        out[f"{tag}_N"] = np.array(N)
        out[f"{tag}_u"] = u
        for par in (0, 1):
            t = sf.dct(u, type=2, axis=0) if par == 0 else sf.dst(u, type=2, axis=0)
            ref = np.fft.rfftn(t, axes=(1, 2)) / (2 * N1 * N2 * N3)
            if par == 1:
                ref = np.concatenate([np.zeros_like(ref[:1]), ref[:-1]], axis=0)
            out[f"{tag}_fwd{par}"] = ref
            C = rng.standard_normal((N1, N2, N3 // 2 + 1)) + 1j * rng.standard_normal((N1, N2, N3 // 2 + 1))
            if par == 1:
                C[0] = 0.0
            out[f"{tag}_spec{par}"] = C
            D = C.copy()
            if par == 1:
                D = np.concatenate([D[1:], np.zeros_like(D[:1])], axis=0)
            r = np.fft.irfftn(D, s=(N2, N3), axes=(1, 2)) * (N2 * N3)
            inv = sf.dct(r, type=3, axis=0) if par == 0 else sf.dst(r, type=3, axis=0)
            out[f"{tag}_inv{par}"] = inv
    np.savez_compressed(os.path.join(HERE, "scft_fft.npz"), **out)


def kat_out1():
    # trunk/test/test-fftw/out1.d: "Total Before IFFT: 1", "AFTER IFFT: tot 80", "AFTER FFT ... (64,0)"
    np.savez_compressed(os.path.join(HERE, "kat_out1.npz"), N=np.array([4, 4, 4]),
                        mode=np.array([0, 0, 2]), tot_before=1.0, tot_after_c2r_incl_padding=80.0,
                        padding_share=16.0, r2c_value=np.array([64.0, 0.0]))


def full_spectrum(A, N):
    """Hermitian extension of a half spectrum (N1,N2,N3/2+1) to a dict {(kx,ky,kz): value}."""
    N1, N2, N3 = N
    modes = {}
    for l1 in range(N1):
        kx = l1 if l1 <= N1 // 2 else l1 - N1
        for l2 in range(N2):
            ky = l2 if l2 <= N2 // 2 else l2 - N2
            for l3 in range(N3 // 2 + 1):
                v = A[l1, l2, l3]
                if v == 0:
                    continue
                modes[(kx, ky, l3)] = v
                if 0 < l3 < N3 // 2:
                    modes[(-kx, -ky, -l3)] = np.conj(v)
    return modes


def nlin_triad():
    N = (12, 12, 12)
    N1, N2, N3 = N
    rng = np.random.default_rng(13)
    cut = N1 // 3          # kept |k_i| <= cut - 1 (Dealias_array, universal_misc.cc:118-159)
    V = [np.zeros((N1, N2, N3 // 2 + 1), complex) for _ in range(3)]
    picks = set()
    while len(picks) < 10:
        kx, ky, kz = rng.integers(-(cut - 1), cut, size=3)
        if kz <= 0:
            continue
        picks.add((int(kx), int(ky), int(kz)))
    for (kx, ky, kz) in picks:
        k = np.array([kx, ky, kz], float)
        a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        a = a - k * (k @ a) / (k @ k)        # divergence free
        for i in range(3):
            V[i][kx % N1, ky % N2, kz] = a[i]
    full = [full_spectrum(V[i], N) for i in range(3)]
    keys = set().union(*[f.keys() for f in full])
    Nk = [np.zeros((N1, N2, N3 // 2 + 1), complex) for _ in range(3)]
    for p in keys:
        for q in keys:
            k = (p[0] + q[0], p[1] + q[1], p[2] + q[2])
            if k[2] < 0 or max(abs(k[0]), abs(k[1]), abs(k[2])) > cut - 1:
                continue          # only kept output modes are compared
            up = [full[i].get(p, 0) for i in range(3)]
            uq = [full[i].get(q, 0) for i in range(3)]
            for i in range(3):
                s = sum(1j * k[j] * up[i] * uq[j] for j in range(3))
                Nk[i][k[0] % N1, k[1] % N2, k[2]] += s
    np.savez_compressed(os.path.join(HERE, "nlin_triad.npz"), N=np.array(N),
                        V=np.stack(V), nlin=np.stack(Nk))


def _band_field(rng, N, kmax, vector):
    """Real random field band-limited to |kx|, |kz| <= kmax[0], kmax[2] and |ky| <= kmax[1] (a set that is
    symmetric under k -> -k, so the half spectrum stays Hermitian on kz = 0), k = 0 removed, vector fields
    projected to K.V = 0; half spectrum (N1, N2, N3/2+1) with Tarang's forward scaling 1/(N1 N2 N3)."""
    N1, N2, N3 = N
    kx = np.fft.fftfreq(N1, 1.0 / N1).reshape(-1, 1, 1)
    ky = np.fft.fftfreq(N2, 1.0 / N2).reshape(1, -1, 1)
    kz = np.arange(N3 // 2 + 1).reshape(1, 1, -1)
    sup = (np.abs(kx) <= kmax[0]) & (np.abs(ky) <= kmax[1]) & (kz <= kmax[2])
    sup = sup & ~((kx == 0) & (ky == 0) & (kz == 0))
    comps = [np.fft.rfftn(rng.standard_normal(N)) / (N1 * N2 * N3) * sup for _ in range(3 if vector else 1)]
    if vector:
        K2 = (kx * kx + ky * ky + kz * kz).astype(float)
        K2[0, 0, 0] = 1.0
        div = kx * comps[0] + ky * comps[1] + kz * comps[2]
        comps = [c - k * div / K2 for c, k in zip(comps, (kx, ky, kz))]
    return comps


def _full_grid(A, N, M):
    """Half spectrum -> full-lattice array on an M^3 grid indexed by k mod M (M > twice the support, so
    sums of two support vectors never wrap)."""
    N1, N2, N3 = N
    F = np.zeros((M, M, M), complex)
    for l1 in range(N1):
        kx = l1 if l1 <= N1 // 2 else l1 - N1
        for l2 in range(N2):
            ky = l2 if l2 <= N2 // 2 else l2 - N2
            for l3 in range(N3 // 2 + 1):
                v = A[l1, l2, l3]
                if v == 0:
                    continue
                F[kx % M, ky % M, l3 % M] = v
                F[-kx % M, -ky % M, -l3 % M] = np.conj(v)
    return F


def et_triad():
    """Energy transfer by explicit triad sums over the full wavenumber lattice (no FFT, no multiplicity
    factors): the receiver-resolved transfer S_R(k | G) = Re[ sum_{p+q=k} i (k . H(q)) (G(p) . R(k)^*) ] for
    helper H, giver G (a sphere / shell / outside-sphere truncation of a field) and receiver R, the form of
    EnergyTr_Compute_nlin (ET_compute_nlin.cc:85-120, N_i = D_j FFT(H_j G_i)) paired with
    Shell_mult_single / Shell_mult_all (four_ET.cc:49-80, 112-151).  Every flux and shell-to-shell entry of
    compute_flux.cc:52-262 and compute_shell_tr.cc:59-117 is then a signed sum of S_R over receiver modes:
      Compute_flux:            flux_self(i)      = - sum_{|k|>r_i} S_U(k | U<)         (helper U)
      Compute_shell_tr:        T(m, n)           = - sum_{k in shell n, |k|<=kin} S_U(k | U_m)
      Compute_flux(W):         U<->B> in_out     = + sum_{|k|>r_i}  S_B(k | U<)        (helper B)
                               in_in             = + sum_{0<|k|<=r_i} S_B(k | U<)
                               B<->B> (W.self)   = - sum_{|k|>r_i}  S_B(k | B<)        (helper U)
                               B<->U> in_out     = + sum_{|k|>r_i}  S_U(k | B<)        (helper U)
                               in_in             = + sum_{0<|k|<=r_i} S_U(k | B<)
                               U>->B> out_out    = + sum_{|k|>r_i}  S_B(k | U>)        (helper B)
                               Zp, Zm (Z+- = U +- B): - sum_{|k|>r_i} S_{Z+}(k | Z+<) with helper Z-, and
                               with the roles exchanged
      Compute_flux_scalar:     flux_SF(i)        = - sum_{|k|>r_i} S_T(k | T<)         (helper U)
      Compute_shell_tr_scalar: T_SF(m, n)        = - sum_{k in shell n, |k|<=kin} S_T(k | T_m)
      Compute_shell_tr(W):     W to W (m, n)     = - sum_{k in shell n, |k|<=kin} S_B(k | B_m)   (helper U)
                               U to W (m, n)     = + sum_{k in shell n, |k|<=kin} S_B(k | U_m)   (helper B)
                               Zp, Zm (m, n)     = - sum S_{Z+}(k | Z+_m) with helper Z-, and exchanged
    Shells follow Get_shell_index (basis_basicfn.cc:288-299): shell n holds r_{n-1} < |k| <= r_n; the giver
    truncations follow Fill_array_shell (universal_ET.cc:45-65): inner < |k| <= outer.  kin is
    Max_radius_inside_dealiased_FOUR (four_inline.h:233-245) = min(N_i / 3) = 4 at 12^3."""
    N = (12, 12, 12)
    M = 16
    rng = np.random.default_rng(14)
    kmax = (3, 2, 3)
    U = _band_field(rng, N, kmax, True)
    B = _band_field(rng, N, kmax, True)
    T = _band_field(rng, N, kmax, False)
    radii = np.array([0.0, 1, 2, 3, 4, 5, 10000.0])
    kin = 4
    FU = [_full_grid(a, N, M) for a in U]
    FB = [_full_grid(a, N, M) for a in B]
    FT = [_full_grid(T[0], N, M)]
    k1 = np.fft.fftfreq(M, 1.0 / M)
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")
    Kv = (KX, KY, KZ)
    Kmag = np.sqrt(KX * KX + KY * KY + KZ * KZ)

    def conv_nlin(H, G):
        # N_i(k) = i sum_j k_j sum_p H_j(k - p) G_i(p), exact on the M-grid (no wrap)
        out = [np.zeros((M, M, M), complex) for _ in G]
        idx = np.argwhere(np.abs(sum(np.abs(g) for g in G)) > 0)
        for (a, b, c) in idx:
            Hs = [np.roll(h, (a, b, c), axis=(0, 1, 2)) for h in H]
            kh = 1j * (KX * Hs[0] + KY * Hs[1] + KZ * Hs[2])
            for i, g in enumerate(G):
                out[i] += g[a, b, c] * kh
        return out

    def S(H, G, R):
        Nl = conv_nlin(H, G)
        return sum((n * np.conj(r)).real for n, r in zip(Nl, R))

    def trunc(F, lo, hi):
        m = (Kmag > lo) & (Kmag <= hi)
        return [f * m for f in F]

    nsph = len(radii) - 1
    out = {"N": np.array(N), "radii": radii, "kin": np.array(kin),
           "U": np.stack(U), "B": np.stack(B), "T": T[0]}
    names = ["u2u", "b2b", "u2b_in_out", "u2b_in_in", "b2u_in_out", "b2u_in_in", "u_out_b_out", "zp", "zm",
             "flux_ns", "flux_sf"]
    res = {k: np.zeros(nsph + 1) for k in names}
    FZp = [u + b for u, b in zip(FU, FB)]
    FZm = [u - b for u, b in zip(FU, FB)]
    for i in range(1, nsph + 1):
        r = radii[i]
        out_m = Kmag > r
        in_m = (Kmag > 0) & (Kmag <= r)
        s = S(FU, trunc(FU, -1, r), FU)
        res["flux_ns"][i] = res["u2u"][i] = -s[out_m].sum()
        s = S(FB, trunc(FU, -1, r), FB)
        res["u2b_in_out"][i] = s[out_m].sum()
        res["u2b_in_in"][i] = s[in_m].sum()
        s = S(FU, trunc(FB, -1, r), FB)
        res["b2b"][i] = -s[out_m].sum()
        s = S(FU, trunc(FB, -1, r), FU)
        res["b2u_in_out"][i] = s[out_m].sum()
        res["b2u_in_in"][i] = s[in_m].sum()
        s = S(FB, trunc(FU, r, 1e300), FB)
        res["u_out_b_out"][i] = s[out_m].sum()
        res["zp"][i] = -S(FZm, trunc(FZp, -1, r), FZp)[out_m].sum()
        res["zm"][i] = -S(FZp, trunc(FZm, -1, r), FZm)[out_m].sum()
        res["flux_sf"][i] = -S(FU, trunc(FT, -1, r), FT)[out_m].sum()
    out.update(res)
    nsh = nsph
    shell_of = np.searchsorted(radii, Kmag, side="left")      # first index with |k| <= r_index
    recv = Kmag <= kin
    Tns = np.zeros((nsh + 1, nsh + 1))
    Tsf = np.zeros((nsh + 1, nsh + 1))
    # Compute_shell_tr(IncVF& W) (compute_shell_tr.cc:154-260): W to W (helper U), U to W (helper W, + sign),
    # Zp to Zp (helper Zm), Zm to Zm (helper Zp)
    mhd = {k: np.zeros((nsh + 1, nsh + 1)) for k in ("ww", "uw", "zp", "zm")}
    for m in range(1, nsh):
        s = S(FU, trunc(FU, radii[m - 1], radii[m]), FU)
        st = S(FU, trunc(FT, radii[m - 1], radii[m]), FT)
        sm = {"ww": -S(FU, trunc(FB, radii[m - 1], radii[m]), FB),
              "uw": S(FB, trunc(FU, radii[m - 1], radii[m]), FB),
              "zp": -S(FZm, trunc(FZp, radii[m - 1], radii[m]), FZp),
              "zm": -S(FZp, trunc(FZm, radii[m - 1], radii[m]), FZm)}
        for n in range(nsh + 1):
            sel = recv & (shell_of == n)
            Tns[m, n] = -s[sel].sum()
            Tsf[m, n] = -st[sel].sum()
            for k in mhd:
                mhd[k][m, n] = sm[k][sel].sum()
    out["shell_tr_ns"] = Tns
    out["shell_tr_sf"] = Tsf
    for k, v in mhd.items():
        out["shell_tr_mhd_" + k] = v
    np.savez_compressed(os.path.join(HERE, "et_triad.npz"), **out)


def ring_triad():
    """Ring-to-ring and cylinder-ring-to-ring transfers (compute_ring_tr.cc:60-173, compute_cylin_ring_tr.cc:54-133)
    by explicit triad sums over the full lattice, as et_triad: T(giver ring, receiver ring) = sign * sum over the
    receiver ring of S_R(k | G) with G the giver field restricted to the giver ring.  Giver rings are selected on
    the half spectrum (Fill_array_ring / Fill_array_cylinder_ring universal_ET.cc:70-127: inner < |k| <= outer
    or inner < k_perp <= outer, and left < theta <= right with theta = Get_polar_angle(k_perp, k_pll) in
    [0, pi] unfolded, or h_lower < k_pll <= h_upper signed; theta = 0 or k_pll = 0 joins the first sector /
    slab), their conjugates following, and on the kz = 0 plane (G(k) + conj G(-k)) / 2; receivers are lattice modes with 0 < |k| <= kin (rings) or
    k_perp <= kperp_max (cylinders) in the ring of their folded angle (Compute_ring_index basis_basicfn.cc:
    329-366) or of |k_pll| (Compute_cylinder_ring_index :413-450)."""
    N = (12, 12, 12)
    M = 16
    rng = np.random.default_rng(15)
    kmax = (3, 3, 3)
    U = _band_field(rng, N, kmax, True)
    T = _band_field(rng, N, kmax, False)[0]
    kin, kpm = 4, 4
    radii = np.array([0.0, 2.0, 3.0, 10000.0])
    sect = np.array([0.0, np.pi / 4, np.pi / 2])
    slabs = np.array([0.0, 2.0, 4.0])
    N1, N2, N3 = N
    kxh = np.fft.fftfreq(N1, 1.0 / N1).reshape(-1, 1, 1) * np.ones((1, N2, N3 // 2 + 1))
    kyh = np.fft.fftfreq(N2, 1.0 / N2).reshape(1, -1, 1) * np.ones((N1, 1, N3 // 2 + 1))
    kzh = np.arange(N3 // 2 + 1).reshape(1, 1, -1) * np.ones((N1, N2, 1))
    k1 = np.fft.fftfreq(M, 1.0 / M)
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")

    def polar(kperp, kpll):
        t = np.where(np.abs(kpll) > 1e-15, np.arctan(kperp / np.where(kpll == 0, 1, kpll)), np.pi / 2)
        t = np.where(np.abs(kperp) > 1e-15, t, np.where(kpll >= 0, 0.0, np.pi))
        return np.where(t >= 0, t, t + np.pi)

    def pll_perp(kx, ky, kz, a):
        c = (kx, ky, kz)
        pll = c[a - 1]
        perp = np.sqrt(sum(v * v for i, v in enumerate(c) if i != a - 1))
        return pll, perp

    def conv_nlin(H, G):
        out = [np.zeros((M, M, M), complex) for _ in G]
        idx = np.argwhere(np.abs(sum(np.abs(g) for g in G)) > 0)
        for (a, b, c) in idx:
            Hs = [np.roll(h, (a, b, c), axis=(0, 1, 2)) for h in H]
            kh = 1j * (KX * Hs[0] + KY * Hs[1] + KZ * Hs[2])
            for i, g in enumerate(G):
                out[i] += g[a, b, c] * kh
        return out

    FU = [_full_grid(a, N, M) for a in U]
    FT = [_full_grid(T, N, M)]
    out = {"N": np.array(N), "U": np.stack(U), "T": T, "radii": radii, "sectors": sect, "slabs": slabs}
    for cyl in (False, True):
        for a in ((3, 1) if not cyl else (3, 2)):
            hp, hq = pll_perp(kxh, kyh, kzh, a)
            lp, lq = pll_perp(KX, KY, KZ, a)
            Kh = np.sqrt(kxh ** 2 + kyh ** 2 + kzh ** 2)
            Kl = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
            edges = slabs if cyl else sect
            if cyl:
                rsel = lq <= kpm
                rsh = np.searchsorted(radii, lq, side="left")
                al = np.abs(lp)
                rse = np.where(al < 1e-15, 1, np.searchsorted(edges, al, side="left"))
                rse = np.maximum(rse, 1)
            else:
                rsel = (Kl <= kin) & (Kl > 1e-15)
                rsh = np.searchsorted(radii, Kl, side="left")
                th = polar(lq, lp)
                th = np.where(th > np.pi / 2, np.pi - th, th)
                rse = np.where(np.abs(th) < 1e-15, 1, np.searchsorted(edges, th, side="left"))
            for name, Fsrc, Ffull, Rfull in (("ns", U, FU, FU), ("sf", [T], FT, FT)):
                res = np.zeros((len(radii), len(edges), len(radii), len(edges)))
                for gs in range(1, len(radii) - 1):
                    for gq in range(1, len(edges)):
                        if cyl:
                            m = (hq > radii[gs - 1]) & (hq <= radii[gs])
                            m &= ((hp > edges[gq - 1]) & (hp <= edges[gq])) | ((np.abs(hp) < 1e-15) & (edges[gq - 1] < 1e-15))
                        else:
                            m = (Kh > radii[gs - 1]) & (Kh <= radii[gs])
                            t = polar(hq, hp)
                            m &= ((t > edges[gq - 1]) & (t <= edges[gq])) | ((np.abs(t) < 1e-15) & (edges[gq - 1] < 1e-15))
                        # A giver that breaks the kz = 0 plane's Hermitian pairs (axis 1 or 2: k_pll < 0 lies in
                        # no sector) enters the inverse transform as FFTW's c2r takes it: the z line's DC
                        # coefficient keeps its real part, i.e. the plane acts as (G(k) + conj G(-k)) / 2.
                        G = []
                        for f in Fsrc:
                            h = f * m
                            p0 = h[:, :, 0]
                            mirror = np.conj(p0[(-np.arange(N1)) % N1][:, (-np.arange(N2)) % N2])
                            h[:, :, 0] = 0.5 * (p0 + mirror)
                            G.append(_full_grid(h, N, M))
                        Nl = conv_nlin(FU, G)
                        Sr = sum((n * np.conj(r)).real for n, r in zip(Nl, Rfull))
                        for rs in range(len(radii)):
                            for rq in range(len(edges)):
                                sel = rsel & (rsh == rs) & (rse == rq)
                                res[gs, gq, rs, rq] = -Sr[sel].sum()
                out[f"{'cyl' if cyl else 'ring'}_{name}_axis{a}"] = res
    np.savez_compressed(os.path.join(HERE, "ring_triad.npz"), **out)


if __name__ == "__main__":
    four_fft()
    scft_fft()
    kat_out1()
    nlin_triad()
    et_triad()
    ring_triad()
    print("golden fixtures written to", HERE)
