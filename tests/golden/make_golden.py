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
Usage: python tests/golden/make_golden.py
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


if __name__ == "__main__":
    four_fft()
    scft_fft()
    kat_out1()
    nlin_triad()
    print("golden fixtures written to", HERE)
