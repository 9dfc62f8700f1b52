// th_reduce.hip -- FP64 shell binning and sums over the reference layout.
//   Compute_local_shell_spectrum_FOUR  trunk/lib/four/four_energy.cc:244-288 (SCFT scft_energy.cc:238-290)
//   Get_local_energy_FOUR              four_energy.cc:48-168 (SCFT scft_energy.cc:51-101)
//   Get_local_Sn_FOUR                  four_energy.cc:179-218
//   Local_shell_mult_single_FOUR       trunk/lib/four/four_ET.cc:49-80
//   Local_shell_mult_all_FOUR (dot)    four_ET.cc:181-229, Get_shell_index basis_basicfn.cc:288-299
// K = sqrt((kx f1)^2 + (ky f2)^2 + (kz f3)^2) is evaluated in FP64 with contraction disabled, so the
// shell index ceil(K) is bit-identical to the CPU's.  Per-block partial sums are FP64 (LDS atomics
// for binned sums, tree reduction for scalars); blocks are merged with Neumaier-compensated sums.
#include <math.h>
#include <mutex>
#include "th_internal.hpp"

namespace th {

struct GeoR {
    int basis, N1, N2, N3, H;
    double f1, f2, f3;
};
static GeoR gr(const Geo& g) { return GeoR{g.basis, g.N1, g.N2, g.N3, g.H, g.kf[0], g.kf[1], g.kf[2]}; }

__device__ __forceinline__ double kmag_r(const GeoR& g, int l1, int l2, int l3)
{
#pragma clang fp contract(off)
    int kx = g.basis == 1 ? l1 : kx_four(l1, g.N1);
    double a = kx * g.f1, b = ky_of(l2, g.N2) * g.f2, c = l3 * g.f3;
    return sqrt(a * a + b * b + c * c);
}

__device__ __forceinline__ double mult_r(const GeoR& g, int l1, int l2, int l3)
{
    int kx = g.basis == 1 ? l1 : kx_four(l1, g.N1);
    int ky = ky_of(l2, g.N2);
    double factor;
    if (g.basis == 0) {
        factor = (l3 > 0) ? 1.0 : 0.5;
        if ((kx == g.N1 / 2) || ((ky == g.N2 / 2) && (g.N2 > 1))) return 2 * factor;
        return factor;
    }
    if (l3 > 0) factor = (kx > 0) ? 2.0 : 1.0;
    else factor = (kx > 0) ? 1.0 : 0.5;
    if ((ky == g.N2 / 2) && (g.N2 > 1)) return 2 * factor;
    return factor;
}

__device__ __forceinline__ double my_pow_r(double x, int n)
{
    if (n == 0) return 1.0;
    if (n == 2) return x * x;
    return x * x * x * x;
}

template <typename R> __device__ __forceinline__ double sq(const cplx<R>& a) { return (double)a.x * a.x + (double)a.y * a.y; }
template <typename R> __device__ __forceinline__ double rdot(const cplx<R>& a, const cplx<R>& b)
{
    return (double)a.x * b.x + (double)a.y * b.y;
}

__device__ __forceinline__ void decode_r(const GeoR& g, long i, int& l1, int& l2, int& l3)
{
    l3 = (int)(i % g.H);
    long r = i / g.H;
    l2 = (int)(r % g.N2);
    l1 = (int)(r / g.N2);
}

constexpr int RBLK = 256;
constexpr int RSEG = 16;      // consecutive modes handled by one thread (run-length accumulation)

// MODE 0: spectrum (bins = ceil(K) <= kkmax; Sk and count), MODE 1: shell_mult_all (bins by radii,
// K <= kin, value 2*mult*sum_c Re(A_c conj(B_c)))
template <typename R, int MODE>
__global__ __launch_bounds__(RBLK) void k_binned(GeoR g, const cplx<R>* __restrict__ A0, const cplx<R>* __restrict__ A1,
                                                 const cplx<R>* __restrict__ A2, const cplx<R>* __restrict__ B0,
                                                 const cplx<R>* __restrict__ B1, const cplx<R>* __restrict__ B2,
                                                 int npow, int kkmax, const double* __restrict__ radii, int nr,
                                                 int nbins, double* __restrict__ partial)
{
    extern __shared__ double sh[];
    double* sS = sh;
    double* sC = sh + nbins;
    for (int b = threadIdx.x; b < 2 * nbins; b += RBLK) sh[b] = 0.0;
    __syncthreads();
    const long n = (long)g.N1 * g.N2 * g.H;
    const long nseg = (n + RSEG - 1) / RSEG;
    for (long s = (long)blockIdx.x * RBLK + threadIdx.x; s < nseg; s += (long)gridDim.x * RBLK) {
        int cur = -1;
        double accS = 0.0, accC = 0.0;
        const long i0 = s * RSEG, i1 = (i0 + RSEG < n) ? i0 + RSEG : n;
        for (long i = i0; i < i1; i++) {
            int l1, l2, l3;
            decode_r(g, i, l1, l2, l3);
            double km = kmag_r(g, l1, l2, l3);
            int idx;
            double val, cnt = 0.0;
            if (MODE == 0) {
                idx = (int)ceil(km);
                if (idx > kkmax) continue;
                double f = mult_r(g, l1, l2, l3);
                double v = B0 ? rdot(A0[i], B0[i]) : sq(A0[i]);
                val = f * my_pow_r(km, npow) * v;
                cnt = (g.basis == 0) ? f : f / 2;
            } else {
                if (km > (double)kkmax) continue;
                idx = 0;
                while (idx < nr && !(km <= radii[idx])) idx++;
                if (idx >= nbins) continue;
                double f = mult_r(g, l1, l2, l3);
                double d = rdot(A0[i], B0[i]);
                if (A1) d += rdot(A1[i], B1[i]);          // null: one-component (scalar) pair
                if (A2) d += rdot(A2[i], B2[i]);
                val = 2 * f * d;
            }
            if (idx != cur) {
                if (cur >= 0) {
                    atomicAdd(&sS[cur], accS);
                    if (MODE == 0) atomicAdd(&sC[cur], accC);
                }
                cur = idx;
                accS = 0.0;
                accC = 0.0;
            }
            accS += val;
            accC += cnt;
        }
        if (cur >= 0) {
            atomicAdd(&sS[cur], accS);
            if (MODE == 0) atomicAdd(&sC[cur], accC);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 2 * nbins; b += RBLK) partial[(long)blockIdx.x * 2 * nbins + b] = sh[b];
}

// MODE 0: energy weights (Get_local_energy), MODE 1: Sn, MODE 2: shell_mult_single over 3 comps (or 1)
template <typename R, int MODE>
__global__ __launch_bounds__(RBLK) void k_scalar(GeoR g, const cplx<R>* __restrict__ A0, const cplx<R>* __restrict__ A1,
                                                 const cplx<R>* __restrict__ A2, const cplx<R>* __restrict__ B0,
                                                 const cplx<R>* __restrict__ B1, const cplx<R>* __restrict__ B2,
                                                 int npow, int kkmax, double inner, double outer,
                                                 double* __restrict__ partial)
{
    __shared__ double red[RBLK];
    double s = 0.0, c = 0.0;                 // Neumaier
    const long n = (long)g.N1 * g.N2 * g.H;
    for (long i = (long)blockIdx.x * RBLK + threadIdx.x; i < n; i += (long)gridDim.x * RBLK) {
        int l1, l2, l3;
        decode_r(g, i, l1, l2, l3);
        double v;
        if (MODE == 0) {
            double w;
            if (g.basis == 0) {
                double ey = (l2 == g.N2 / 2 && g.N2 > 1) ? 1.0 : 0.0;
                double ex = (l1 == g.N1 / 2) ? 1.0 : 0.0;
                w = 1.0 + ey + ex;
                if (l3 == 0) w -= 0.5 * (1.0 + ey + ex);
            } else {
                w = 2.0 - (l3 == 0 ? 1.0 : 0.0) - (l1 == 0 ? 1.0 : 0.0) + ((l1 == 0 && l3 == 0) ? 0.5 : 0.0);
                if (l2 == g.N2 / 2 && g.N2 > 2) w *= 2.0;
            }
            if (i == 0) w -= 0.5;
            v = w * (B0 ? rdot(A0[i], B0[i]) : sq(A0[i]));
        } else if (MODE == 1) {
            double km = kmag_r(g, l1, l2, l3);
            if (!(km <= (double)kkmax)) continue;
            v = mult_r(g, l1, l2, l3) * my_pow_r(km, npow) * sq(A0[i]);
            if (i == 0 && npow == 0) v -= mult_r(g, 0, 0, 0) * sq(A0[0]);
        } else {
            double km = kmag_r(g, l1, l2, l3);
            if (!((km > inner) && (km <= outer))) continue;
            double d = rdot(A0[i], B0[i]);
            if (A1) d += rdot(A1[i], B1[i]);              // null: one-component (scalar) pair
            if (A2) d += rdot(A2[i], B2[i]);
            v = 2 * mult_r(g, l1, l2, l3) * d;
        }
        double t = s + v;
        c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    red[threadIdx.x] = s + c;
    __syncthreads();
    for (int w = RBLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[b] = compensated sum over blocks of partial[blk * stride + b]
// Helicity and spectral entropy (FOUR basis; the glob.d / spectrum.d columns).
//   MODE 0: Compute_local_helicity_FOUR four_energy.cc:453-506 over |K| <= kkmax:
//           modal = 2 mult K.(Vr x Vi); values {sum modal, sum modal/K^2 (K^2 > 1e-15), sum K^2 modal}
//   MODE 1: Get_local_entropy_FOUR four_energy.cc:687-732: p = (sum_c |V_c|^2 / 2) / E,
//           value sum 2 mult (-p log2 p) over p > 1e-15
// SCFT boxes hold V1 in the sin basis along x; -i V1 is its Fourier counterpart (scft_energy.cc:470-474)
template <typename R> __device__ __forceinline__ cplx<R> rot_sin(const GeoR& g, const cplx<R>& a)
{
    return g.basis == 1 ? mk<R>(a.y, -a.x) : a;
}
__device__ __forceinline__ int kx_r(const GeoR& g, int l1) { return g.basis == 1 ? l1 : kx_four(l1, g.N1); }
template <typename R> __device__ __forceinline__ void vr_x_vi(const cplx<R>& a, const cplx<R>& b, const cplx<R>& c,
                                                             double* x)
{
    x[0] = (double)b.x * c.y - (double)c.x * b.y;
    x[1] = (double)c.x * a.y - (double)a.x * c.y;
    x[2] = (double)a.x * b.y - (double)b.x * a.y;
}
template <typename R, int MODE>
__global__ __launch_bounds__(RBLK) void k_modal(GeoR g, const cplx<R>* __restrict__ A0, const cplx<R>* __restrict__ A1,
                                                const cplx<R>* __restrict__ A2, int kkmax, double E,
                                                double* __restrict__ partial)
{
    constexpr int NV = MODE == 0 ? 3 : 1;
    __shared__ double red[NV][RBLK];
    double acc[NV] = {};
    const long n = (long)g.N1 * g.N2 * g.H;
    for (long i = (long)blockIdx.x * RBLK + threadIdx.x; i < n; i += (long)gridDim.x * RBLK) {
        int l1, l2, l3;
        decode_r(g, i, l1, l2, l3);
        if (MODE == 0) {
            double km = kmag_r(g, l1, l2, l3);
            if (km > (double)kkmax) continue;
            double x[3];
            vr_x_vi(rot_sin(g, A0[i]), A1[i], A2[i], x);
            const double kx = kx_r(g, l1) * g.f1, ky = ky_of(l2, g.N2) * g.f2, kz = l3 * g.f3;
            const double ksq = km * km;
            const double modal = 2 * mult_r(g, l1, l2, l3) * (kx * x[0] + ky * x[1] + kz * x[2]);
            acc[0] += modal;
            if (ksq > 1e-15) acc[NV > 1 ? 1 : 0] += modal / ksq;
            acc[NV > 2 ? 2 : 0] += ksq * modal;
        } else {
            // Modal_energy_SCFT (scft_inline.h:347-354) is |A|^2 for kx > 0; SCFT weighs by mult, FOUR by 2 mult
            const double v2 = sq(A0[i]) + (A1 ? sq(A1[i]) + sq(A2[i]) : 0.0);     // A1 == nullptr: one field
            const double me = 0.5 * v2 * ((g.basis == 1 && l1 > 0) ? 2.0 : 1.0);
            const double p = me / E;
            if (p > 1e-15) acc[0] += (g.basis == 1 ? 1.0 : 2.0) * mult_r(g, l1, l2, l3) * (-p * log(p) / log(2.0));
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int w = RBLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int v = 0; v < NV; v++) red[v][threadIdx.x] += red[v][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < NV) partial[(long)blockIdx.x * NV + threadIdx.x] = red[threadIdx.x][0];
}

// Compute_local_shell_spectrum_helicity_FOUR four_energy.cc:551-599 (_SCFT scft_energy.cc:546-599): bin ceil(K) <= kkmax,
// H1k_c += 2 mult K_c (Vr x Vi)_c and count += 2 mult; partial[block][4 nbins]
template <typename R>
__global__ __launch_bounds__(RBLK) void k_binned_hel(GeoR g, const cplx<R>* __restrict__ A0, const cplx<R>* __restrict__ A1,
                                                     const cplx<R>* __restrict__ A2, int kkmax, int nbins,
                                                     double* __restrict__ partial)
{
    extern __shared__ double sh[];
    for (int b = threadIdx.x; b < 4 * nbins; b += RBLK) sh[b] = 0.0;
    __syncthreads();
    const long n = (long)g.N1 * g.N2 * g.H;
    for (long i = (long)blockIdx.x * RBLK + threadIdx.x; i < n; i += (long)gridDim.x * RBLK) {
        int l1, l2, l3;
        decode_r(g, i, l1, l2, l3);
        const int idx = (int)ceil(kmag_r(g, l1, l2, l3));
        if (idx > kkmax || idx >= nbins) continue;
        double x[3];
        vr_x_vi(rot_sin(g, A0[i]), A1[i], A2[i], x);
        const double f = 2 * mult_r(g, l1, l2, l3);
        atomicAdd(&sh[idx], f * (kx_r(g, l1) * g.f1 * x[0]));
        atomicAdd(&sh[nbins + idx], f * (ky_of(l2, g.N2) * g.f2 * x[1]));
        atomicAdd(&sh[2 * nbins + idx], f * (l3 * g.f3 * x[2]));
        atomicAdd(&sh[3 * nbins + idx], g.basis == 1 ? f / 2 : f);      // SCFT counts factor/2
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 4 * nbins; b += RBLK) partial[(long)blockIdx.x * 4 * nbins + b] = sh[b];
}

// Anisotropic spectra (FOUR, 3-D) about axis a = globalvar_anisotropy_switch (1, 2, 3), bins
// (row, col) flattened row-major into rows x cols; a column past the last lands on the next row's
// column 0 and past the array it is dropped (the reference's unchecked Blitz store).
//   MODE 0, ring: Compute_local_ring_spectrum_FOUR four_energy.cc:828-954; row ceil|K| <= kmaxi for
//     |K| > MYEPS, column Get_sector_index basis_basicfn.cc:304-322 of the polar angle
//     (Get_polar_angle basis_basicfn_inline.h:89-112; sector edges i dpar, the last pmax).
//     Off the axis S1 += w |(V x K)_a / K_perp|^2, S2 += w (|V|^2 - that); on the axis (a = 1) S1 += w |V_y|^2,
//     S2 += w |V_z|^2, (a = 2) S1 += w |V_z|^2, S2 += w |V_y|^2, (a = 3) nothing (the components are never set).
//   MODE 1, cylinder: Compute_local_cylinder_ring_spectrum_FOUR four_energy.cc:1450-1515; row ceil(K_perp)
//     <= kmaxi, column Get_slab_index basis_basicfn.cc:373-390 of |K_pll| (slab edges i dpar, the last pmax);
//     S1 += w |V_a|^2 (0 for a = 3), S2 += w (|V|^2 - |V_a|^2).
//   w = Multiplicity_factor_FOUR |K|^n.
struct AnisoP {
    int axis, nbin, rows, cols, n, kmaxi;
    double dpar, pmax;
};

//   HEL: the helicity tables instead (Compute_local_ring_spectrum_helicity_FOUR four_energy.cc:1357-1412,
//   Compute_local_cylinder_ring_spectrum_helicity_FOUR :1825-1880): S1 += 2 mult K.(Vr x Vi), S2 untouched.
template <typename R, int MODE, bool HEL>
__global__ __launch_bounds__(RBLK) void k_aniso(GeoR g, const cplx<R>* __restrict__ A0, const cplx<R>* __restrict__ A1,
                                                const cplx<R>* __restrict__ A2, AnisoP ap, double* __restrict__ partial)
{
    extern __shared__ double sh[];
    const int nb = ap.rows * ap.cols;
    for (int b = threadIdx.x; b < 2 * nb; b += RBLK) sh[b] = 0.0;
    __syncthreads();
    const long n = (long)g.N1 * g.N2 * g.H;
    const double eps = 1e-15;
    for (long i = (long)blockIdx.x * RBLK + threadIdx.x; i < n; i += (long)gridDim.x * RBLK) {
        int l1, l2, l3;
        decode_r(g, i, l1, l2, l3);
        const double kk[3] = {kx_four(l1, g.N1) * g.f1, ky_of(l2, g.N2) * g.f2, l3 * g.f3};
        const double km = kmag_r(g, l1, l2, l3);
        const int a = ap.axis - 1, p = ap.axis % 3, q = (ap.axis + 1) % 3;
        const double kpll = kk[a];
        double kperp;
        {
#pragma clang fp contract(off)
            kperp = sqrt(kk[p] * kk[p] + kk[q] * kk[q]);
        }
        int row, col;
        if (MODE == 0) {
            row = (int)ceil(km);
            if (!(km > eps && row <= ap.kmaxi)) continue;
            double t;
            if (fabs(kperp) > eps) {
                if (fabs(kpll) > eps) {
                    const double r = kperp / kpll;
                    t = (r == 1.0) ? M_PI / 4 : atan(r);       // atan(1) correctly rounded, as libm gives it
                } else {
                    t = M_PI / 2;
                }
            } else {
                t = (kpll >= 0) ? 0.0 : M_PI;
            }
            if (t < 0) t += M_PI;
            if (t > M_PI / 2) t = M_PI - t;
            if (fabs(t) < eps) {
                col = 1;
            } else {
                for (col = 0; col < ap.cols; col++)
                    if (t <= (col < ap.nbin ? col * ap.dpar : ap.pmax)) break;
            }
        } else {
            row = (int)ceil(kperp);
            if (row > ap.kmaxi) continue;
            const double ak = fabs(kpll);
            if (ak < eps) {
                col = 1;
            } else {
                for (col = 1; col < ap.cols; col++)
                    if (ak <= (col < ap.nbin ? col * ap.dpar : ap.pmax)) break;
            }
        }
        const long flat = (long)row * ap.cols + col;
        if (flat >= nb) continue;
        if (HEL) {
            double x[3];
            vr_x_vi(A0[i], A1[i], A2[i], x);
            atomicAdd(&sh[flat], 2 * mult_r(g, l1, l2, l3) * (kk[0] * x[0] + kk[1] * x[1] + kk[2] * x[2]));
            continue;
        }
        const cplx<R> v[3] = {A0[i], A1[i], A2[i]};
        const double w = mult_r(g, l1, l2, l3) * my_pow_r(km, ap.n);
        const double v2 = sq(v[0]) + sq(v[1]) + sq(v[2]);
        double s1, s2;
        if (MODE == 0) {
            if (kperp < eps) {
                if (ap.axis == 3) continue;
                s1 = sq(v[ap.axis == 1 ? 1 : 2]);
                s2 = sq(v[ap.axis == 1 ? 2 : 1]);
            } else {
                const double xr = ((double)v[p].x * kk[q] - (double)v[q].x * kk[p]) / kperp;
                const double xi = ((double)v[p].y * kk[q] - (double)v[q].y * kk[p]) / kperp;
                s1 = xr * xr + xi * xi;
                s2 = v2 - s1;
            }
        } else {
            s1 = ap.axis == 3 ? 0.0 : sq(v[a]);
            s2 = v2 - s1;
        }
        atomicAdd(&sh[flat], w * s1);
        atomicAdd(&sh[nb + flat], w * s2);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 2 * nb; b += RBLK) partial[(long)blockIdx.x * 2 * nb + b] = sh[b];
}


// ---- ring-to-ring and cylinder-ring-to-ring transfers (FOUR, 3-D) ----
// Polar angle of K about the axis (Get_polar_angle basis_basicfn_inline.h:89-112), in [0, pi]; a ratio of
// exactly +-1 takes +-pi/4, what a correctly rounded atan returns.
__device__ __forceinline__ double polar_angle_r(double kperp, double kpll)
{
    const double eps = 1e-15;
    double t;
    if (fabs(kperp) > eps) {
        if (fabs(kpll) > eps) {
            const double r = kperp / kpll;
            t = (r == 1.0) ? M_PI / 4 : (r == -1.0) ? -(M_PI / 4) : atan(r);
        } else {
            t = M_PI / 2;
        }
    } else {
        t = (kpll >= 0) ? 0.0 : M_PI;
    }
    return (t >= 0) ? t : t + M_PI;
}
__device__ __forceinline__ void pll_perp_r(const GeoR& g, int axis, int l1, int l2, int l3, double& kpll, double& kperp)
{
#pragma clang fp contract(off)
    const double kk[3] = {kx_four(l1, g.N1) * g.f1, ky_of(l2, g.N2) * g.f2, l3 * g.f3};
    const int p = axis % 3, q = (axis + 1) % 3;
    kpll = kk[axis - 1];
    kperp = sqrt(kk[p] * kk[p] + kk[q] * kk[q]);
}

// Fill_array_ring / Fill_array_cylinder_ring universal_ET.cc:70-127: inner < |K| (K_perp) <= outer and
// left < theta <= right with theta unfolded in [0, pi] (or h_lower < K_pll <= h_upper, K_pll signed); the
// theta = 0 (K_pll = 0) line belongs to the first sector (slab).
template <typename R>
__global__ void k_fill_ring(GeoR g, const cplx<R>* __restrict__ F, cplx<R>* __restrict__ G, RingP rp, double lo,
                            double hi, double e0, double e1)
{
    const long n = (long)g.N1 * g.N2 * g.H;
    const double eps = 1e-15;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int l1, l2, l3;
        decode_r(g, i, l1, l2, l3);
        double kpll, kperp;
        pll_perp_r(g, rp.axis, l1, l2, l3, kpll, kperp);
        bool in;
        if (rp.cyl) {
            in = (kperp > lo) && (kperp <= hi) && (((kpll > e0) && (kpll <= e1)) || ((fabs(kpll) < eps) && (e0 < eps)));
        } else {
            const double km = kmag_r(g, l1, l2, l3);
            const double th = polar_angle_r(kperp, kpll);
            in = (km > lo) && (km <= hi) && (((th > e0) && (th <= e1)) || ((fabs(th) < eps) && (e0 < eps)));
        }
        G[i] = in ? F[i] : mk<R>(R(0), R(0));
    }
}

// Local_ring_mult_all_FOUR four_ET.cc:259-299 (receivers 0 < |K| <= Max_radius_inside, Compute_ring_index
// basis_basicfn.cc:329-366: the first radius >= |K|, the folded angle's sector, theta = 0 in sector 1) and
// Local_cyl_ring_mult_all_FOUR :422-460 (K_perp <= Anis_max_Krho_radius_inside, Compute_cylinder_ring_index
// :413-450: the first K_perp radius, the slab of |K_pll|, K_pll = 0 in slab 1); value 2 mult Re(A . B*).
// A bin past the last column lands on the next row's column 0 and past the table it is dropped, as the
// reference's unchecked Blitz store.
template <typename R, bool CYL>
__global__ __launch_bounds__(RBLK) void k_ring_mult(GeoR g, const cplx<R>* __restrict__ A0, const cplx<R>* __restrict__ A1,
                                                    const cplx<R>* __restrict__ A2, const cplx<R>* __restrict__ B0,
                                                    const cplx<R>* __restrict__ B1, const cplx<R>* __restrict__ B2,
                                                    RingP rp, double* __restrict__ partial)
{
    extern __shared__ double sh[];
    const int rows = rp.nr + 1, cols = rp.nsec + 1, nb = rows * cols;
    for (int b = threadIdx.x; b < nb; b += RBLK) sh[b] = 0.0;
    __syncthreads();
    const long n = (long)g.N1 * g.N2 * g.H;
    const double eps = 1e-15;
    for (long i = (long)blockIdx.x * RBLK + threadIdx.x; i < n; i += (long)gridDim.x * RBLK) {
        int l1, l2, l3;
        decode_r(g, i, l1, l2, l3);
        double kpll, kperp;
        pll_perp_r(g, rp.axis, l1, l2, l3, kpll, kperp);
        int row, col;
        if (CYL) {
            if (!(kperp <= rp.kpm)) continue;
            for (row = 0; row < rows; row++)
                if (kperp <= rp.r[row]) break;
            const double ak = fabs(kpll);
            if (ak < eps) {
                col = 1;
            } else {
                for (col = 1; col < cols; col++)
                    if (ak <= rp.e[col]) break;
            }
        } else {
            const double km = kmag_r(g, l1, l2, l3);
            if (!((km <= rp.kin) && (km > eps))) continue;
            for (row = 0; row < rows; row++)
                if (km <= rp.r[row]) break;
            double th = polar_angle_r(kperp, kpll);
            if (th > M_PI / 2) th = M_PI - th;
            if (fabs(th) < eps) {
                col = 1;
            } else {
                for (col = 0; col < cols; col++)
                    if (th <= rp.e[col]) break;
            }
        }
        const long flat = (long)row * cols + col;
        if (flat >= nb) continue;
        double d = rdot(A0[i], B0[i]);
        if (A1) d += rdot(A1[i], B1[i]) + rdot(A2[i], B2[i]);
        atomicAdd(&sh[flat], 2 * mult_r(g, l1, l2, l3) * d);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += RBLK) partial[(long)blockIdx.x * nb + b] = sh[b];
}

int RedScratch::ensure(size_t n_dev, size_t n_host)
{
    if (n_dev > cap) {
        if (d) (void)hipFree(d);
        d = nullptr;
        TH_CHECK_HIP(hipMalloc(&d, n_dev * sizeof(double)));
        cap = n_dev;
    }
    if (n_host > hcap) {
        if (h) (void)hipHostFree(h);
        h = nullptr;
        TH_CHECK_HIP(hipHostMalloc(&h, n_host * sizeof(double)));
        hcap = n_host;
    }
    return OK;
}
void RedScratch::release()
{
    if (d) (void)hipFree(d);
    if (h) (void)hipHostFree(h);
    d = nullptr;
    h = nullptr;
    cap = hcap = 0;
}

static int nblocks_for(size_t n)
{
    size_t b = (n + RBLK * RSEG - 1) / (RBLK * RSEG);
    if (b > 1024) b = 1024;
    return (int)(b ? b : 1);
}

int ring_params(const Geo& g, int cyl, int axis, const double* radii, int nr, const double* edges, int nsec, RingP* rp)
{
    if (g.basis != 0 || g.N2 < 2 || g.N3 < 4) { set_error("ring transfers: 3-D FOUR boxes only"); return E_UNSUPPORTED; }
    if (axis < 1 || axis > 3 || nr < 2 || nr > RING_MAX || nsec < 1 || nsec > RING_MAX || !radii || !edges) {
        set_error("ring transfers: axis 1..3, 2 <= shells <= %d, 1 <= sectors <= %d", RING_MAX, RING_MAX);
        return E_ARG;
    }
    rp->cyl = cyl ? 1 : 0;
    rp->axis = axis;
    rp->nr = nr;
    rp->nsec = nsec;
    rp->kin = max_radius_inside(g);
    rp->kpm = cylinder_shell_ek_size(g, axis) - 1;     // Anis_max_Krho_radius_inside four_inline.h:517-546
    for (int i = 0; i <= RING_MAX; i++) {
        rp->r[i] = i <= nr ? radii[i] : 0.0;
        rp->e[i] = i <= nsec ? edges[i] : 0.0;
    }
    return OK;
}

template <typename R>
int fill_ring(const Geo& g, const cplx<R>* F, cplx<R>* G, const RingP& rp, int gs, int gq, hipStream_t st)
{
    const int nblk = nblocks_for(g.ncplx() / 4 + 1);
    hipLaunchKernelGGL((k_fill_ring<R>), dim3(nblk), dim3(RBLK), 0, st, gr(g), F, G, rp, rp.r[gs - 1], rp.r[gs],
                       rp.e[gq - 1], rp.e[gq]);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
int reduce_ring_mult(const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3], const RingP& rp, double* out,
                     RedScratch& rs, hipStream_t st)
{
    const int nb = (rp.nr + 1) * (rp.nsec + 1);
    const int nblk = nblocks_for(g.ncplx() / 4 + 1);
    int r = rs.ensure((size_t)nblk * nb + nb, nb);
    if (r) return r;
    double* dout = rs.d + (size_t)nblk * nb;
    const size_t lds = nb * sizeof(double);
    if (rp.cyl)
        hipLaunchKernelGGL((k_ring_mult<R, true>), dim3(nblk), dim3(RBLK), lds, st, gr(g), A[0], A[1], A[2], B[0], B[1],
                           B[2], rp, rs.d);
    else
        hipLaunchKernelGGL((k_ring_mult<R, false>), dim3(nblk), dim3(RBLK), lds, st, gr(g), A[0], A[1], A[2], B[0], B[1],
                           B[2], rp, rs.d);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nblk, nb, nb, dout, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, dout, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < nb; k++) out[k] = rs.h[k];
    return OK;
}

int min_radius_outside(const Geo& g)
{
    auto p2 = [](double x) { return x * x; };
    int N1 = g.N1, N2 = g.N2, N3 = g.N3;
    if (g.basis == 0) {
        if (g.dealias) return (int)ceil(sqrt(p2((N1 / 3) * g.kf[0]) + p2(N2 / 3 * g.kf[1]) + p2(N3 / 3 * g.kf[2])));
        return (int)ceil(sqrt(p2((N1 / 2) * g.kf[0]) + p2(N2 / 2 * g.kf[1]) + p2(N3 / 2 * g.kf[2])));
    }
    if (g.dealias) return (int)ceil(sqrt(p2(2 * N1 / 3 * g.kf[0]) + p2(N2 / 3 * g.kf[1]) + p2(N3 / 3 * g.kf[2])));
    return (int)ceil(sqrt(p2((N1 - 1) * g.kf[0]) + p2(N2 / 2 * g.kf[1]) + p2(N3 / 2 * g.kf[2])));
}

int max_radius_inside(const Geo& g)
{
    auto mn = [](double a, double b) { return a < b ? a : b; };
    int N1 = g.N1, N2 = g.N2, N3 = g.N3;
    double k;
    if (g.basis == 0) {
        if (g.dealias) { k = mn((N1 / 3) * g.kf[0], (N2 / 3) * g.kf[1]); if (N3 > 2) k = mn(k, (N3 / 3) * g.kf[2]); }
        else { k = mn((N1 / 2) * g.kf[0], (N2 / 2) * g.kf[1]); if (N3 > 2) k = mn(k, (N3 / 2) * g.kf[2]); }
    } else {
        if (g.dealias) { k = mn(2 * N1 / 3 * g.kf[0], (N2 / 3) * g.kf[1]); if (N3 > 2) k = mn(k, (N3 / 3) * g.kf[2]); }
        else { k = mn((N1 - 1) * g.kf[0], (N2 / 2) * g.kf[1]); if (N3 > 2) k = mn(k, (N3 / 2) * g.kf[2]); }
    }
    return (int)k;
}

double approx_modes_in_shell(const Geo& g, int r)
{
    double v = M_PI * r * r / (g.kf[0] * g.kf[1] * g.kf[2]);
    return g.basis == 0 ? 2 * v : v;
}

template <typename R>
int reduce_shell(const Geo& g, const cplx<R>* A, const cplx<R>* B, double n, int nbins, double* Sk, double* cnt,
                 RedScratch& rs, hipStream_t st, bool rescale)
{
    int nb = nblocks_for(g.ncplx());
    int kkmax = min_radius_outside(g);
    if (nbins != kkmax + 1) { set_error("reduce_shell: nbins %d != shell_ek_size %d", nbins, kkmax + 1); return E_ARG; }
    int r = rs.ensure((size_t)nb * 2 * nbins + 2 * nbins, 2 * nbins);
    if (r) return r;
    double* out = rs.d + (size_t)nb * 2 * nbins;
    hipLaunchKernelGGL((k_binned<R, 0>), dim3(nb), dim3(RBLK), 2 * nbins * sizeof(double), st, gr(g), A,
                       (const cplx<R>*)nullptr, (const cplx<R>*)nullptr, B, (const cplx<R>*)nullptr,
                       (const cplx<R>*)nullptr, (int)n, kkmax, (const double*)nullptr, 0, nbins, rs.d);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nb, 2 * nbins, 2 * nbins, out, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, out, 2 * nbins * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    int kin = max_radius_inside(g);
    for (int k = 0; k < nbins; k++) {
        double s = rs.h[k], c = rs.h[nbins + k];
        if (rescale && k > kin && k <= kkmax && c >= 1) s = s * approx_modes_in_shell(g, k) / c;   // four_energy.cc:324-336
        Sk[k] = s;
        if (cnt) cnt[k] = c;
    }
    return OK;
}

// ---- th_shell_spectrum_noalloc: the shell spectrum with no device or host allocation ----
// Its per-block partials (nblocks_for(n) <= 1024 blocks x 2 NOALLOC_BINS values) and finalized sums and
// counts (2 NOALLOC_BINS values) live in this file-scope array, one copy per device for the life of the
// process (8.4 MB).  Host code never passes the array's name to a kernel: each device's own address comes
// from hipGetSymbolAddress, looked up once per device with that device current.  Calls that overlap on one
// device share the array and race.
constexpr int NOALLOC_BINS = 512;
constexpr size_t NOALLOC_DOUBLES = (size_t)1024 * 2 * NOALLOC_BINS + 2 * NOALLOC_BINS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

static int noalloc_scratch(double** out)
{
    int dev = -1;
    TH_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { set_error("noalloc scratch: device %d out of range", dev); return E_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return E_HIP;
        }
        g_scratch_ptr[dev] = (double*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

// The partial-shell rescale of four_energy.cc:324-336 on the device, with the host formula's operation order
// (Sk * Approx_number_modes_in_shell / Sk_count), so the values equal those of reduce_shell bit for bit.
__global__ void k_rescale_shells(double* __restrict__ sums, int nbins, int kin, int kkmax, int basis, double kf0,
                                 double kf1, double kf2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbins) return;
    const double c = sums[nbins + k];
    if (k > kin && k <= kkmax && c >= 1) {
        double v = M_PI * k * k / (kf0 * kf1 * kf2);
        double a = basis == 0 ? 2 * v : v;
        sums[k] = sums[k] * a / c;
    }
}

template <typename R>
int reduce_shell_noalloc(const Geo& g, const cplx<R>* A, const cplx<R>* B, double n, int nbins, double* Sk, double* cnt,
                         hipStream_t st)
{
    int nb = nblocks_for(g.ncplx());
    int kkmax = min_radius_outside(g);
    if (nbins != kkmax + 1) { set_error("reduce_shell_noalloc: nbins %d != shell_ek_size %d", nbins, kkmax + 1); return E_ARG; }
    if (nbins > NOALLOC_BINS || (size_t)nb * 2 * nbins + 2 * nbins > NOALLOC_DOUBLES) {
        set_error("th_shell_spectrum_noalloc: shell_ek_size %d exceeds the maximum %d of its static scratch", nbins,
                  NOALLOC_BINS);
        return E_ARG;
    }
    double* scr = nullptr;
    TRY(noalloc_scratch(&scr));
    double* out = scr + (size_t)nb * 2 * nbins;
    // every block writes all 2 nbins of its partial row, and the finalize writes all of `out`: nothing here
    // reads what an earlier call left in the array
    hipLaunchKernelGGL((k_binned<R, 0>), dim3(nb), dim3(RBLK), 2 * nbins * sizeof(double), st, gr(g), A,
                       (const cplx<R>*)nullptr, (const cplx<R>*)nullptr, B, (const cplx<R>*)nullptr,
                       (const cplx<R>*)nullptr, (int)n, kkmax, (const double*)nullptr, 0, nbins, scr);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(scr, nb, 2 * nbins, 2 * nbins, out, st);
    TH_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rescale_shells, dim3((nbins + 255) / 256), dim3(256), 0, st, out, nbins, max_radius_inside(g),
                       kkmax, g.basis, g.kf[0], g.kf[1], g.kf[2]);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(Sk, out, nbins * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cnt) TH_CHECK_HIP(hipMemcpyAsync(cnt, out + nbins, nbins * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    return OK;
}

template <typename R>
static int run_scalar(int mode, const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3], int npow,
                      double inner, double outer, double* out, RedScratch& rs, hipStream_t st)
{
    int nb = nblocks_for(g.ncplx() / 4 + 1);
    int r = rs.ensure((size_t)nb + 1, 1);
    if (r) return r;
    int kkmax = min_radius_outside(g);
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nb), dim3(RBLK), 0, st, gr(g), A[0], A[1], A[2], B[0], B[1], B[2], npow, kkmax,
                           inner, outer, rs.d);
    };
    if (mode == 0) launch(k_scalar<R, 0>);
    else if (mode == 1) launch(k_scalar<R, 1>);
    else launch(k_scalar<R, 2>);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nb, 1, 1, rs.d + nb, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, rs.d + nb, sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    *out = rs.h[0];
    return OK;
}

template <typename R>
int reduce_energy(const Geo& g, const cplx<R>* A, const cplx<R>* B, double* out, RedScratch& rs, hipStream_t st)
{
    const cplx<R>* As[3] = {A, nullptr, nullptr};
    const cplx<R>* Bs[3] = {B, nullptr, nullptr};
    return run_scalar<R>(0, g, As, Bs, 0, 0, 0, out, rs, st);
}
template <typename R>
int reduce_Sn(const Geo& g, const cplx<R>* A, double n, double* out, RedScratch& rs, hipStream_t st)
{
    const cplx<R>* As[3] = {A, nullptr, nullptr};
    const cplx<R>* Bs[3] = {nullptr, nullptr, nullptr};
    return run_scalar<R>(1, g, As, Bs, (int)n, 0, 0, out, rs, st);
}
template <typename R>
int reduce_shell_mult_single(const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3], double inner,
                             double outer, double* out, RedScratch& rs, hipStream_t st)
{
    return run_scalar<R>(2, g, A, B, 0, inner, outer, out, rs, st);
}

template <typename R>
int reduce_shell_mult_all(const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3], const double* radii_dev,
                          int nr, double* out, RedScratch& rs, hipStream_t st)
{
    int nb = nblocks_for(g.ncplx());
    int nbins = nr;
    int r = rs.ensure((size_t)nb * 2 * nbins + 2 * nbins, 2 * nbins);
    if (r) return r;
    double* dout = rs.d + (size_t)nb * 2 * nbins;
    int kin = max_radius_inside(g);
    hipLaunchKernelGGL((k_binned<R, 1>), dim3(nb), dim3(RBLK), 2 * nbins * sizeof(double), st, gr(g), A[0], A[1], A[2],
                       B[0], B[1], B[2], 0, kin, radii_dev, nr, nbins, rs.d);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nb, 2 * nbins, nbins, dout, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, dout, nbins * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < nbins; k++) out[k] = rs.h[k];
    return OK;
}

template <typename R> static int run_modal(int mode, const Geo& g, const cplx<R>* const V[3], double E, double* out,
                                           RedScratch& rs, hipStream_t st)
{
    const int nv = mode == 0 ? 3 : 1;
    int nb = nblocks_for(g.ncplx() / 4 + 1);
    int r = rs.ensure((size_t)nb * nv + nv, nv);
    if (r) return r;
    int kkmax = min_radius_outside(g);
    if (mode == 0)
        hipLaunchKernelGGL((k_modal<R, 0>), dim3(nb), dim3(RBLK), 0, st, gr(g), V[0], V[1], V[2], kkmax, E, rs.d);
    else
        hipLaunchKernelGGL((k_modal<R, 1>), dim3(nb), dim3(RBLK), 0, st, gr(g), V[0], V[1], V[2], kkmax, E, rs.d);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nb, nv, nv, rs.d + (size_t)nb * nv, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, rs.d + (size_t)nb * nv, nv * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    for (int v = 0; v < nv; v++) out[v] = rs.h[v];
    return OK;
}

// Compute_total_helicity_FOUR: out = {H1, H2, DH1 = 2 sum K^2 modal, DH2 = 2 H1}
template <typename R> int reduce_helicity(const Geo& g, const cplx<R>* const V[3], double* out, RedScratch& rs, hipStream_t st)
{
    double v[3];
    TRY(run_modal<R>(0, g, V, 0.0, v, rs, st));
    out[0] = v[0];
    out[1] = v[1];
    out[2] = 2.0 * v[2];
    out[3] = 2.0 * v[0];
    return OK;
}

// Get_entropy_FOUR / _SCFT with E = the three components' Get_total_energy
template <typename R> int reduce_entropy(const Geo& g, const cplx<R>* const V[3], double* out, RedScratch& rs, hipStream_t st)
{
    double E = 0, e;
    for (int c = 0; c < 3; c++) {
        TRY(reduce_energy<R>(g, V[c], nullptr, &e, rs, st));
        E += e;
    }
    return run_modal<R>(1, g, V, E, out, rs, st);
}

// Get_entropy_scalar_FOUR four_energy.cc:761-826 / _SCFT scft_energy.cc:750-800: one field
template <typename R> int reduce_entropy_scalar(const Geo& g, const cplx<R>* A, double* out, RedScratch& rs, hipStream_t st)
{
    double E;
    TRY(reduce_energy<R>(g, A, nullptr, &E, rs, st));
    const cplx<R>* V[3] = {A, nullptr, nullptr};
    return run_modal<R>(1, g, V, E, out, rs, st);
}

// Compute_shell_spectrum_helicity_FOUR four_energy.cc:602-671 / _SCFT scft_energy.cc:601-677 (with the
// partial-shell rescale)
template <typename R>
int reduce_shell_helicity(const Geo& g, const cplx<R>* const V[3], int nbins, double* H1, double* H2, double* H3,
                          RedScratch& rs, hipStream_t st)
{
    int nb = nblocks_for(g.ncplx() / 4 + 1);
    int kkmax = min_radius_outside(g);
    if (nbins != kkmax + 1) { set_error("reduce_shell_helicity: nbins %d != shell_ek_size %d", nbins, kkmax + 1); return E_ARG; }
    int r = rs.ensure((size_t)nb * 4 * nbins + 4 * nbins, 4 * nbins);
    if (r) return r;
    double* out = rs.d + (size_t)nb * 4 * nbins;
    hipLaunchKernelGGL((k_binned_hel<R>), dim3(nb), dim3(RBLK), 4 * nbins * sizeof(double), st, gr(g), V[0], V[1], V[2],
                       kkmax, nbins, rs.d);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nb, 4 * nbins, 4 * nbins, out, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, out, 4 * nbins * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    int kin = max_radius_inside(g);
    for (int k = 0; k < nbins; k++) {
        double c = rs.h[3 * nbins + k], s = 1.0;
        if (k > kin && k <= kkmax && c >= 1) s = approx_modes_in_shell(g, k) / c;
        H1[k] = rs.h[k] * s;
        H2[k] = rs.h[nbins + k] * s;
        H3[k] = rs.h[2 * nbins + k] * s;
    }
    return OK;
}

// Anis_max_Krho_radius_inside_FOUR four_inline.h:517-546 (+1: the rows of the cylindrical spectrum)
int cylinder_shell_ek_size(const Geo& g, int axis)
{
    const int d = g.dealias ? 3 : 2;
    const int N[3] = {g.N1, g.N2, g.N3};
    const int p = axis == 1 ? 1 : 0, q = axis == 3 ? 1 : 2;
    const double k1 = (N[p] / d) * g.kf[p], k2 = (N[q] / d) * g.kf[q];
    return (int)(k1 < k2 ? k1 : k2) + 1;
}

// mode 0 ring / 1 cylinder energy spectra (S1, S2); mode 2 / 3 the ring / cylinder helicity tables (S1 only)
template <typename R>
int reduce_aniso(const Geo& g, const cplx<R>* const V[3], int mode, int nbin, int n, int axis, double* S1, double* S2,
                 RedScratch& rs, hipStream_t st)
{
    const bool hel = mode >= 2;
    mode &= 1;
    if (g.basis != 0 || g.N2 < 2 || g.N3 < 4) { set_error("anisotropic spectra: 3-D FOUR boxes only"); return E_UNSUPPORTED; }
    if (axis < 1 || axis > 3 || nbin < 1 || (n != 0 && n != 2 && n != 4)) {
        set_error("anisotropic spectra: axis must be 1..3, bins >= 1, n in {0, 2, 4}");
        return E_ARG;
    }
    AnisoP ap;
    ap.axis = axis;
    ap.nbin = nbin;
    ap.cols = nbin + 1;
    ap.n = n;
    if (mode == 0) {
        ap.kmaxi = max_radius_inside(g);
        ap.rows = ap.kmaxi + 1;
        ap.pmax = M_PI / 2;                               // Get_max_polar_angle_FOUR four_inline.h:549-554
        ap.dpar = ap.pmax / nbin;                         // Cvf.cc:85-106, linear sectors
    } else {
        ap.rows = cylinder_shell_ek_size(g, axis);
        ap.kmaxi = ap.rows - 1;
        const int d = g.dealias ? 3 : 2;
        const int N[3] = {g.N1, g.N2, g.N3};
        ap.pmax = (N[axis - 1] / d) * g.kf[axis - 1];     // Anis_max_Kpll_FOUR four_inline.h:484-514
        ap.dpar = (ap.pmax - 0.0) / nbin;                 // Cvf.cc:166-190, linear slabs from 0
    }
    const int nb2 = 2 * ap.rows * ap.cols;
    if ((size_t)nb2 * sizeof(double) > 64 * 1024) {
        set_error("anisotropic spectra: %d x %d bins exceed one workgroup's 64 KiB of LDS", ap.rows, ap.cols);
        return E_ARG;
    }
    const int nblk = nblocks_for(g.ncplx() / 4 + 1);
    int r = rs.ensure((size_t)nblk * nb2 + nb2, nb2);
    if (r) return r;
    double* out = rs.d + (size_t)nblk * nb2;
    const size_t lds = nb2 * sizeof(double);
    if (mode == 0 && !hel)
        hipLaunchKernelGGL((k_aniso<R, 0, false>), dim3(nblk), dim3(RBLK), lds, st, gr(g), V[0], V[1], V[2], ap, rs.d);
    else if (mode == 1 && !hel)
        hipLaunchKernelGGL((k_aniso<R, 1, false>), dim3(nblk), dim3(RBLK), lds, st, gr(g), V[0], V[1], V[2], ap, rs.d);
    else if (mode == 0)
        hipLaunchKernelGGL((k_aniso<R, 0, true>), dim3(nblk), dim3(RBLK), lds, st, gr(g), V[0], V[1], V[2], ap, rs.d);
    else
        hipLaunchKernelGGL((k_aniso<R, 1, true>), dim3(nblk), dim3(RBLK), lds, st, gr(g), V[0], V[1], V[2], ap, rs.d);
    TH_CHECK_HIP(hipGetLastError());
    launch_finalize(rs.d, nblk, nb2, nb2, out, st);
    TH_CHECK_HIP(hipGetLastError());
    TH_CHECK_HIP(hipMemcpyAsync(rs.h, out, nb2 * sizeof(double), hipMemcpyDeviceToHost, st));
    TH_CHECK_HIP(hipStreamSynchronize(st));
    const int nb = ap.rows * ap.cols;
    for (int k = 0; k < nb; k++) {
        S1[k] = rs.h[k];
        if (S2) S2[k] = rs.h[nb + k];
    }
    return OK;
}

#define TH_RINST(R)                                                                                                  \
    template int reduce_shell<R>(const Geo&, const cplx<R>*, const cplx<R>*, double, int, double*, double*,         \
                                 RedScratch&, hipStream_t, bool);                                                    \
    template int reduce_energy<R>(const Geo&, const cplx<R>*, const cplx<R>*, double*, RedScratch&, hipStream_t);   \
    template int reduce_Sn<R>(const Geo&, const cplx<R>*, double, double*, RedScratch&, hipStream_t);               \
    template int reduce_shell_mult_single<R>(const Geo&, const cplx<R>* const*, const cplx<R>* const*, double,     \
                                             double, double*, RedScratch&, hipStream_t);                             \
    template int reduce_shell_mult_all<R>(const Geo&, const cplx<R>* const*, const cplx<R>* const*, const double*, \
                                          int, double*, RedScratch&, hipStream_t);                                  \
    template int reduce_helicity<R>(const Geo&, const cplx<R>* const*, double*, RedScratch&, hipStream_t);          \
    template int reduce_entropy<R>(const Geo&, const cplx<R>* const*, double*, RedScratch&, hipStream_t);           \
    template int reduce_entropy_scalar<R>(const Geo&, const cplx<R>*, double*, RedScratch&, hipStream_t);         \
    template int reduce_shell_helicity<R>(const Geo&, const cplx<R>* const*, int, double*, double*, double*,         \
                                          RedScratch&, hipStream_t);                                                \
    template int reduce_aniso<R>(const Geo&, const cplx<R>* const*, int, int, int, int, double*, double*,             \
                                 RedScratch&, hipStream_t);                                                          \
    template int reduce_shell_noalloc<R>(const Geo&, const cplx<R>*, const cplx<R>*, double, int, double*, double*, \
                                         hipStream_t);                                                               \
    template int fill_ring<R>(const Geo&, const cplx<R>*, cplx<R>*, const RingP&, int, int, hipStream_t);             \
    template int reduce_ring_mult<R>(const Geo&, const cplx<R>* const*, const cplx<R>* const*, const RingP&, double*, \
                                     RedScratch&, hipStream_t);
TH_RINST(float)
TH_RINST(double)

}  // namespace th
