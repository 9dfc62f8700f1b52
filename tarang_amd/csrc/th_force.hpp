// th_force.hpp -- device helpers of the energy / helicity forcing (force_field_proc 1 and 2), shared by
// the reference-layout band kernels (th_force.hip) and the fused step (th_fused.hip).
#pragma once
#include <math.h>
#include "th_internal.hpp"

namespace th {

constexpr double MYEPS = 1e-15;    // extern_def_vars.h:81
constexpr double MYEPS2 = 1e-5;    // extern_def_vars.h:82

__host__ __device__ __forceinline__ double kmag3(double a, double b, double c)
{
#pragma clang fp contract(off)
    return sqrt(a * a + b * b + c * c);
}

// Force of one mode (FP64) from its stored velocity v, before the reality condition: F = alpha v +
// beta (i K x v) on inner < |K| <= outer with modal energy > MYEPS, alpha and beta of scheme 1 or 2,
// the vorticity term only for beta > MYEPS (Force_alpha_beta force_alpha_beta.cc:60-89), zero where
// Dealias_force applies.
template <typename R>
__device__ __forceinline__ void mode_force(const ForceBand& b, const cplx<R> v[3], int lx, int ly, int lz,
                                           double fr[3], double fi[3])
{
    fr[0] = fr[1] = fr[2] = fi[0] = fi[1] = fi[2] = 0.0;
    if (b.proc == 3) {
        // Setup_Taylor_Green_force_field force_model.cc:57-88 (FOUR): F1 = -i s_x amp/8, F2 = i s_y amp/8 on
        // (s_x k0, s_y k0, k0), the field amp (sin k0x cos k0y cos k0z, -cos k0x sin k0y cos k0z, 0)
        const int kx = kx_four(lx, b.N1), ky = ky_of(ly, b.N2);
        if (lz == b.k0 && (kx == b.k0 || kx == -b.k0) && (ky == b.k0 || ky == -b.k0)) {
            const double a = b.per_mode / 8;
            fi[0] = kx > 0 ? -a : a;
            fi[1] = ky > 0 ? a : -a;
        }
        return;
    }
    if (b.proc == 4) {
        // Setup_ABC_force_field force_model.cc:122-160 (FOUR): amp (B cos k0y + C sin k0z, A sin k0x + C cos k0z,
        // A cos k0x + B sin k0y) on (0, +-k0, 0), (0, 0, k0) and (+-k0, 0, 0), each value X * amp / 2
        const int kx = kx_four(lx, b.N1), ky = ky_of(ly, b.N2), k0 = b.k0;
        const double amp = b.per_mode, A = b.abc[0] * amp / 2, B = b.abc[1] * amp / 2, C = b.abc[2] * amp / 2;
        if (kx == 0 && lz == 0 && (ky == k0 || ky == -k0)) { fr[0] = B; fi[2] = ky > 0 ? -B : B; }
        if (kx == 0 && ky == 0 && lz == k0) { fi[0] = -C; fr[1] = C; }
        if (ky == 0 && lz == 0 && (kx == k0 || kx == -k0)) { fi[1] = kx > 0 ? -A : A; fr[2] = A; }
        return;
    }
    const double kk[3] = {kx_four(lx, b.N1) * b.f1, ky_of(ly, b.N2) * b.f2, lz * b.f3};
    const double km = kmag3(kk[0], kk[1], kk[2]);
    if (!(km > b.inner && km <= b.outer)) return;
    if (b.proc == 2 && !(b.per_mode > MYEPS)) return;
    if (b.dealias_force && !kept_mode(0, 1, b.N1, b.N2, b.N3, lx, ly, lz)) return;
    double vr[3], vi[3], E = 0.0;
    for (int c = 0; c < 3; c++) {
        vr[c] = (double)v[c].x;
        vi[c] = (double)v[c].y;
        E += vr[c] * vr[c] + vi[c] * vi[c];
    }
    E = E / 2;                                               // CV_Modal_energy Cvf.cc:595-600
    if (!(E > b.emin)) return;
    const double hel = kk[0] * (vr[1] * vi[2] - vr[2] * vi[1]) + kk[1] * (vr[2] * vi[0] - vr[0] * vi[2]) +
                       kk[2] * (vr[0] * vi[1] - vr[1] * vi[0]);       // Get_Modal_helicity_FOUR
    const double sk = hel / (km * E);
    double alpha, beta;
    if (b.proc == 1) {
        const double temp = b.per_mode / (2 * E);
        if (fabs(sk * sk - 1) > MYEPS2) {
            alpha = temp * (1 - sk * b.p4) / (1 - sk * sk);
            beta = temp * (b.p4 - sk) / (1 - sk * sk);
        } else {
            alpha = temp / 2;
            beta = alpha / (sk * km);
        }
    } else {
        const double temp1 = sqrt(b.per_mode / E);
        if (fabs(sk * sk - 1) > MYEPS2) {
            const double temp2 = sqrt((1 + b.p4) / (1 + sk)), temp3 = sqrt((1 - b.p4) / (1 - sk));
            alpha = (temp1 / 2) * (temp2 + temp3);
            beta = (temp1 / (2 * km)) * (temp2 - temp3);
        } else {
            alpha = temp1;
            beta = 0.0;
        }
    }
    for (int c = 0; c < 3; c++) {
        fr[c] = alpha * vr[c];
        fi[c] = alpha * vi[c];
    }
    if (beta > MYEPS) {                                      // vorticity i K x V only for beta > MYEPS
        for (int c = 0; c < 3; c++) {
            const int a1 = (c + 1) % 3, a2 = (c + 2) % 3;
            const double cr = kk[a1] * vr[a2] - kk[a2] * vr[a1], ci = kk[a1] * vi[a2] - kk[a2] * vi[a1];
            fr[c] += beta * -ci;
            fi[c] += beta * cr;
        }
    }
}

// Box thread t -> stored mode (lx, ly, lz): kx = tx - kx_max mapped by Get_lx_FOUR four_inline.h:103-106,
// ky likewise by Get_ly3D_FOUR :138 (or every index when the box spans the axis).
__host__ __device__ __forceinline__ void band_mode(const ForceBand& b, long t, int& lx, int& ly, int& lz)
{
    lz = (int)(t % b.nz);
    const long r = t / b.nz;
    const int ty = (int)(r % b.ny), tx = (int)(r / b.ny);
    if (b.fullx) {
        lx = tx;
    } else {
        const int kx = tx - b.kxm;
        lx = kx >= 0 ? kx : kx + b.N1;
    }
    if (b.fully) {
        ly = ty;
    } else {
        const int ky = ty - b.kym;
        ly = ky >= 0 ? ky : ky + b.N2;
    }
}

// Satisfy_reality_array (universal_misc.cc:51-109) on the kz = 0 and kz = N3/2 planes: true when the
// mode is overwritten with the conjugate of mode (sx, sy, lz); else (sx, sy) = (lx, ly).
__host__ __device__ __forceinline__ bool reality_source(const ForceBand& b, int lx, int ly, int lz, int& sx, int& sy)
{
    sx = lx;
    sy = ly;
    if (lz != 0 && lz != b.N3 / 2) return false;
    const int kx = kx_four(lx, b.N1);
    if (ly > b.N2 / 2) {
        sx = (-kx >= 0) ? -kx : -kx + b.N1;
        sy = -ky_of(ly, b.N2);
        return true;
    }
    if (ly == 0 && kx < 0) {
        sx = -kx;
        return true;
    }
    return false;
}

}  // namespace th
