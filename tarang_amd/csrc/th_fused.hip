// th_fused.hip -- dealias-pruned, fused pseudo-spectral RK stage (FOUR basis).
//
// State layout: only the modes kept by Dealias_array (universal_misc.cc:118-159) are stored, in the
// reference's (kx, ky, kz) order, compacted per axis:  S[cx][cy][cz] with cz padded to KZP (128-byte
// rows).  A stage of the IF-RK4 scheme (time_advance.cc:100-155) is five passes, each a batch of
// LDS-staged line FFTs over 128-byte column tiles; nothing else touches HBM:
//   1  x-inverse   S[cx][m]    -> A[x][m]          (m = (cy,cz); zero rows of the gap are never read)
//   2  y-inverse   A[x][cy][z] -> B[x][y][z]
//   3  z-fused     B[x][y][:]  -> c2r, real-space products (Array_real_mult), r2c -> C[x][y][kz<N3/3]
//   4  y-forward   C[x][y][z]  -> D[x][cy][z]      (only kept ky written)
//   5  x-forward   D[x][m]     -> kept kx only; epilogue: N_i = sum_j i K_j F(u_i u_j) (RSprod.cc),
//                  projection -P(K) N (compute_pressure.cc, compute_rhs.cc), FP64 integrating factor
//                  exp(-nu K^2 a dt) and the RK4 stage update (single_time_step.cc, Compute_RK4_parts)
// MHD uses the Elsasser form (compute_nlin_VF_MHD.cc) with 6 inverse and 9 product lines; the energy
// transfer (ET_compute_nlin.cc) uses helper U and a shell-masked giver G and bins 2 mult Re(N U*)
// (four_ET.cc) in the pass-5 epilogue instead of writing N.
#include <math.h>
#include <string.h>
#include <array>
#include <vector>
#include "../../include/tarang_hip.h"
#include "th_fft.hpp"
#include "th_internal.hpp"
#include "th_plan.hpp"

namespace th {

// Line of logical length N with nk stored entries: logical [0,h) and [N-(nk-h), N).
struct LMap {
    int N, nk, h;
};
__host__ __device__ __forceinline__ bool lkept(const LMap& m, int l) { return l < m.h || l >= m.N - (m.nk - m.h); }
__host__ __device__ __forceinline__ int lq(const LMap& m, int l) { return l < m.h ? l : l - (m.N - m.nk); }
__host__ __device__ __forceinline__ int ll(const LMap& m, int q) { return q < m.h ? q : q + (m.N - m.nk); }

// Workgroups are dispatched round-robin over the 8 XCDs (linear id mod 8), each XCD with its own
// L2.  xcd_tile renumbers a linear workgroup id so that every XCD gets one contiguous range of
// logical tiles: neighbouring tiles, which share 128-byte lines, then meet in the same L2.
// A zero the compiler cannot see through.  Offsetting a table pointer by it inside a rolled loop
// keeps the table loads in the loop (L1 hits) instead of hoisting every twiddle into registers
// for the whole kernel, which would cost occupancy.
__device__ __forceinline__ int opaque_zero()
{
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

// Fused states are 2/3-dealiased, so along x only l < L/3 and l > L - L/3 are kept.  Lane t of an
// x line holds positions t + P m (m = 0..E-1, register i = (m % (E/RL)) RL + m / (E/RL)); xdrop
// tells, at compile time, whether register i holds no kept position for any lane.
template <int L, int P, int E, int RL> __host__ __device__ constexpr bool xdrop(int i)
{
    const int m = i / RL + (i % RL) * (E / RL), h = L / 3;
    return m * P >= h && m * P + P - 1 <= L - h;
}

__device__ __forceinline__ int xcd_tile(int bid, int nblk)
{
    const int xcd = bid & 7, loc = bid >> 3;
    const int per = nblk >> 3, rem = nblk & 7;
    return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + loc;
}

struct FG {
    int N1, N2, N3, LZ;
    LMap mx, my;
    int nkz, KZP;
    long sx;     // x-row stride of the state and of A/D: nky * KZP
    long sB;     // x-plane stride of B/C: N2 * KZP
    long cS;     // component stride of the state: nkx * sx
    long cA;     // component stride of A/D: N1 * sx
    long cB;     // component stride of B/C: N1 * sB
    long bS, bA, bB, bC;   // box strides of the state, A/D, B and C buffers (batched boxes of one size)
    int nbox;              // boxes of this size advanced by one launch per pass
    double f1, f2, f3;
    double norm; // 1 / (N1 N2 N3)
};

struct StageCoef {
    int stage;          // 1..4
    double dt;
    double fh[2];       // (-diss dt) * 0.5 for field groups 0 (U) and 1 (B / scalar)
    double f1[2];       // (-diss dt) * 1.0
};

struct ShellSel {       // giver mask inner < |K| <= outer (Fill_array_shell universal_ET.cc:45-65)
    double inner, outer;
};

__device__ __forceinline__ double kmag_f(double a, double b, double c)
{
#pragma clang fp contract(off)
    return sqrt(a * a + b * b + c * c);
}
__device__ __forceinline__ double ksq_f(double a, double b, double c)
{
#pragma clang fp contract(off)
    return a * a + b * b + c * c;
}

template <typename R> __device__ __forceinline__ cplx<R> ld(const cplx<R>* p) { return *p; }

// ------------------------------------------------------------------------------------------
// pass 1: x-inverse.  MODE 0: copy comps; 1: Elsasser Z+ = U + B (comps 0-2), Z- = U - B (3-5);
// 2: giver G = U masked to the shell sel.
// ------------------------------------------------------------------------------------------
template <typename R, int L, int P, int T, int MODE>
__global__ __launch_bounds__(T * P) void k_fx_inv(FG g, const cplx<R>* __restrict__ S, cplx<R>* __restrict__ A,
                                                  const cplx<R>* __restrict__ tw, ShellSel sel)
{
    using F = LineFFT<R, L, P, +1>;
    __shared__ cplx<R> lds[L * T];
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int ntz = (g.nkz + T - 1) / T;
    const int cy = blockIdx.x / ntz;
    const int cz = (blockIdx.x % ntz) * T + c;
    const bool ok = cz < g.nkz;
    const long m = (long)cy * g.KZP + cz;
    const int oc = blockIdx.y;
    S += blockIdx.z * g.bS;
    A += blockIdx.z * g.bA;
    double Ky = 0, Kz = 0;
    if (MODE == 2) {
        Ky = ky_of(ll(g.my, cy), g.N2) * g.f2;
        Kz = cz * g.f3;
    }
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int l = F::in_pos(t, i);
        cplx<R> a = czero<R>();
        if (ok && lkept(g.mx, l)) {
            const long off = (long)lq(g.mx, l) * g.sx + m;
            if (MODE == 0) {
                a = S[oc * g.cS + off];
            } else if (MODE == 1) {
                const int cc = oc % 3;
                cplx<R> u = S[cc * g.cS + off], b = S[(3 + cc) * g.cS + off];
                a = (oc < 3) ? (u + b) : (u - b);
            } else {
                double km = kmag_f(kx_four(l, g.N1) * g.f1, Ky, Kz);
                if ((km > sel.inner) && (km <= sel.outer)) a = S[oc * g.cS + off];
            }
        }
        v[i] = a;
    }
    F::template run<T>(v, lds + c, t, tw);
    if (ok) {
        cplx<R>* out = A + oc * g.cA + m;
#pragma unroll
        for (int i = 0; i < F::E; i++) out[(long)F::out_pos(t, i) * g.sx] = v[i];
    }
}

// pass 2: y-inverse, A[x][cy][cz] -> B[x][y][cz]
template <typename R, int L, int P, int T>
__global__ __launch_bounds__(T * P) void k_fy_inv(FG g, const cplx<R>* __restrict__ A, cplx<R>* __restrict__ B,
                                                  const cplx<R>* __restrict__ tw, int ncomp)
{
    using F = LineFFT<R, L, P, +1>;
    __shared__ cplx<R> lds[L * T];
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int cz = blockIdx.x * T + c;
    const bool ok = cz < g.nkz;
    const int x = blockIdx.y, comp = blockIdx.z % ncomp, box = blockIdx.z / ncomp;
    A += box * g.bA;
    B += box * g.bB;
    const cplx<R>* a = A + comp * g.cA + (long)x * g.sx + cz;
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int l = F::in_pos(t, i);
        v[i] = (ok && lkept(g.my, l)) ? a[(long)lq(g.my, l) * g.KZP] : czero<R>();
    }
    F::template run<T>(v, lds + c, t, tw);
    if (ok) {
        cplx<R>* b = B + comp * g.cB + (long)x * g.sB + cz;
#pragma unroll
        for (int i = 0; i < F::E; i++) b[(long)F::out_pos(t, i) * g.KZP] = v[i];
    }
}

// pass 3: fused z.  Input lines hold kz < nkz; c2r to the z-packed real line, products, r2c.
// MODE 0 (NS):  in u0..u2,            out u0u0 u1u1 u2u2 u0u1 u1u2 u2u0          (RSprod.cc:52-77)
// MODE 1 (MHD): in Z+0..2, Z-0..2,    out P(i,j) = Z+_i Z-_j at 3i+j             (RSprod.cc:117-151)
// MODE 2 (ET):  in H0..2,  G0..2,     out P(i,j) = G_i H_j at 3i+j               (ET_RSprod.cc:49-85)
template <int MODE> struct ZProd;
template <> struct ZProd<0> {
    static constexpr int NIN = 3, NOUT = 6;
    static constexpr int a(int p) { return p < 3 ? p : p - 3; }
    static constexpr int b(int p) { return p < 3 ? p : (p == 5 ? 0 : p - 2); }
};
template <> struct ZProd<1> {
    static constexpr int NIN = 6, NOUT = 9;
    static constexpr int a(int p) { return p / 3; }
    static constexpr int b(int p) { return 3 + p % 3; }
};
template <> struct ZProd<2> {
    static constexpr int NIN = 6, NOUT = 9;
    static constexpr int a(int p) { return 3 + p / 3; }
    static constexpr int b(int p) { return p % 3; }
};

template <typename R> __device__ __forceinline__ cplx<R> shfl_c(const cplx<R>& v, int src, int width)
{
    return mk<R>(__shfl(v.x, src, width), __shfl(v.y, src, width));
}

// Lane t of a z line holds the points at positions t + P m, m = 0..E-1, in a register order fixed
// by the radix Rr of the FFT stage that reads (first stage) or writes (last stage) them.
template <int E, int Rr> __device__ __forceinline__ constexpr int zreg(int m)
{
    return (m % (E / Rr)) * Rr + m / (E / Rr);
}

template <typename R, int LZ, int P, int TZ, int MODE>
__global__ __launch_bounds__(TZ * P) void k_fz(FG g, const cplx<R>* __restrict__ B, cplx<R>* __restrict__ Cout,
                                               const cplx<R>* __restrict__ twL, const cplx<R>* __restrict__ twN)
{
    using FI = LineFFT<R, LZ, P, +1>;
    using FF = LineFFT<R, LZ, P, -1>;
    using ZP = ZProd<MODE>;
    constexpr int E = FI::E;
    static_assert(64 % P == 0, "a z line's threads must share one wavefront");
    constexpr int LZS = ZPad<LZ>::line;   // padded line stride in LDS
    __shared__ cplx<R> lds[TZ * LZS];
    const int t = threadIdx.x % P, c = threadIdx.x / P;
    const long q = (long)blockIdx.x * TZ + c;
    const bool ok = q < (long)g.N1 * g.N2;
    const long off = q * g.KZP;
    B += blockIdx.y * g.bB;
    Cout += blockIdx.y * g.bC;
    cplx<R>* Z = lds + c * LZS;
    // The mirror point LZ - (t + P m) sits in lane (P - t) mod P at m' = E - 1 - m (t > 0) or
    // m' = (E - m) mod E (t = 0): the half-length packing reads it by a lane shuffle.
    const int tp = (P - t) % P;
    cplx<R> twk[E];   // W_N3^(t + P m)
#pragma unroll
    for (int m = 0; m < E; m++) twk[m] = twN[t + P * m];
    cplx<R> u[ZP::NIN][E];
#pragma unroll
    for (int a = 0; a < ZP::NIN; a++) {
        const cplx<R>* X = B + a * g.cB + off;
        cplx<R> xa[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + P * m;
            xa[m] = (ok && k < g.nkz) ? X[k] : czero<R>();
        }
        const cplx<R> xny = (ok && t == 0 && LZ < g.nkz) ? X[LZ] : czero<R>();   // kz = N3/2 (undealiased)
#pragma unroll
        for (int m = 0; m < E; m++) {
            cplx<R> xr = shfl_c(xa[E - 1 - m], tp, P);
            cplx<R> xb = (t == 0) ? (m == 0 ? xny : xa[(E - m) % E]) : xr;
            cplx<R> xv = xa[m];
            if (t == 0 && m == 0) { xv.y = R(0); xb.y = R(0); }
            xb.y = -xb.y;
            cplx<R> s = xv + xb, d = xv - xb;
            cplx<R> w = conjg(twk[m]);
            u[a][zreg<E, FI::R0>(m)] = s + mul_i(cmul(w, d));
        }
        FI::template run<1, true, ZPad<LZ>>(u[a], Z, t, twL);
    }
    // One product line: forward half-length FFT, r2c unpacking, store of the kept kz.
    auto product_out = [&](cplx<R> (&w)[E], int p) {
        if constexpr (FI::R0 != FI::RL) {
            wave_sync();
#pragma unroll
            for (int i = 0; i < E; i++) Z[FI::out_pos(t, i)] = w[i];
            wave_sync();
#pragma unroll
            for (int i = 0; i < E; i++) w[i] = Z[FF::in_pos(t, i)];
        }
        FF::template run<1, true, ZPad<LZ>>(w, Z, t, twL);
        const cplx<R>* twp = twN + opaque_zero();   // W_N3^(t + P m), reloaded per product (L1)
        cplx<R>* X = Cout + p * g.cB + off;
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + P * m;
            cplx<R> yr = shfl_c(w[zreg<E, FF::RL>(E - 1 - m)], tp, P);
            cplx<R> zk = w[zreg<E, FF::RL>(m)];
            cplx<R> zc = conjg((t == 0) ? w[zreg<E, FF::RL>((E - m) % E)] : yr);
            cplx<R> e = cscale(zk + zc, R(0.5));
            cplx<R> tt = cmul(twp[t + P * m], zk - zc);
            if (ok && k < g.nkz) X[k] = mk<R>(e.x + R(0.5) * tt.y, e.y - R(0.5) * tt.x);
        }
        if (ok && t == 0 && LZ < g.nkz) {   // kz = N3/2 (undealiased): Y[0] - ... with W_N3^(N3/2) = -1
            cplx<R> z0 = w[zreg<E, FF::RL>(0)];
            X[LZ] = mk<R>(z0.x - z0.y, R(0));
        }
    };
    if constexpr (sizeof(R) == 4) {
        // FP32: a rolled loop with operands picked by selects keeps the compiler from
        // interleaving products and hoisting their twiddles (registers, occupancy).
#pragma unroll 1
        for (int p = 0; p < ZP::NOUT; p++) {
            const int ia = ZP::a(p), ib = ZP::b(p);
            cplx<R> w[E];
#pragma unroll
            for (int i = 0; i < E; i++) {
                cplx<R> x = u[0][i], y = u[0][i];
#pragma unroll
                for (int a = 1; a < ZP::NIN; a++) {
                    x = (ia == a) ? u[a][i] : x;
                    y = (ib == a) ? u[a][i] : y;
                }
                w[i] = mk<R>(x.x * y.x, x.y * y.y);
            }
            product_out(w, p);
        }
    } else {
        // FP64: selects over the arrays would become dynamic indexing (scratch); unroll instead
#pragma unroll
        for (int p = 0; p < ZP::NOUT; p++) {
            const int ia = ZP::a(p), ib = ZP::b(p);
            cplx<R> w[E];
#pragma unroll
            for (int i = 0; i < E; i++) w[i] = mk<R>(u[ia][i].x * u[ib][i].x, u[ia][i].y * u[ib][i].y);
            product_out(w, p);
        }
    }
}

// pass 4: y-forward, C[x][y][cz] -> D[x][cy][cz] (kept ky only)
template <typename R, int L, int P, int T>
__global__ __launch_bounds__(T * P) void k_fy_fwd(FG g, const cplx<R>* __restrict__ Cin, cplx<R>* __restrict__ D,
                                                  const cplx<R>* __restrict__ tw, int ncomp)
{
    using F = LineFFT<R, L, P, -1>;
    __shared__ cplx<R> lds[L * T];
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int cz = blockIdx.x * T + c;
    const bool ok = cz < g.nkz;
    const int x = blockIdx.y, comp = blockIdx.z % ncomp, box = blockIdx.z / ncomp;
    Cin += box * g.bC;
    D += box * g.bA;
    const cplx<R>* a = Cin + comp * g.cB + (long)x * g.sB + cz;
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) v[i] = ok ? a[(long)F::in_pos(t, i) * g.KZP] : czero<R>();
    F::template run<T>(v, lds + c, t, tw);
    if (ok) {
        cplx<R>* d = D + comp * g.cA + (long)x * g.sx + cz;
#pragma unroll
        for (int i = 0; i < F::E; i++) {
            const int l = F::out_pos(t, i);
            if (lkept(g.my, l)) d[(long)lq(g.my, l) * g.KZP] = v[i];
        }
    }
}

// ------------------------------------------------------------------------------------------
// pass 5: x-forward of the product lines with the nonlinear-term contraction accumulated in
// registers, then the projection and the IF-RK4 stage update.
//   MODE 0 (NS):  N_i = sum_j i K_j P_ij                         (Derivative_RSprod_VV RSprod.cc:94-105)
//   MODE 1 (MHD): N+_i = sum_j i K_j P(i,j), N-_i = sum_j i K_j P(j,i); N_u = (N+ + N-)/2,
//                 N_b = (N+ - N-)/2 (Elsasser_to_UB_nlin elsasser.cc:117-130); U projected, B not.
// ------------------------------------------------------------------------------------------
template <typename R> __device__ __forceinline__ void acc_iK(cplx<R>& n, R K, const cplx<R>& v)
{
    n.x -= K * v.y;
    n.y += K * v.x;
}

struct Dc {
    double x, y;
};
__device__ __forceinline__ Dc dmk(double a, double b) { Dc r; r.x = a; r.y = b; return r; }
template <typename R> __device__ __forceinline__ Dc dl(const cplx<R>& a) { return dmk((double)a.x, (double)a.y); }
template <typename R> __device__ __forceinline__ cplx<R> ds(const Dc& a) { return mk<R>((R)a.x, (R)a.y); }
__device__ __forceinline__ Dc dsc(const Dc& a, double s) { return dmk(a.x * s, a.y * s); }
__device__ __forceinline__ Dc dadd(const Dc& a, const Dc& b) { return dmk(a.x + b.x, a.y + b.y); }

// One RK4 stage update of a field value V0 (state), stage field Vs and accumulator tot
// (single_time_step.cc:54-86 with a,b,c of time_advance.cc:120-153).
template <typename R>
__device__ __forceinline__ void rk_update(int stage, double dt, double eh, double e1, const Dc& rhs, cplx<R>* S,
                                          cplx<R>* Vs, cplx<R>* acc)
{
    Dc V0 = dl(*S);
    if (stage == 1) {
        Dc v = dadd(V0, dsc(rhs, 0.5 * dt));
        *Vs = ds<R>(dsc(v, eh));
        *acc = ds<R>(dsc(dsc(rhs, e1), dt / 6));
    } else if (stage == 2) {
        *Vs = ds<R>(dadd(dsc(V0, eh), dsc(rhs, 0.5 * dt)));
        Dc a = dl(*acc);
        *acc = ds<R>(dadd(a, dsc(dsc(rhs, eh), dt / 3)));
    } else if (stage == 3) {
        Dc v = dadd(dsc(V0, eh), dsc(rhs, 1.0 * dt));
        *Vs = ds<R>(dsc(v, eh));
        Dc a = dl(*acc);
        *acc = ds<R>(dadd(a, dsc(dsc(rhs, eh), dt / 3)));
    } else {
        Dc a = dadd(dl(*acc), dsc(rhs, dt / 6));
        *S = ds<R>(dadd(dsc(V0, e1), a));
    }
}

template <typename R, int L, int P, int T, int MODE>
__global__ __launch_bounds__(T * P) __attribute__((amdgpu_waves_per_eu((MODE == 0 && sizeof(R) == 4) ? 4 : 1))) void k_fx_fwd(FG g, const cplx<R>* __restrict__ D, cplx<R>* S, cplx<R>* Vs,
                                                  cplx<R>* acc, const cplx<R>* __restrict__ tw,
                                                  const StageCoef* __restrict__ coefs)
{
    using F = LineFFT<R, L, P, -1>;
    constexpr int E = F::E;
    constexpr int NP = (MODE == 0) ? 6 : 9;
    constexpr int NA = (MODE == 0) ? 3 : 6;
    __shared__ cplx<R> lds[L * T];
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int ntz = (g.nkz + T - 1) / T;
    const int lin = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, box = lin / gridDim.x;
    const int cy = tile / ntz;
    const int cz = (tile % ntz) * T + c;
    const bool ok = cz < g.nkz;
    const long m = (long)cy * g.KZP + cz;
    const StageCoef sc = coefs[box];
    D += box * g.bA;
    S += box * g.bS;
    Vs += box * g.bS;
    acc += box * g.bS;
    const R Ky = (R)(ky_of(ll(g.my, cy), g.N2) * g.f2);
    const R Kz = (R)(cz * g.f3);
    cplx<R> Nacc[NA][E];
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int i = 0; i < E; i++) Nacc[a][i] = czero<R>();
    // The product loop stays rolled (one product line live at a time).  Each product p adds
    // i (alpha_a Kx + beta_a) P_p to accumulator a, so accumulators are indexed statically.
    // Addresses split into a uniform row base (SGPRs) and one 32-bit lane offset: lane (c, t)
    // reads column cz = tile * T + c of row in_pos(t, i) = t + in_pos(0, i).
    const long mu = (long)cy * g.KZP + (long)(tile % ntz) * T;
    const int lane_off = c + t * (int)g.sx;
#pragma unroll 1
    for (int p = 0; p < NP; p++) {
        cplx<R> v[E];
        const cplx<R>* src = D + p * g.cA + mu;
#pragma unroll
        for (int i = 0; i < E; i++) {
            const cplx<R>* row = src + (long)F::in_pos(0, i) * g.sx;
            v[i] = ok ? row[lane_off] : czero<R>();
        }
        F::template run<T>(v, lds + c, t, tw + opaque_zero());
        // contributions (a1 <- K_k1) and (a2 <- K_k2) of product p; k = 0 means Kx (per row)
        int a1, k1, a2, k2;
        if (MODE == 0) {           // P_ij, (i,j) = (0,0),(1,1),(2,2),(0,1),(1,2),(2,0): N_i += iK_j P, N_j += iK_i P
            const int pi = p < 3 ? p : p - 3, pj = p < 3 ? p : (p == 5 ? 0 : p - 2);
            a1 = pi; k1 = pj;
            a2 = (pi != pj) ? pj : -1; k2 = pi;
        } else {                   // P(pi,pj) = Z+_pi Z-_pj: N+_pi += iK_pj P, N-_pj += iK_pi P
            const int pi = p / 3, pj = p % 3;
            a1 = pi; k1 = pj;
            a2 = 3 + pj; k2 = pi;
        }
        R al[NA], be[NA];
#pragma unroll
        for (int a = 0; a < NA; a++) {
            al[a] = R(0);
            be[a] = R(0);
            if (a == a1) { if (k1 == 0) al[a] += R(1); else be[a] += (k1 == 1 ? Ky : Kz); }
            if (a == a2) { if (k2 == 0) al[a] += R(1); else be[a] += (k2 == 1 ? Ky : Kz); }
        }
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (xdrop<L, P, E, F::RL>(i)) continue;
            const R Kx = (R)(kx_four(F::out_pos(t, i), g.N1) * g.f1);
#pragma unroll
            for (int a = 0; a < NA; a++) acc_iK(Nacc[a][i], al[a] * Kx + be[a], v[i]);
        }
    }
    if (!ok) return;
    const double dKy = ky_of(ll(g.my, cy), g.N2) * g.f2, dKz = cz * g.f3;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const int l = F::out_pos(t, i);
        if (xdrop<L, P, E, F::RL>(i) || !lkept(g.mx, l)) continue;
        const double dKx = kx_four(l, g.N1) * g.f1;
        const double K2 = ksq_f(dKx, dKy, dKz);
        // compact row of l = t + l0: the kept lanes of one register all sit on one side of the
        // dealiased gap (P < L/3), so the row shift is uniform per register
        const int l0 = F::out_pos(0, i);
        const long idx = (long)(l0 - (l0 >= g.mx.h ? g.mx.N - g.mx.nk : 0)) * g.sx + mu + lane_off;
        Dc Nu[3];
        if (MODE == 0) {
#pragma unroll
            for (int a = 0; a < 3; a++) Nu[a] = dsc(dl(Nacc[a][i]), g.norm);
        } else {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                Dc np = dl(Nacc[a][i]), nm = dl(Nacc[3 + a][i]);
                Nu[a] = dsc(dadd(np, nm), 0.5 * g.norm);
            }
        }
        Dc rhs[3];
        if (K2 > 0) {
            Dc kn = dmk(dKx * Nu[0].x + dKy * Nu[1].x + dKz * Nu[2].x, dKx * Nu[0].y + dKy * Nu[1].y + dKz * Nu[2].y);
            const double Kc[3] = {dKx, dKy, dKz};
#pragma unroll
            for (int a = 0; a < 3; a++)
                rhs[a] = dmk(-(Nu[a].x - Kc[a] * kn.x / K2), -(Nu[a].y - Kc[a] * kn.y / K2));
        } else {
#pragma unroll
            for (int a = 0; a < 3; a++) rhs[a] = dmk(-Nu[a].x, -Nu[a].y);
        }
        const double eh = exp(sc.fh[0] * K2), e1 = exp(sc.f1[0] * K2);
#pragma unroll
        for (int a = 0; a < 3; a++)
            rk_update<R>(sc.stage, sc.dt, eh, e1, rhs[a], S + a * g.cS + idx, Vs + a * g.cS + idx, acc + a * g.cS + idx);
        if (MODE == 1) {
            const double ehb = exp(sc.fh[1] * K2), e1b = exp(sc.f1[1] * K2);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                Dc np = dl(Nacc[a][i]), nm = dl(Nacc[3 + a][i]);
                Dc nb = dsc(dmk(np.x - nm.x, np.y - nm.y), 0.5 * g.norm);
                Dc rb = dmk(-nb.x, -nb.y);
                rk_update<R>(sc.stage, sc.dt, ehb, e1b, rb, S + (3 + a) * g.cS + idx, Vs + (3 + a) * g.cS + idx,
                             acc + (3 + a) * g.cS + idx);
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // one mode at a time: bounds the live state loads
    }
}

// ------------------------------------------------------------------------------------------
// energy-transfer epilogue: x-forward of the 9 ET product lines, N_i = sum_j i K_j F(H_j G_i), then
// 2 mult Re(N . U*) binned by shell (Get_shell_index over radii) separately for |K| <= kmax_inside
// (Shell_mult_all) and |K| > kmax_inside (which only Shell_mult_single counts).
// ------------------------------------------------------------------------------------------
template <typename R, int L, int P, int T>
__global__ __launch_bounds__(T * P) void k_fx_fwd_et(FG g, const cplx<R>* __restrict__ D, const cplx<R>* __restrict__ U,
                                                     const cplx<R>* __restrict__ tw, const double* __restrict__ radii,
                                                     int nr, int kin, double* __restrict__ partial)
{
    using F = LineFFT<R, L, P, -1>;
    constexpr int E = F::E;
    __shared__ cplx<R> lds[L * T];
    extern __shared__ double bins[];      // 2 nr doubles: |K| <= kin binned by shell, then |K| > kin by shell
    for (int b = threadIdx.x; b < 2 * nr; b += T * P) bins[b] = 0.0;
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int ntz = (g.nkz + T - 1) / T;
    const int cy = blockIdx.x / ntz;
    const int cz = (blockIdx.x % ntz) * T + c;
    const bool ok = cz < g.nkz;
    const long m = (long)cy * g.KZP + cz;
    const R Ky = (R)(ky_of(ll(g.my, cy), g.N2) * g.f2);
    const R Kz = (R)(cz * g.f3);
    cplx<R> Nacc[3][E];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int i = 0; i < E; i++) Nacc[a][i] = czero<R>();
#pragma unroll 1
    for (int p = 0; p < 9; p++) {
        cplx<R> v[E];
        const cplx<R>* src = D + p * g.cA + m;
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = ok ? src[(long)F::in_pos(t, i) * g.sx] : czero<R>();
        F::template run<T>(v, lds + c, t, tw + opaque_zero());
        const int pi = p / 3, pj = p % 3;        // P(pi,pj) = G_pi H_pj -> N_pi += i K_pj P
        R al[3], be[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            al[a] = (a == pi && pj == 0) ? R(1) : R(0);
            be[a] = (a == pi && pj != 0) ? (pj == 1 ? Ky : Kz) : R(0);
        }
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (xdrop<L, P, E, F::RL>(i)) continue;
            const R Kx = (R)(kx_four(F::out_pos(t, i), g.N1) * g.f1);
#pragma unroll
            for (int a = 0; a < 3; a++) acc_iK(Nacc[a][i], al[a] * Kx + be[a], v[i]);
        }
    }
    __syncthreads();
    if (ok) {
        const double dKy = ky_of(ll(g.my, cy), g.N2) * g.f2, dKz = cz * g.f3;
        const int l2 = ll(g.my, cy);
#pragma unroll
        for (int i = 0; i < E; i++) {
            const int l = F::out_pos(t, i);
            if (xdrop<L, P, E, F::RL>(i) || !lkept(g.mx, l)) continue;
            const int kx = kx_four(l, g.N1);
            const double km = kmag_f(kx * g.f1, dKy, dKz);
            // Multiplicity_factor_FOUR four_inline.h:347-365
            double f = (cz > 0) ? 1.0 : 0.5;
            if (kx == g.N1 / 2 || ky_of(l2, g.N2) == g.N2 / 2) f *= 2;
            const long idx = (long)lq(g.mx, l) * g.sx + m;
            double s = 0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                cplx<R> u = U[a * g.cS + idx];
                s += (double)Nacc[a][i].x * (double)u.x + (double)Nacc[a][i].y * (double)u.y;
            }
            s *= 2 * f * g.norm;
            int b = 0;
            while (b < nr && !(km <= radii[b])) b++;
            if (b >= nr) continue;
            atomicAdd(&bins[(km <= (double)kin) ? b : nr + b], s);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 2 * nr; b += T * P) partial[(long)blockIdx.x * (2 * nr) + b] = bins[b];
}

// compact <-> reference layout
template <typename R>
__global__ void k_gather(FG g, const cplx<R>* __restrict__ ref, cplx<R>* __restrict__ S)
{
    long n = (long)g.mx.nk * g.sx;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int cz = (int)(i % g.KZP);
        long r = i / g.KZP;
        int cy = (int)(r % g.my.nk), cx = (int)(r / g.my.nk);
        if (cz >= g.nkz) { S[i] = czero<R>(); continue; }
        int l1 = ll(g.mx, cx), l2 = ll(g.my, cy);
        S[i] = ref[((long)l1 * g.N2 + l2) * (g.N3 / 2 + 1) + cz];
    }
}
template <typename R>
__global__ void k_scatter(FG g, const cplx<R>* __restrict__ S, cplx<R>* __restrict__ ref)
{
    const int H = g.N3 / 2 + 1;
    long n = (long)g.N1 * g.N2 * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int l3 = (int)(i % H);
        long r = i / H;
        int l2 = (int)(r % g.N2), l1 = (int)(r / g.N2);
        cplx<R> v = czero<R>();
        if (l3 < g.nkz && lkept(g.mx, l1) && lkept(g.my, l2)) v = S[(long)lq(g.mx, l1) * g.sx + (long)lq(g.my, l2) * g.KZP + l3];
        ref[i] = v;
    }
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
struct Fused {};

void release_fused(th_plan* p)
{
    delete p->fused;
    p->fused = nullptr;
}

template <typename R> constexpr int tileT() { return 128 / (int)sizeof(cplx<R>); }

// kernel dispatch on line length
#define TH_FLENS(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#define TH_ZLENS(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

template <typename R, int L, int MODE>
static int launch_xinv(const FG& g, const cplx<R>* S, cplx<R>* A, int ncomp, const cplx<R>* tw, ShellSel sel, hipStream_t st)
{
    constexpr int E = pts_per_thread(L), P = L / E, T = tileT<R>();
    const int ntz = (g.nkz + T - 1) / T;
    hipLaunchKernelGGL((k_fx_inv<R, L, P, T, MODE>), dim3(g.my.nk * ntz, ncomp, g.nbox), dim3(T * P), 0, st, g, S, A, tw, sel);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int L>
static int launch_yinv(const FG& g, const cplx<R>* A, cplx<R>* B, int ncomp, const cplx<R>* tw, hipStream_t st)
{
    constexpr int E = pts_per_thread(L), P = L / E, T = tileT<R>();
    hipLaunchKernelGGL((k_fy_inv<R, L, P, T>), dim3((g.nkz + T - 1) / T, g.N1, ncomp * g.nbox), dim3(T * P), 0, st, g, A, B,
                       tw, ncomp);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int LZ, int MODE>
static int launch_z(const FG& g, const cplx<R>* B, cplx<R>* Cc, const cplx<R>* twL, const cplx<R>* twN, hipStream_t st)
{
    constexpr int E = pts_per_thread(LZ), P = LZ / E;
    constexpr int TZ = (256 / P) < 1 ? 1 : 256 / P;
    const long lines = (long)g.N1 * g.N2;
    hipLaunchKernelGGL((k_fz<R, LZ, P, TZ, MODE>), dim3((unsigned)((lines + TZ - 1) / TZ), g.nbox), dim3(TZ * P), 0, st, g, B, Cc,
                       twL, twN);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int L>
static int launch_yfwd(const FG& g, const cplx<R>* Cc, cplx<R>* D, int ncomp, const cplx<R>* tw, hipStream_t st)
{
    constexpr int E = pts_per_thread(L), P = L / E, T = tileT<R>();
    hipLaunchKernelGGL((k_fy_fwd<R, L, P, T>), dim3((g.nkz + T - 1) / T, g.N1, ncomp * g.nbox), dim3(T * P), 0, st, g, Cc, D,
                       tw, ncomp);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
// The x-forward pass keeps 3 (NS) or 6 (MHD) accumulator lines per thread next to the FFT line, so
// it uses fewer points per thread than the other passes to stay inside the VGPR file (no spills).
constexpr int xfwd_E(int L) { return L < 8 ? L : ((L % 3 == 0) ? (L >= 12 ? 12 : L) : 8); }
template <typename R, int L, int MODE> constexpr int xfwd_T()
{
    constexpr int P = L / xfwd_E(L);
    constexpr int t = (MODE == 0 && sizeof(R) == 4) ? 16 : 8;
    return (t * P > 512) ? (512 / P < 1 ? 1 : 512 / P) : t;
}

template <typename R, int L, int MODE>
static int launch_xfwd(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw,
                       const StageCoef* sc, hipStream_t st)
{
    constexpr int E = xfwd_E(L), P = L / E, T = xfwd_T<R, L, MODE>();
    const int ntz = (g.nkz + T - 1) / T;
    hipLaunchKernelGGL((k_fx_fwd<R, L, P, T, MODE>), dim3(g.my.nk * ntz, g.nbox), dim3(T * P), 0, st, g, D, S, Vs, acc, tw,
                       sc);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int L>
static int launch_xfwd_et(const FG& g, const cplx<R>* D, const cplx<R>* U, const cplx<R>* tw, const double* radii, int nr,
                          int kin, double* partial, int* nblk, hipStream_t st)
{
    constexpr int E = xfwd_E(L), P = L / E, T = xfwd_T<R, L, 0>();
    const int ntz = (g.nkz + T - 1) / T;
    *nblk = g.my.nk * ntz;
    if (!partial) return OK;
    hipLaunchKernelGGL((k_fx_fwd_et<R, L, P, T>), dim3(*nblk), dim3(T * P), 2 * nr * sizeof(double), st, g, D, U, tw,
                       radii, nr, kin, partial);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R, int MODE>
static int d_xinv(const FG& g, const cplx<R>* S, cplx<R>* A, int nc, const cplx<R>* tw, ShellSel sel, hipStream_t st)
{
#define C_(n) case n: return launch_xinv<R, n, MODE>(g, S, A, nc, tw, sel, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
static int d_yinv(const FG& g, const cplx<R>* A, cplx<R>* B, int nc, const cplx<R>* tw, hipStream_t st)
{
#define C_(n) case n: return launch_yinv<R, n>(g, A, B, nc, tw, st);
    switch (g.N2) { TH_FLENS(C_) default: set_error("fused: N2 %d unsupported", g.N2); return E_UNSUPPORTED; }
#undef C_
}
template <typename R, int MODE>
static int d_z(const FG& g, const cplx<R>* B, cplx<R>* Cc, const cplx<R>* twL, const cplx<R>* twN, hipStream_t st)
{
#define C_(n) case n: return launch_z<R, n, MODE>(g, B, Cc, twL, twN, st);
    switch (g.LZ) { TH_ZLENS(C_) default: set_error("fused: N3 %d unsupported", g.N3); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
static int d_yfwd(const FG& g, const cplx<R>* Cc, cplx<R>* D, int nc, const cplx<R>* tw, hipStream_t st)
{
#define C_(n) case n: return launch_yfwd<R, n>(g, Cc, D, nc, tw, st);
    switch (g.N2) { TH_FLENS(C_) default: set_error("fused: N2 %d unsupported", g.N2); return E_UNSUPPORTED; }
#undef C_
}
template <typename R, int MODE>
static int d_xfwd(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw,
                  const StageCoef* sc, hipStream_t st)
{
#define C_(n) case n: return launch_xfwd<R, n, MODE>(g, D, S, Vs, acc, tw, sc, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
static int d_xfwd_et(const FG& g, const cplx<R>* D, const cplx<R>* U, const cplx<R>* tw, const double* radii, int nr,
                     int kin, double* partial, int* nblk, hipStream_t st)
{
#define C_(n) case n: return launch_xfwd_et<R, n>(g, D, U, tw, radii, nr, kin, partial, nblk, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}

static bool fused_len_ok(int n)
{
    const int ok[] = {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};
    for (int v : ok)
        if (v == n) return true;
    return false;
}

FG make_fg(const Geo& geo, int prec)
{
    FG g;
    memset(&g, 0, sizeof(g));
    g.N1 = geo.N1;
    g.N2 = geo.N2;
    g.N3 = geo.N3;
    g.LZ = geo.N3 / 2;
    if (geo.dealias) {
        int hx = geo.N1 / 3, nkx = 2 * (geo.N1 / 3) - 1;
        g.mx = LMap{geo.N1, nkx, hx};
        int ylo = geo.N2 / 3, yhi = geo.N2 - (2 * geo.N2 / 3 + 2);
        g.my = LMap{geo.N2, ylo + yhi, ylo};
        g.nkz = geo.N3 / 3;
    } else {
        g.mx = LMap{geo.N1, geo.N1, geo.N1};
        g.my = LMap{geo.N2, geo.N2, geo.N2};
        g.nkz = geo.N3 / 2 + 1;
    }
    int T = prec ? 16 : 8;
    g.KZP = (g.nkz + T - 1) / T * T;
    g.sx = (long)g.my.nk * g.KZP;
    g.sB = (long)g.N2 * g.KZP;
    g.cS = (long)g.mx.nk * g.sx;
    g.cA = (long)g.N1 * g.sx;
    g.cB = (long)g.N1 * g.sB;
    g.f1 = geo.kf[0];
    g.f2 = geo.kf[1];
    g.f3 = geo.kf[2];
    g.norm = 1.0 / ((double)geo.N1 * (double)geo.N2 * (double)geo.N3);
    g.nbox = 1;
    return g;
}

}  // namespace th

using namespace th;

struct th_state {
    th_plan* plan = nullptr;
    int physics = TH_NS;
    int ncomp = 3;
    int nin = 3, nout = 6;       // pipeline component counts
    int nbox = 1;                // boxes of the plan's size held box-major in every buffer
    FG fg;
    size_t esz = 16;
    void* S = nullptr;           // state (compact), ncomp comps
    void* Vs = nullptr;          // stage field
    void* acc = nullptr;         // RK4 accumulator
    void* X1 = nullptr;          // A / D layout, max(nin, nout) comps
    void* X2 = nullptr;          // B layout, nin comps (ET: 6)
    void* X3 = nullptr;          // C layout, nout comps (ET: 9)
    double* partial = nullptr;
    size_t partial_cap = 0;
    double* radii = nullptr;
    int radii_cap = 0;
    StageCoef* dcoef = nullptr;  // device [4 stages][nbox]; the captured step reads it
    StageCoef* hcoef = nullptr;  // pinned staging for dcoef
    std::vector<StageCoef> cur;  // what dcoef holds once the last upload has run
    hipEvent_t cev = nullptr;    // recorded after the last upload from hcoef
    bool cev_live = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    ~th_state()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* q : {S, Vs, acc, X1, X2, X3}) if (q) (void)hipFree(q);
        if (partial) (void)hipFree(partial);
        if (radii) (void)hipFree(radii);
        if (dcoef) (void)hipFree(dcoef);
        if (hcoef) (void)hipHostFree(hcoef);
        if (cev) (void)hipEventDestroy(cev);
    }
};

#define TRY(x)             \
    do {                   \
        int r_ = (x);      \
        if (r_) return r_; \
    } while (0)

// Fused state for nbox boxes of the plan's size. Single states (nbox == 1) also keep room for the
// energy-transfer pipeline (6 B-layout and 9 C-layout comps); batched states hold only what the
// RK step needs.
static int state_create(th_plan* p, th_state** out, int physics, int nbox)
{
    if (!p || !out || nbox < 1) { set_error("th_state_create: bad args"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    if (p->g.basis != TH_FOUR) { set_error("th_state_create: the fused path covers the FOUR basis; use th_step_plain for SCFT"); return E_UNSUPPORTED; }
    if (physics != TH_NS && physics != TH_MHD) { set_error("th_state_create: physics %d not fused (use th_step_plain)", physics); return E_UNSUPPORTED; }
    if (!p->g.dealias) { set_error("th_state_create: the fused path stores 2/3-dealiased modes only; create the plan with dealias = 1"); return E_UNSUPPORTED; }
    if (!fused_len_ok(p->g.N1) || !fused_len_ok(p->g.N2) || !fused_len_ok(p->g.N3)) {
        set_error("th_state_create: fused path needs N in {16,24,32,48,64,96,128,192,256,384,512}");
        return E_UNSUPPORTED;
    }
    th_state* s = new (std::nothrow) th_state();
    if (!s) return E_NOMEM;
    s->plan = p;
    s->physics = physics;
    s->nbox = nbox;
    s->ncomp = physics == TH_NS ? 3 : 6;
    s->nin = physics == TH_NS ? 3 : 6;
    s->nout = physics == TH_NS ? 6 : 9;
    s->fg = make_fg(p->g, p->prec);
    s->esz = p->prec ? 8 : 16;
    FG& g = s->fg;
    const int nA = nbox == 1 ? 9 : s->nout, nBin = nbox == 1 ? 6 : s->nin, nC = nbox == 1 ? 9 : s->nout;
    g.nbox = nbox;
    g.bS = s->ncomp * g.cS;
    g.bA = nA * g.cA;
    g.bB = nBin * g.cB;
    g.bC = nC * g.cB;
    auto alloc = [&](void** q, size_t elems) -> int {
        if (hipMalloc(q, elems * s->esz) != hipSuccess) { set_error("th_state_create: out of device memory"); return E_NOMEM; }
        if (hipMemset(*q, 0, elems * s->esz) != hipSuccess) return E_HIP;
        return OK;
    };
    int r = OK;
    const size_t nb = (size_t)nbox;
    if ((r = alloc(&s->S, nb * g.bS)) || (r = alloc(&s->Vs, nb * g.bS)) || (r = alloc(&s->acc, nb * g.bS)) ||
        (r = alloc(&s->X1, nb * g.bA)) || (r = alloc(&s->X2, nb * g.bB)) || (r = alloc(&s->X3, nb * g.bC))) {
        delete s;
        return r;
    }
    const size_t cb = 4 * nb * sizeof(StageCoef);
    if (hipMalloc((void**)&s->dcoef, cb) != hipSuccess || hipHostMalloc((void**)&s->hcoef, cb, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->cev, hipEventDisableTiming) != hipSuccess) {
        set_error("th_state_create: coefficient buffers");
        delete s;
        return E_NOMEM;
    }
    *out = s;
    return OK;
}

extern "C" int th_state_create(th_plan* p, th_state** out, int physics) { return state_create(p, out, physics, 1); }

extern "C" int th_state_destroy(th_state* s)
{
    if (!s) return OK;
    (void)hipSetDevice(s->plan->ctx->device);
    (void)hipStreamSynchronize(s->plan->ctx->stream);
    delete s;
    return OK;
}

static dim3 g1d(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return dim3((unsigned)(b ? b : 1));
}

static int state_upload(th_state* s, int box, int comp, const void* src, int src_on_host, hipStream_t st)
{
    if (!s || !src || comp < 0 || comp >= s->ncomp || box < 0 || box >= s->nbox) { set_error("th_state_upload: bad args"); return E_ARG; }
    th_plan* p = s->plan;
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    const void* dsrc = src;
    void* tmp = nullptr;
    size_t bytes = p->g.ncplx() * s->esz;
    if (src_on_host) {
        TH_CHECK_HIP(hipMalloc(&tmp, bytes));
        TH_CHECK_HIP(hipMemcpyAsync(tmp, src, bytes, hipMemcpyHostToDevice, st));
        dsrc = tmp;
    }
    const FG& g = s->fg;
    const size_t n = (size_t)g.mx.nk * g.sx;
    const long off = box * g.bS + comp * g.cS;
    if (p->prec)
        hipLaunchKernelGGL((k_gather<float>), g1d(n), dim3(256), 0, st, g, (const cplx<float>*)dsrc, (cplx<float>*)s->S + off);
    else
        hipLaunchKernelGGL((k_gather<double>), g1d(n), dim3(256), 0, st, g, (const cplx<double>*)dsrc, (cplx<double>*)s->S + off);
    hipError_t e = hipGetLastError();
    if (tmp) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(tmp);
    }
    if (e != hipSuccess) { set_error("k_gather: %s", hipGetErrorString(e)); return E_HIP; }
    return OK;
}

static int state_download(th_state* s, int box, int comp, void* dst, int dst_on_host, hipStream_t st)
{
    if (!s || !dst || comp < 0 || comp >= s->ncomp || box < 0 || box >= s->nbox) { set_error("th_state_download: bad args"); return E_ARG; }
    th_plan* p = s->plan;
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    size_t bytes = p->g.ncplx() * s->esz;
    void* ddst = dst;
    void* tmp = nullptr;
    if (dst_on_host) {
        TH_CHECK_HIP(hipMalloc(&tmp, bytes));
        ddst = tmp;
    }
    const FG& g = s->fg;
    const size_t n = p->g.ncplx();
    const long off = box * g.bS + comp * g.cS;
    if (p->prec)
        hipLaunchKernelGGL((k_scatter<float>), g1d(n), dim3(256), 0, st, g, (const cplx<float>*)s->S + off, (cplx<float>*)ddst);
    else
        hipLaunchKernelGGL((k_scatter<double>), g1d(n), dim3(256), 0, st, g, (const cplx<double>*)s->S + off, (cplx<double>*)ddst);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && tmp) {
        e = hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) { set_error("th_state_download: %s", hipGetErrorString(e)); return E_HIP; }
    return OK;
}

extern "C" int th_state_upload(th_state* s, int comp, const void* src, int src_on_host)
{
    if (!s) { set_error("th_state_upload: null"); return E_ARG; }
    return state_upload(s, 0, comp, src, src_on_host, s->plan->ctx->stream);
}

extern "C" int th_state_download(th_state* s, int comp, void* dst, int dst_on_host)
{
    if (!s) { set_error("th_state_download: null"); return E_ARG; }
    return state_download(s, 0, comp, dst, dst_on_host, s->plan->ctx->stream);
}

static StageCoef stage_coef(int stage, double dt, double nu, double eta)
{
    StageCoef c;
    memset(&c, 0, sizeof(c));
    c.stage = stage;
    c.dt = dt;
    c.fh[0] = (-nu * dt) * 0.5;
    c.f1[0] = (-nu * dt) * 1.0;
    c.fh[1] = (-eta * dt) * 0.5;
    c.f1[1] = (-eta * dt) * 1.0;
    return c;
}

static bool same_coef(const StageCoef& a, const StageCoef& b)
{
    return a.stage == b.stage && a.dt == b.dt && a.fh[0] == b.fh[0] && a.fh[1] == b.fh[1] && a.f1[0] == b.f1[0] &&
           a.f1[1] == b.f1[1];
}

// Per-box stage coefficients (dt, -nu dt / 2, -nu dt, ...) live in device memory, laid out
// [stage][box], so a captured step serves any dt and viscosity. A new set is staged through
// pinned memory and copied on stream st ahead of the step that uses it.
static int set_coefs(th_state* s, const double* dt, const double* nu, const double* eta, hipStream_t st)
{
    const int nb = s->nbox;
    std::vector<StageCoef> c((size_t)4 * nb);
    for (int k = 0; k < 4; k++)
        for (int b = 0; b < nb; b++) c[(size_t)k * nb + b] = stage_coef(k + 1, dt[b], nu[b], eta ? eta[b] : 0.0);
    if (s->cur.size() == c.size()) {
        bool same = true;
        for (size_t i = 0; i < c.size() && same; i++) same = same_coef(c[i], s->cur[i]);
        if (same) return OK;
    }
    if (s->cev_live) TH_CHECK_HIP(hipEventSynchronize(s->cev));   // the previous copy has read hcoef
    memcpy(s->hcoef, c.data(), c.size() * sizeof(StageCoef));
    TH_CHECK_HIP(hipMemcpyAsync(s->dcoef, s->hcoef, c.size() * sizeof(StageCoef), hipMemcpyHostToDevice, st));
    TH_CHECK_HIP(hipEventRecord(s->cev, st));
    s->cev_live = true;
    s->cur.swap(c);
    return OK;
}

// One RK4 step of every box in s, as 4 stages x 5 pass launches on stream st.
template <typename R> static int enqueue_step(th_state* s, hipStream_t st)
{
    th_plan* p = s->plan;
    const FG& g = s->fg;
    const cplx<R>* twx = p->twd<R>(g.N1);
    const cplx<R>* twy = p->twd<R>(g.N2);
    const cplx<R>* twz = p->twd<R>(g.LZ);
    const cplx<R>* twn = p->twd<R>(g.N3);
    cplx<R>* S = (cplx<R>*)s->S;
    cplx<R>* Vs = (cplx<R>*)s->Vs;
    cplx<R>* acc = (cplx<R>*)s->acc;
    cplx<R>* X1 = (cplx<R>*)s->X1;
    cplx<R>* X2 = (cplx<R>*)s->X2;
    cplx<R>* X3 = (cplx<R>*)s->X3;
    ShellSel sel{0, 0};
    for (int stage = 1; stage <= 4; stage++) {
        const StageCoef* sc = s->dcoef + (size_t)(stage - 1) * s->nbox;
        const cplx<R>* src = stage == 1 ? S : Vs;
        if (s->physics == TH_NS) {
            TRY((d_xinv<R, 0>(g, src, X1, 3, twx, sel, st)));
            TRY(d_yinv<R>(g, X1, X2, 3, twy, st));
            TRY((d_z<R, 0>(g, X2, X3, twz, twn, st)));
            TRY(d_yfwd<R>(g, X3, X1, 6, twy, st));
            TRY((d_xfwd<R, 0>(g, X1, S, Vs, acc, twx, sc, st)));
        } else {
            TRY((d_xinv<R, 1>(g, src, X1, 6, twx, sel, st)));
            TRY(d_yinv<R>(g, X1, X2, 6, twy, st));
            TRY((d_z<R, 1>(g, X2, X3, twz, twn, st)));
            TRY(d_yfwd<R>(g, X3, X1, 9, twy, st));
            TRY((d_xfwd<R, 1>(g, X1, S, Vs, acc, twx, sc, st)));
        }
    }
    return OK;
}

static int ensure_twiddles(th_plan* p, const FG& g, hipStream_t st)
{
    for (int L : {g.N1, g.N2, g.LZ, g.N3})
        if (!p->tw.get(L, p->prec, st)) { set_error("twiddle table allocation failed"); return E_NOMEM; }
    return OK;
}

extern "C" int th_step(th_plan* p, th_state* s, double dt, const th_phys* ph, int nsteps)
{
    if (!p || !s || !ph || s->plan != p || nsteps < 0) { set_error("th_step: bad args"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    hipStream_t st = p->ctx->stream;
    // twiddle tables must exist before graph capture (their creation allocates)
    TRY(ensure_twiddles(p, s->fg, st));
    TRY(set_coefs(s, &dt, &ph->nu, &ph->eta, st));
    if (!s->gexec) {
        TH_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int r = p->prec ? enqueue_step<float>(s, st) : enqueue_step<double>(s, st);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
        if (e != hipSuccess) { set_error("hipStreamEndCapture: %s", hipGetErrorString(e)); return E_HIP; }
        s->graph = graph;
        TH_CHECK_HIP(hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < nsteps; i++) TH_CHECK_HIP(hipGraphLaunch(s->gexec, st));
    return OK;
}

static __global__ void th_finalize_bins(const double* __restrict__ partial, int nblk, int stride, int nval,
                                        double* __restrict__ out)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nval) return;
    double s = 0.0, c = 0.0;
    for (int k = 0; k < nblk; k++) {
        double v = partial[(long)k * stride + b];
        double t = s + v;
        c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    out[b] = s + c;
}

// Energy transfer from the state's velocity: one giver-masked fused pipeline per shell n gives
// T(n, m) (Shell_mult_all, |K| <= kmax_inside) and the same sums over |K| > kmax_inside ("corner",
// which Shell_mult_single counts); then Pi(R_i) = sum_{n<=i} sum_{m>i} [T(n,m) - corner(n,m)]
// (sphere i = union of shells 1..i and N[G] is linear in G).
template <typename R> static int flux_impl(th_state* s, const double* radii, int nsh, double* Pi, double* T)
{
    th_plan* p = s->plan;
    hipStream_t st = p->ctx->stream;
    const FG& g = s->fg;
    const cplx<R>* twx = p->twd<R>(g.N1);
    const cplx<R>* twy = p->twd<R>(g.N2);
    const cplx<R>* twz = p->twd<R>(g.LZ);
    const cplx<R>* twn = p->twd<R>(g.N3);
    cplx<R>* S = (cplx<R>*)s->S;
    cplx<R>* X1 = (cplx<R>*)s->X1;
    cplx<R>* X2 = (cplx<R>*)s->X2;
    cplx<R>* X3 = (cplx<R>*)s->X3;
    const int nr = nsh + 1;
    if (s->radii_cap < nr) {
        if (s->radii) (void)hipFree(s->radii);
        s->radii = nullptr;
        TH_CHECK_HIP(hipMalloc(&s->radii, nr * sizeof(double)));
        s->radii_cap = nr;
    }
    TH_CHECK_HIP(hipMemcpyAsync(s->radii, radii, nr * sizeof(double), hipMemcpyHostToDevice, st));
    int nblk = 0;
    TRY(d_xfwd_et<R>(g, X1, S, twx, s->radii, nr, 0, nullptr, &nblk, st));
    size_t need = (size_t)nblk * (2 * nr) + (2 * nr);
    if (need > s->partial_cap) {
        if (s->partial) (void)hipFree(s->partial);
        s->partial = nullptr;
        TH_CHECK_HIP(hipMalloc(&s->partial, need * sizeof(double)));
        s->partial_cap = need;
    }
    const int kin = max_radius_inside(p->g);
    ShellSel all{-1.0, 1e300};
    // helper H = U: inverse x, y once, kept in X2 comps 0..2 (B layout)
    TRY((d_xinv<R, 0>(g, S, X1, 3, twx, all, st)));
    TRY(d_yinv<R>(g, X1, X2, 3, twy, st));
    std::vector<double> hostbuf(2 * nr);
    std::vector<double> Tm((size_t)nr * nr, 0.0), corner((size_t)nr * nr, 0.0);
    double* outd = s->partial + (size_t)nblk * (2 * nr);
    for (int n = 1; n < nsh; n++) {
        ShellSel sel{radii[n - 1], radii[n]};
        TRY((d_xinv<R, 2>(g, S, X1, 3, twx, sel, st)));
        TRY(d_yinv<R>(g, X1, X2 + 3 * g.cB, 3, twy, st));
        TRY((d_z<R, 2>(g, X2, X3, twz, twn, st)));
        TRY(d_yfwd<R>(g, X3, X1, 9, twy, st));
        TRY(d_xfwd_et<R>(g, X1, S, twx, s->radii, nr, kin, s->partial, &nblk, st));
        hipLaunchKernelGGL(th_finalize_bins, dim3((2 * nr + 255) / 256), dim3(256), 0, st, (const double*)s->partial,
                           nblk, 2 * nr, 2 * nr, outd);
        TH_CHECK_HIP(hipGetLastError());
        TH_CHECK_HIP(hipMemcpyAsync(hostbuf.data(), outd, 2 * nr * sizeof(double), hipMemcpyDeviceToHost, st));
        TH_CHECK_HIP(hipStreamSynchronize(st));
        for (int m = 0; m < nr; m++) {
            Tm[(size_t)n * nr + m] = -hostbuf[m];
            corner[(size_t)n * nr + m] = hostbuf[nr + m];
        }
    }
    if (T)
        for (size_t i = 0; i < (size_t)nr * nr; i++) T[i] = Tm[i];
    if (Pi) {
        Pi[0] = 0;
        for (int i = 1; i <= nsh; i++) {
            double acc_ = 0;
            for (int n = 1; n <= i && n < nsh; n++)
                for (int m = i + 1; m < nr; m++) acc_ += Tm[(size_t)n * nr + m] - corner[(size_t)n * nr + m];
            Pi[i] = (i >= nsh) ? 0.0 : acc_;
        }
    }
    return OK;
}

extern "C" int th_flux_shell_tr(th_plan* p, th_state* s, const double* radii, int nsh, double* Pi, double* T)
{
    if (!p || !s || s->plan != p || !radii || nsh < 2) { set_error("th_flux_shell_tr: bad args"); return E_ARG; }
    if (s->physics != TH_NS) { set_error("th_flux_shell_tr: NS state required"); return E_UNSUPPORTED; }
    if (radii[nsh] < 1e3) { set_error("th_flux_shell_tr: radii[nsh] must be the infinite radius"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? flux_impl<float>(s, radii, nsh, Pi, T) : flux_impl<double>(s, radii, nsh, Pi, T);
}

// ------------------------------------------------------------------------------------------
// Per-pass timing of one fused RK stage with HIP events on the plan's stream, and each pass's
// algorithmic HBM bytes (compulsory reads + writes of kept data; padding excluded).
// ------------------------------------------------------------------------------------------
template <typename R> static int profile_impl(th_state* s, double dt, double nu, double eta, int stage, int nrep,
                                              double* ms, double* bytes)
{
    th_plan* p = s->plan;
    hipStream_t st = p->ctx->stream;
    const FG& g = s->fg;
    const cplx<R>* twx = p->twd<R>(g.N1);
    const cplx<R>* twy = p->twd<R>(g.N2);
    const cplx<R>* twz = p->twd<R>(g.LZ);
    const cplx<R>* twn = p->twd<R>(g.N3);
    cplx<R>* S = (cplx<R>*)s->S;
    cplx<R>* Vs = (cplx<R>*)s->Vs;
    cplx<R>* acc = (cplx<R>*)s->acc;
    cplx<R>* X1 = (cplx<R>*)s->X1;
    cplx<R>* X2 = (cplx<R>*)s->X2;
    cplx<R>* X3 = (cplx<R>*)s->X3;
    TRY(set_coefs(s, &dt, &nu, &eta, st));
    const StageCoef* sc = s->dcoef + (stage - 1);
    ShellSel sel{0, 0};
    const bool mhd = s->physics == TH_MHD;
    const int nin = mhd ? 6 : 3, nout = mhd ? 9 : 6, nst = mhd ? 6 : 3;
    const cplx<R>* src = stage == 1 ? S : Vs;
    hipEvent_t ev[6];
    for (int i = 0; i < 6; i++) TH_CHECK_HIP(hipEventCreate(&ev[i]));
    double tot[5] = {0, 0, 0, 0, 0};
    for (int rep = 0; rep < nrep; rep++) {
        TH_CHECK_HIP(hipEventRecord(ev[0], st));
        if (mhd) TRY((d_xinv<R, 1>(g, src, X1, 6, twx, sel, st)));
        else TRY((d_xinv<R, 0>(g, src, X1, 3, twx, sel, st)));
        TH_CHECK_HIP(hipEventRecord(ev[1], st));
        TRY(d_yinv<R>(g, X1, X2, nin, twy, st));
        TH_CHECK_HIP(hipEventRecord(ev[2], st));
        if (mhd) TRY((d_z<R, 1>(g, X2, X3, twz, twn, st)));
        else TRY((d_z<R, 0>(g, X2, X3, twz, twn, st)));
        TH_CHECK_HIP(hipEventRecord(ev[3], st));
        TRY(d_yfwd<R>(g, X3, X1, nout, twy, st));
        TH_CHECK_HIP(hipEventRecord(ev[4], st));
        if (mhd) TRY((d_xfwd<R, 1>(g, X1, S, Vs, acc, twx, sc, st)));
        else TRY((d_xfwd<R, 0>(g, X1, S, Vs, acc, twx, sc, st)));
        TH_CHECK_HIP(hipEventRecord(ev[5], st));
        TH_CHECK_HIP(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; k++) {
            float t = 0;
            TH_CHECK_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            tot[k] += t;
        }
    }
    for (int i = 0; i < 6; i++) (void)hipEventDestroy(ev[i]);
    const double e = (double)sizeof(cplx<R>);
    const double cS = (double)g.mx.nk * g.my.nk * g.nkz, cA = (double)g.N1 * g.my.nk * g.nkz,
                 cB = (double)g.N1 * g.N2 * g.nkz;
    const int state_rw = (stage == 1 || stage == 4) ? 3 : 4;   // V0 (+acc) read, Vs / acc / S written
    const double b[5] = {e * (nst * cS + nin * cA), e * (nin * cA + nin * cB), e * (nin * cB + nout * cB),
                         e * (nout * cB + nout * cA), e * (nout * cA + state_rw * nst * cS)};
    for (int k = 0; k < 5; k++) {
        ms[k] = tot[k] / nrep;
        bytes[k] = b[k];
    }
    return OK;
}

extern "C" int th_profile_passes(th_plan* p, th_state* s, double dt, const th_phys* ph, int stage, int nrep, double* ms,
                                 double* bytes)
{
    if (!p || !s || s->plan != p || !ph || stage < 1 || stage > 4 || nrep < 1 || !ms || !bytes) {
        set_error("th_profile_passes: bad args");
        return E_ARG;
    }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? profile_impl<float>(s, dt, ph->nu, ph->eta, stage, nrep, ms, bytes)
                   : profile_impl<double>(s, dt, ph->nu, ph->eta, stage, nrep, ms, bytes);
}

// ------------------------------------------------------------------------------------------
// Ensemble of independent NS boxes of uneven sizes. Boxes of one size form a class held in one
// batched state (box-major buffers, one launch per pass covers every box of the class); the
// classes run as concurrent branches of one captured RK4 step. Per-box dt and viscosity come from
// the classes' device coefficient tables, so the captured step is reused for any dt / nu.
// ------------------------------------------------------------------------------------------
struct th_ensemble {
    th_ctx* ctx = nullptr;
    int nbox = 0, prec = 0;
    struct Cls {
        th_plan* plan = nullptr;
        th_state* st = nullptr;
        std::vector<int> boxes;       // ensemble box ids, in slot order
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
    };
    std::vector<Cls> cls;
    std::vector<int> cls_of, slot_of;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<double> dtc, nuc;      // per-class scratch for coefficient upload
    ~th_ensemble()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (Cls& c : cls) {
            if (c.st) (void)th_state_destroy(c.st);
            if (c.plan) (void)th_plan_destroy(c.plan);
            if (c.done) (void)hipEventDestroy(c.done);
            if (c.stream) (void)hipStreamDestroy(c.stream);
        }
        if (fork) (void)hipEventDestroy(fork);
    }
};

extern "C" int th_ensemble_create(th_ctx* ctx, th_ensemble** out, int nbox, const int* N, int precision)
{
    if (!ctx || !out || nbox < 1 || !N || (precision != TH_FP64 && precision != TH_FP32_MIXED)) {
        set_error("th_ensemble_create: bad args");
        return E_ARG;
    }
    TH_CHECK_HIP(hipSetDevice(ctx->device));
    th_ensemble* e = new (std::nothrow) th_ensemble();
    if (!e) return E_NOMEM;
    e->ctx = ctx;
    e->nbox = nbox;
    e->prec = precision;
    e->cls_of.assign(nbox, -1);
    e->slot_of.assign(nbox, -1);
    std::vector<std::array<int, 3>> shapes;
    for (int b = 0; b < nbox; b++) {
        std::array<int, 3> n{N[3 * b], N[3 * b + 1], N[3 * b + 2]};
        int c = -1;
        for (size_t i = 0; i < shapes.size(); i++)
            if (shapes[i] == n) c = (int)i;
        if (c < 0) {
            c = (int)shapes.size();
            shapes.push_back(n);
            e->cls.emplace_back();
        }
        e->cls_of[b] = c;
        e->slot_of[b] = (int)e->cls[c].boxes.size();
        e->cls[c].boxes.push_back(b);
    }
    const double kf[3] = {1.0, 1.0, 1.0};
    for (size_t c = 0; c < e->cls.size(); c++) {
        th_ensemble::Cls& k = e->cls[c];
        int r = th_plan_create(ctx, &k.plan, TH_FOUR, shapes[c][0], shapes[c][1], shapes[c][2], kf, precision, 1);
        if (!r) r = state_create(k.plan, &k.st, TH_NS, (int)k.boxes.size());
        if (!r && (hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking) != hipSuccess ||
                   hipEventCreateWithFlags(&k.done, hipEventDisableTiming) != hipSuccess)) {
            set_error("th_ensemble_create: stream / event creation failed");
            r = E_HIP;
        }
        if (r) {
            delete e;
            return r;
        }
    }
    if (hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
        delete e;
        set_error("th_ensemble_create: event creation failed");
        return E_HIP;
    }
    *out = e;
    return OK;
}

extern "C" int th_ensemble_destroy(th_ensemble* e)
{
    if (!e) return OK;
    (void)hipSetDevice(e->ctx->device);
    (void)hipStreamSynchronize(e->ctx->stream);
    delete e;
    return OK;
}

extern "C" int th_ensemble_plan(th_ensemble* e, int box, th_plan** plan)
{
    if (!e || !plan || box < 0 || box >= e->nbox) { set_error("th_ensemble_plan: bad args"); return E_ARG; }
    *plan = e->cls[e->cls_of[box]].plan;
    return OK;
}

extern "C" int th_ensemble_upload(th_ensemble* e, int box, int comp, const void* src, int src_on_host)
{
    if (!e || box < 0 || box >= e->nbox) { set_error("th_ensemble_upload: bad args"); return E_ARG; }
    return state_upload(e->cls[e->cls_of[box]].st, e->slot_of[box], comp, src, src_on_host, e->ctx->stream);
}

extern "C" int th_ensemble_download(th_ensemble* e, int box, int comp, void* dst, int dst_on_host)
{
    if (!e || box < 0 || box >= e->nbox) { set_error("th_ensemble_download: bad args"); return E_ARG; }
    return state_download(e->cls[e->cls_of[box]].st, e->slot_of[box], comp, dst, dst_on_host, e->ctx->stream);
}

extern "C" int th_ensemble_step(th_ensemble* e, const double* dt, const double* nu, int nsteps)
{
    if (!e || !dt || !nu || nsteps < 0) { set_error("th_ensemble_step: bad args"); return E_ARG; }
    for (int b = 0; b < e->nbox; b++)
        if (!(dt[b] > 0.0) || !(nu[b] >= 0.0)) { set_error("th_ensemble_step: box %d has dt %g nu %g", b, dt[b], nu[b]); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(e->ctx->device));
    hipStream_t st = e->ctx->stream;
    for (th_ensemble::Cls& k : e->cls) {
        TRY(ensure_twiddles(k.plan, k.st->fg, st));
        e->dtc.clear();
        e->nuc.clear();
        for (int b : k.boxes) {
            e->dtc.push_back(dt[b]);
            e->nuc.push_back(nu[b]);
        }
        TRY(set_coefs(k.st, e->dtc.data(), e->nuc.data(), nullptr, st));
    }
    if (!e->gexec) {
        TH_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int r = OK;
        hipError_t he = hipEventRecord(e->fork, st);
        for (size_t c = 0; c < e->cls.size() && !r && he == hipSuccess; c++) {
            th_ensemble::Cls& k = e->cls[c];
            he = hipStreamWaitEvent(k.stream, e->fork, 0);
            if (he != hipSuccess) break;
            r = e->prec ? enqueue_step<float>(k.st, k.stream) : enqueue_step<double>(k.st, k.stream);
            if (r) break;
            he = hipEventRecord(k.done, k.stream);
            if (he == hipSuccess) he = hipStreamWaitEvent(st, k.done, 0);
        }
        hipGraph_t graph = nullptr;
        hipError_t ee = hipStreamEndCapture(st, &graph);
        if (r || he != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (r) return r;
            set_error("th_ensemble_step capture: %s", hipGetErrorString(he != hipSuccess ? he : ee));
            return E_HIP;
        }
        e->graph = graph;
        TH_CHECK_HIP(hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < nsteps; i++) TH_CHECK_HIP(hipGraphLaunch(e->gexec, st));
    return OK;
}
