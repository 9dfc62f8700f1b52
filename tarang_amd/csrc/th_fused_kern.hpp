// th_fused_kern.hpp -- kernels of the dealias-pruned, fused pseudo-spectral RK stage (FOUR basis).
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
// (four_ET.cc) in the pass-5 epilogue instead of writing N.  Chained NS and MHD steps fold pass 1 of the
// next stage into pass 5 (k_fx_fwd CHAIN).  Each pipeline buffer is written once and read once by the next
// pass, so outputs go out with non-temporal stores and the y and z passes load non-temporally.

#pragma once
#include <type_traits>
#include "th_fused.hpp"

namespace th {

template <typename R> __device__ __forceinline__ cplx<R> ld(const cplx<R>* p) { return *p; }

// ------------------------------------------------------------------------------------------
// pass 1: x-inverse.  MODE 0: copy comps; 1: Elsasser Z+ = U + B (comps 0-2), Z- = U - B (3-5);
// 2: giver G = U masked to the shell sel.
// ------------------------------------------------------------------------------------------
// Twiddle tables are staged in LDS once per workgroup: read from global memory in every FFT stage
// they would cost as many vector-memory instructions as the line data itself.
template <typename R, int N> __device__ __forceinline__ void stage_tw(cplx<R>* dst, const cplx<R>* __restrict__ src)
{
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
}

template <typename R, int L, int P, int T, int MODE>
__global__ __launch_bounds__(T * P) void k_fx_inv(FG g, const cplx<R>* __restrict__ S, cplx<R>* __restrict__ A,
                                                  const cplx<R>* __restrict__ tw, ShellSel sel)
{
    using F = LineFFT<R, L, P, +1>;
    __shared__ cplx<R> lds[L * T];
    __shared__ cplx<R> twl[L];
    stage_tw<R, L>(twl, tw);
    __syncthreads();
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
    if constexpr (MODE == 1) {
        // component cc of U and B read once: Z+ = U + B into output cc, then Z- = U - B into output 3 + cc
        const int cc = oc;
        cplx<R> u[F::E], b[F::E];
#pragma unroll
        for (int i = 0; i < F::E; i++) {
            const int l = F::in_pos(t, i);
            u[i] = b[i] = czero<R>();
            if (ok && lkept(g.mx, l)) {
                const long off = (long)lq(g.mx, l) * g.sx + m;
                u[i] = S[cc * g.cS + off];
                b[i] = S[(3 + cc) * g.cS + off];
            }
            v[i] = u[i] + b[i];
        }
        F::template run<T, false, NoPad, false, 0>(v, lds + c, t, twl);   // lds unused before
        if (ok) {
            cplx<R>* out = A + cc * g.cA + m;
#pragma unroll
            for (int i = 0; i < F::E; i++) st_nt(out + (long)F::out_pos(t, i) * g.sx, v[i]);
        }
#pragma unroll
        for (int i = 0; i < F::E; i++) v[i] = u[i] - b[i];
        // the first transform's last exchange reads precede this one's first writes: barrier first
        F::template run<T, false, NoPad, false, 1>(v, lds + c, t, twl);
        if (ok) {
            cplx<R>* out = A + (3 + cc) * g.cA + m;
#pragma unroll
            for (int i = 0; i < F::E; i++) st_nt(out + (long)F::out_pos(t, i) * g.sx, v[i]);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int l = F::in_pos(t, i);
        cplx<R> a = czero<R>();
        if (ok && lkept(g.mx, l)) {
            const long off = (long)lq(g.mx, l) * g.sx + m;
            if (MODE == 0) {
                a = S[oc * g.cS + off];
            } else {
                double km = kmag_f(kx_four(l, g.N1) * g.f1, Ky, Kz);
                if ((km > sel.inner) && (km <= sel.outer)) a = S[oc * g.cS + off];
            }
        }
        v[i] = a;
    }
    F::template run<T, false, NoPad, false, 0>(v, lds + c, t, twl);   // lds unused before: no first pre-write barrier
    if (ok) {
        cplx<R>* out = A + oc * g.cA + m;
#pragma unroll
        for (int i = 0; i < F::E; i++) st_nt(out + (long)F::out_pos(t, i) * g.sx, v[i]);
    }
}

// pass 2: y-inverse, A[x][cy][cz] -> B[x][y][cz]
template <typename R, int L, int P, int T>
__global__ __launch_bounds__(T * P) void k_fy_inv(FG g, const cplx<R>* __restrict__ A, cplx<R>* __restrict__ B,
                                                  const cplx<R>* __restrict__ tw, int ncomp)
{
    using F = LineFFT<R, L, P, +1>;
    __shared__ cplx<R> lds[L * T];
    __shared__ cplx<R> twl[L];
    stage_tw<R, L>(twl, tw);
    __syncthreads();
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
        v[i] = (ok && lkept(g.my, l)) ? ld_nt(a + (long)lq(g.my, l) * g.KZP) : czero<R>();
    }
    F::template run<T, false, NoPad, false, 0>(v, lds + c, t, twl);   // lds unused before: no first pre-write barrier
    if (ok) {
        cplx<R>* b = B + comp * g.cB + (long)x * g.sB + cz;
#pragma unroll
        for (int i = 0; i < F::E; i++) st_nt(b + (long)F::out_pos(t, i) * g.KZP, v[i]);
    }
}

// pass 3: fused z.  Input lines hold kz < nkz; c2r to the z-packed real line, products, r2c.
// MODE 0 (NS):  in u0..u2,            out u0u0 u1u1 u2u2 u0u1 u1u2 u2u0          (RSprod.cc:52-77)
// MODE 1 (MHD): in Z+0..2, Z-0..2,    out P(i,j) = Z+_i Z-_j at 3i+j             (RSprod.cc:117-151)
// MODE 2 (ET):  in H0..2,  G0..2,     out P(i,j) = G_i H_j at 3i+j               (ET_RSprod.cc:49-85)
template <int MODE> struct ZProd;
// Products are also listed as 3 rotation steps r of NPR products each: product j of step r multiplies
// x = u[G0 + r] by operand yop(j) (-1 - k: u[G0 + (r + k) % 3]; k >= 0: input k) into output pout(r, j).
template <> struct ZProd<0> {
    static constexpr int NIN = 3, NOUT = 6;
    static constexpr int a(int p) { return p < 3 ? p : p - 3; }
    static constexpr int b(int p) { return p < 3 ? p : (p == 5 ? 0 : p - 2); }
    static constexpr int G0 = 0, NPR = 2;
    static constexpr int yop(int j) { return j == 0 ? -1 : -2; }
    static constexpr int pout(int r, int j) { return j == 0 ? r : 3 + r; }
};
template <> struct ZProd<1> {
    static constexpr int NIN = 6, NOUT = 9;
    static constexpr int a(int p) { return p / 3; }
    static constexpr int b(int p) { return 3 + p % 3; }
    static constexpr int G0 = 0, NPR = 3;
    static constexpr int yop(int j) { return 3 + j; }
    static constexpr int pout(int r, int j) { return 3 * r + j; }
};
template <> struct ZProd<2> {
    static constexpr int NIN = 6, NOUT = 9;
    static constexpr int a(int p) { return 3 + p / 3; }
    static constexpr int b(int p) { return p % 3; }
    static constexpr int G0 = 3, NPR = 3;
    static constexpr int yop(int j) { return j; }
    static constexpr int pout(int r, int j) { return 3 * r + j; }
};
// MODE 3 (RB): in V1 V2 V3 T, out V1V1 V2V2 V3V3 V1V2 V2V3 V3V1 V1T V2T V3T (compute_nlin_SF.cc:80-177)
template <> struct ZProd<3> {
    static constexpr int NIN = 4, NOUT = 9;
    static constexpr int a(int p) { return p % 3; }
    static constexpr int b(int p) { return p < 3 ? p : (p < 6 ? (p - 2) % 3 : 3); }
    static constexpr int G0 = 0, NPR = 3;
    static constexpr int yop(int j) { return j == 0 ? -1 : j == 1 ? -2 : 3; }
    static constexpr int pout(int r, int j) { return 3 * j + r; }
};

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
    constexpr bool R16 = (LZ / P) % 16 == 0;   // 16 points per thread: radix-16 stages
    using FI = LineFFT<R, LZ, P, +1, R16>;
    using FF = LineFFT<R, LZ, P, -1, R16>;
    using Lay = std::conditional_t<R16, ZPad16<LZ>, ZPad<LZ>>;
    using ZP = ZProd<MODE>;
    constexpr int E = FI::E;
    static_assert(64 % P == 0, "a z line's threads must share one wavefront");
    constexpr int LZS = Lay::line;   // padded line stride in LDS
    static_assert(LZS >= LZ + 1, "the mirror exchange uses positions 0..LZ of the line's LDS region");
    __shared__ cplx<R> lds[TZ * LZS];
    // Twiddles live in LDS: W_LZ^j for the line FFTs, then W_N3^k (k < LZ) for the half-length
    // packing.  Global twiddle loads inside the product loop would queue behind the product stores
    // (one in-order vector-memory counter) and stall every FFT stage on store completion.
    __shared__ cplx<R> tws[2 * LZ];
    for (int i = threadIdx.x; i < 2 * LZ; i += TZ * P) tws[i] = i < LZ ? twL[i] : twN[i - LZ];
    __syncthreads();
    const int t = threadIdx.x % P, c = threadIdx.x / P;
    const long q = (long)blockIdx.x * TZ + c;
    const bool ok = q < (long)g.N1 * g.N2;
    const long off = q * g.KZP;
    B += blockIdx.y * g.bB;
    Cout += blockIdx.y * g.bC;
    cplx<R>* Z = lds + c * LZS;
    // The half-length packing pairs point k = t + P m with LZ - k.  The line's points go to
    // Z[0..LZ) (plain order) with Z[LZ] = the k = N3/2 coefficient (inverse) or point 0 (forward,
    // period LZ), so every lane reads its mirror at Z[LZ - t - P m] with no lane-0 special case.
    cplx<R> twk[E];   // W_N3^(t + P m)
#pragma unroll
    for (int m = 0; m < E; m++) twk[m] = tws[LZ + t + P * m];
    // The state is 2/3-dealiased (th_state_create refuses other plans): a line holds kz < NKZ = N3/3, so
    // register m (points k = t + P m) is zero for every lane when P m >= NKZ, and its mirror LZ - k is
    // zero for every lane when LZ - (P - 1) - P m >= NKZ (LZ itself, the kz = N3/2 coefficient, included).
    // Both are known at compile time: those loads, stores and terms are left out.
    constexpr int NKZ = (2 * LZ) / 3;
    auto zin = [](int m) constexpr { return P * m >= NKZ; };
    auto zmir = [](int m) constexpr { return LZ - (P - 1) - P * m >= NKZ; };
    static_assert(LZ - (P - 1) >= NKZ, "register 0's mirrors (position LZ included) must be dealiased zeros");
    cplx<R> u[ZP::NIN][E];
#pragma unroll
    for (int a = 0; a < ZP::NIN; a++) {
        const cplx<R>* X = B + a * g.cB + off;
        cplx<R> xa[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + P * m;
            xa[m] = (!zin(m) && ok && k < NKZ) ? ld_nt(X + k) : czero<R>();
        }
        wave_sync();
        // every position is written (the zeros too): mixed registers read mirrors up to LZ - NKZ + P
#pragma unroll
        for (int m = 0; m < E; m++) Z[t + P * m] = xa[m];
        wave_sync();
#pragma unroll
        for (int m = 0; m < E; m++) {
            cplx<R> w = conjg(twk[m]);
            if (zmir(m)) {                    // s = d = x_k
                cplx<R> xv = xa[m];
                if (t == 0 && m == 0) xv.y = R(0);
                u[a][zreg<E, FI::R0>(m)] = add_iwd(xv, w, xv);
            } else if (zin(m)) {              // s = conj(x_{LZ-k}), d = -s
                cplx<R> xb = conjg(Z[LZ - t - P * m]);
                u[a][zreg<E, FI::R0>(m)] = add_iwd(xb, mk<R>(-w.x, -w.y), xb);
            } else {
                cplx<R> xb = Z[LZ - t - P * m];
                cplx<R> xv = xa[m];
                xb.y = -xb.y;
                cplx<R> s = xv + xb, d = xv - xb;
                u[a][zreg<E, FI::R0>(m)] = add_iwd(s, w, d);
            }
        }
        FI::template run<1, true, Lay>(u[a], Z, t, tws);
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
        // twiddles re-read per product (opaque offset): hoisting them would hold 2 E more VGPRs
        const cplx<R>* tw = tws + opaque_zero();
        FF::template run<1, true, Lay>(w, Z, t, tw);
        const cplx<R>* twp = tw + LZ;   // W_N3^(t + P m)
        cplx<R>* X = Cout + p * g.cB + off;
        wave_sync();
        // only mirrors LZ - k of kept k < NKZ are read: positions above LZ - NKZ
#pragma unroll
        for (int m = 0; m < E; m++)
            if (P * m + (P - 1) > LZ - NKZ) Z[t + P * m] = w[zreg<E, FF::RL>(m)];
        if (t == 0) Z[LZ] = w[zreg<E, FF::RL>(0)];
        wave_sync();
        // 2 X_k = (z_k + conj z_{LZ-k}) - i W^k (z_k - conj z_{LZ-k}); the factor 1/2 is left to the
        // consumers of the product lines, whose normalisation FG::norm carries it
        auto unpack = [&](int m) {
            cplx<R> zk = w[zreg<E, FF::RL>(m)];
            cplx<R> zc = conjg(Z[LZ - t - P * m]);
            const cplx<R> tw = twp[t + P * m];
            return add_iwd(zk + zc, mk<R>(-tw.x, -tw.y), zk - zc);
        };
        // registers whose points all lie at kz >= NKZ are neither unpacked nor stored
        if constexpr (sizeof(R) == 4) {
            // unpack every point first, then store: no LDS read waits between the stores
            cplx<R> o[E];
#pragma unroll
            for (int m = 0; m < E; m++)
                if (!zin(m)) o[m] = unpack(m);
#pragma unroll
            for (int m = 0; m < E; m++) {
                const int k = t + P * m;
                if (!zin(m) && ok && k < NKZ) st_nt(X + k, o[m]);
            }
        } else {
            // FP64: holding every unpacked point would cost 4 E VGPRs
#pragma unroll
            for (int m = 0; m < E; m++) {
                if (zin(m)) continue;
                const int k = t + P * m;
                const cplx<R> o = unpack(m);
                if (ok && k < NKZ) st_nt(X + k, o);
            }
        }
    };
    if constexpr (sizeof(R) == 4 && E <= 8) {
        // FP32, 8 or fewer points per thread: a rolled loop over the 3 rotation steps (fewer live
        // registers than the unrolled form below: more waves per SIMD); rotating the group's registers
        // by one between steps puts every operand in a fixed register, so no per-product selects.
        constexpr int G0 = ZP::G0;
#pragma unroll 1
        for (int r = 0; r < 3; r++) {
#pragma unroll
            for (int j = 0; j < ZP::NPR; j++) {
                const int yo = ZP::yop(j);
                cplx<R> w[E];
#pragma unroll
                for (int i = 0; i < E; i++) {
                    const cplx<R> x = u[G0][i];
                    const cplx<R> y = yo < 0 ? u[G0 - 1 - yo][i] : u[yo][i];
                    w[i] = mk<R>(x.x * y.x, x.y * y.y);
                }
                product_out(w, ZP::pout(r, j));
            }
#pragma unroll
            for (int i = 0; i < E; i++) {
                const cplx<R> t0 = u[G0][i];
                u[G0][i] = u[G0 + 1][i];
                u[G0 + 1][i] = u[G0 + 2][i];
                u[G0 + 2][i] = t0;
            }
        }
    } else if constexpr (sizeof(R) == 4) {
        // FP32: products one after another, each fenced off for the scheduler, so that the compiler
        // neither interleaves products nor holds two products' lines and twiddles at once (registers,
        // occupancy); every operand is a compile-time register.
        constexpr int G0 = ZP::G0;
#pragma unroll
        for (int r = 0; r < 3; r++) {
#pragma unroll
            for (int j = 0; j < ZP::NPR; j++) {
                const int yo = ZP::yop(j);
                const int xi = G0 + r, yi = yo < 0 ? G0 + (r - 1 - yo) % 3 : yo;
                cplx<R> w[E];
#pragma unroll
                for (int i = 0; i < E; i++) w[i] = mk<R>(u[xi][i].x * u[yi][i].x, u[xi][i].y * u[yi][i].y);
                product_out(w, ZP::pout(r, j));
                __builtin_amdgcn_sched_barrier(0);
            }
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
    __shared__ cplx<R> twl[L];
    stage_tw<R, L>(twl, tw);
    __syncthreads();
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int cz = blockIdx.x * T + c;
    const bool ok = cz < g.nkz;
    const int x = blockIdx.y, comp = blockIdx.z % ncomp, box = blockIdx.z / ncomp;
    Cin += box * g.bC;
    D += box * g.bA;
    const cplx<R>* a = Cin + comp * g.cB + (long)x * g.sB + cz;
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) v[i] = ok ? ld_nt(a + (long)F::in_pos(t, i) * g.KZP) : czero<R>();
    F::template run<T, false, NoPad, false, 0>(v, lds + c, t, twl);   // lds unused before: no first pre-write barrier
    if (ok) {
        cplx<R>* d = D + comp * g.cA + (long)x * g.sx + cz;
#pragma unroll
        for (int i = 0; i < F::E; i++) {
            const int l = F::out_pos(t, i);
            if (lkept(g.my, l)) st_nt(d + (long)lq(g.my, l) * g.KZP, v[i]);
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
// s0 and a0 are the values of S and acc at this point (a0 unused in stage 1), loaded by the caller.
// Returns the stage's output, the field the next stage transforms: the stage field (stages 1-3) or the new
// state (stage 4), as stored.  WVS = false leaves the stage field unwritten (its only reader, the next x-inverse,
// then runs on the returned value in registers).
template <typename R>
__device__ __forceinline__ cplx<R> rk_store(int stage, double dt, double eh, double e1, const Dc& rhs, const cplx<R>& s0,
                                            const cplx<R>& a0, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, bool wvs = true)
{
    const Dc V0 = dl(s0);
    cplx<R> out;
    if (stage == 1) {
        Dc v = dadd(V0, dsc(rhs, 0.5 * dt));
        out = ds<R>(dsc(v, eh));
        st_nt(acc, ds<R>(dsc(dsc(rhs, e1), dt / 6)));
    } else if (stage == 2) {
        out = ds<R>(dadd(dsc(V0, eh), dsc(rhs, 0.5 * dt)));
        st_nt(acc, ds<R>(dadd(dl(a0), dsc(dsc(rhs, eh), dt / 3))));
    } else if (stage == 3) {
        Dc v = dadd(dsc(V0, eh), dsc(rhs, 1.0 * dt));
        out = ds<R>(dsc(v, eh));
        st_nt(acc, ds<R>(dadd(dl(a0), dsc(dsc(rhs, eh), dt / 3))));
    } else {
        Dc a = dadd(dl(a0), dsc(rhs, dt / 6));
        out = ds<R>(dadd(dsc(V0, e1), a));
        *S = out;
        return out;
    }
    if (wvs) *Vs = out;
    return out;
}
template <typename R>
__device__ __forceinline__ void rk_update(int stage, double dt, double eh, double e1, const Dc& rhs, cplx<R>* S,
                                          cplx<R>* Vs, cplx<R>* acc)
{
    const cplx<R> s0 = *S;
    const cplx<R> a0 = stage > 1 ? *acc : czero<R>();
    rk_store<R>(stage, dt, eh, e1, rhs, s0, a0, S, Vs, acc);
}

// first register after (i) that xdrop keeps, or -1
template <int L, int P, int E, int RL> constexpr int xnext(int i)
{
    for (int j = i + 1; j < E; j++)
        if (!xdrop<L, P, E, RL>(j)) return j;
    return -1;
}

// CHAIN: the epilogue also performs the next stage's x-inverse: the stage output (stage field, or the new
// state in stage 4) stays in the accumulator registers in place of N and is transformed back along x, and the
// A lines are written over this tile's columns of D (every product of the tile has been read by then, and no
// other workgroup reads these columns); the stage field itself is not written, except the stage-3 output (the
// last stage's input) of a forced state (FG::keep3), from which the host evaluates the outputs' force.  NS: components 0-2 = U.
// MHD: the Elsasser lines of k_fx_inv MODE 1, Z+ = U + B into components 0-2 and Z- = U - B into 3-5.
template <typename R, int L, int P, int T, int MODE, bool CHAIN = false>
__global__ __launch_bounds__(T * P) __attribute__((amdgpu_waves_per_eu((MODE == 0 && sizeof(R) == 4 && L % 3 != 0) ? 4 : 1))) void k_fx_fwd(FG g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs,
                                                  cplx<R>* acc, const cplx<R>* __restrict__ tw,
                                                  const StageCoef* __restrict__ coefs)
{
    using F = LineFFT<R, L, P, -1>;
    constexpr int E = F::E;
    constexpr int NP = (MODE == 0) ? 6 : 9;
    constexpr int NA = (MODE == 0) ? 3 : 6;
    constexpr int NX = (MODE == 1) ? 4 : 2;
    __shared__ cplx<R> lds[2 * L * T];   // two exchange buffers (LineFFT DBUF)
    __shared__ cplx<R> twl[L];
    // exp(c Kx^2) per row and exp(c (Ky^2 + Kz^2)) per column for the integrating factors c = fh[0], f1[0]
    // (MHD: and fh[1], f1[1]): exp(c K^2) = exp(c Kx^2) exp(c (Ky^2 + Kz^2))
    __shared__ double exl[NX][L];
    __shared__ double eyz[NX][T];
    stage_tw<R, L>(twl, tw);
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int ntz = (g.nkz + T - 1) / T;
    const int lin = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, box = lin / gridDim.x;
    const int cy = tile / ntz;
    const int cz = (tile % ntz) * T + c;
    const bool ok = cz < g.nkz;
    const StageCoef sc = coefs[box];
    for (int l = threadIdx.x; l < L; l += T * P) {
        const double kx = kx_four(l, g.N1) * g.f1, kx2 = kx * kx;
        exl[0][l] = exp(sc.fh[0] * kx2);
        exl[1][l] = exp(sc.f1[0] * kx2);
        if constexpr (MODE == 1) {
            exl[2][l] = exp(sc.fh[1] * kx2);
            exl[3][l] = exp(sc.f1[1] * kx2);
        }
    }
    if (t == 0) {   // one thread per column
        const double ky = ky_of(ll(g.my, cy), g.N2) * g.f2, kz = cz * g.f3, kyz2 = ky * ky + kz * kz;
        eyz[0][c] = exp(sc.fh[0] * kyz2);
        eyz[1][c] = exp(sc.f1[0] * kyz2);
        if constexpr (MODE == 1) {
            eyz[2][c] = exp(sc.fh[1] * kyz2);
            eyz[3][c] = exp(sc.f1[1] * kyz2);
        }
    }
    __syncthreads();
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
    const unsigned lane_b = (unsigned)lane_off * (unsigned)sizeof(cplx<R>);
    // Where T divides the KZP padding (KZP is nkz rounded up to 16 FP32 / 8 FP64 elements) a column past nkz
    // (!ok) still lies inside its row, so the loads are unconditional: such a column's values only ever reach its
    // own lines, which are never stored.  Wider tiles (short lines) keep the guard.
    constexpr bool PADOK = (sizeof(R) == 4 ? 16 : 8) % T == 0;
#pragma unroll 1
    for (int p = 0; p < NP; p++) {
        cplx<R> v[E];
        const cplx<R>* src = D + p * g.cA + mu;
#pragma unroll
        for (int i = 0; i < E; i++) {
            const cplx<R>* row = src + (long)F::in_pos(0, i) * g.sx;
            v[i] = (PADOK || ok) ? *at_byte(row, lane_b) : czero<R>();
        }
        F::template run<T, false, NoPad, true>(v, lds + c, t, twl + opaque_zero());
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
    static_assert(!CHAIN || (MODE <= 1 && F::R0 == F::RL), "chained x-inverse: NS or MHD, first radix = last radix");
    if constexpr (!CHAIN)
        if (!ok) return;
    const double dKy = ky_of(ll(g.my, cy), g.N2) * g.f2, dKz = cz * g.f3;
    // Epilogue, software-pipelined over the kept registers: the state and accumulator values of
    // the next register are loaded before this register's results are stored.  Vector-memory
    // operations retire through one in-order counter, so a load issued after a store would wait
    // for the store to complete.
    constexpr int NC = (MODE == 1) ? 6 : 3;
    const int stage = sc.stage;
    const bool wvs = !CHAIN || (g.keep3 && stage == 3);   // uniform
    auto row_u = [&](int i) -> long {
        // compact row of l = t + l0: the kept lanes of one register all sit on one side of the
        // dealiased gap (P < L/3), so the row shift is uniform per register; the lane adds lane_off
        const int l0 = F::out_pos(0, i);
        return (long)(l0 - (l0 >= g.mx.h ? g.mx.N - g.mx.nk : 0)) * g.sx;
    };
    // Unconditional loads (a branch around them would pull the first use, and its wait, up to the
    // load): lanes in the dealiased gap read row 0 of their column, and stage 1, which has no
    // accumulator yet, reads S a second time (a cache hit); neither value is used.
    const cplx<R>* accr = stage > 1 ? acc : S;
    // byte offset of register i's row for this lane from the tile's column base (S + a cS + mu); gap lanes
    // (not loaded from, see above) take row 0
    auto row_b = [&](int i) -> unsigned { return lane_b + (unsigned)(row_u(i) * (long)sizeof(cplx<R>)); };
    auto fetch = [&](int i, cplx<R> (&s0)[NC], cplx<R> (&a0)[NC]) {
        const unsigned b = lkept(g.mx, F::out_pos(t, i)) ? row_b(i) : lane_b;
#pragma unroll
        for (int a = 0; a < NC; a++) {
            s0[a] = *at_byte(S + a * g.cS + mu, b);
            a0[a] = *at_byte(accr + a * g.cS + mu, b);
        }
    };
    // CHAIN: no prefetch (its registers would push the 1024-thread FP32 kernel past 128 VGPRs; the chained
    // epilogue stores no stage field for a load to wait on in stages 1-3)
    cplx<R> nS[NC], nA[NC];
    if (!CHAIN && ok) fetch(xnext<L, P, E, F::RL>(-1), nS, nA);
#pragma unroll
    for (int i = 0; i < E; i++) {
        if (xdrop<L, P, E, F::RL>(i)) {
            if constexpr (CHAIN)
#pragma unroll
                for (int a = 0; a < NA; a++) Nacc[a][i] = czero<R>();
            continue;
        }
        cplx<R> s0[NC], a0[NC];
        if constexpr (CHAIN) {
            if (ok) fetch(i, s0, a0);
        } else {
#pragma unroll
            for (int a = 0; a < NC; a++) {
                s0[a] = nS[a];
                a0[a] = nA[a];
            }
            if (ok && xnext<L, P, E, F::RL>(i) >= 0) fetch(xnext<L, P, E, F::RL>(i), nS, nA);
        }
        const int l = F::out_pos(t, i);
        cplx<R> keep[3] = {czero<R>(), czero<R>(), czero<R>()};
        cplx<R> keepb[3] = {czero<R>(), czero<R>(), czero<R>()};   // MHD: B's stage output
        if (ok && lkept(g.mx, l)) {
            const double dKx = kx_four(l, g.N1) * g.f1;
            const double K2 = ksq_f(dKx, dKy, dKz);
            const unsigned bk = row_b(i);
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
                const double iK2 = 1.0 / K2;
                Dc kn = dmk((dKx * Nu[0].x + dKy * Nu[1].x + dKz * Nu[2].x) * iK2,
                            (dKx * Nu[0].y + dKy * Nu[1].y + dKz * Nu[2].y) * iK2);
                const double Kc[3] = {dKx, dKy, dKz};
#pragma unroll
                for (int a = 0; a < 3; a++) rhs[a] = dmk(-(Nu[a].x - Kc[a] * kn.x), -(Nu[a].y - Kc[a] * kn.y));
            } else {
#pragma unroll
                for (int a = 0; a < 3; a++) rhs[a] = dmk(-Nu[a].x, -Nu[a].y);
            }
            const double eh = exl[0][l] * eyz[0][c], e1 = exl[1][l] * eyz[1][c];
#pragma unroll
            for (int a = 0; a < 3; a++)
                keep[a] = rk_store<R>(stage, sc.dt, eh, e1, rhs[a], s0[a], a0[a], at_byte(S + a * g.cS + mu, bk),
                                      at_byte(Vs + a * g.cS + mu, bk), at_byte(acc + a * g.cS + mu, bk), wvs);
            if constexpr (MODE == 1) {
                const double ehb = exl[2][l] * eyz[2][c], e1b = exl[3][l] * eyz[3][c];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    Dc np = dl(Nacc[a][i]), nm = dl(Nacc[3 + a][i]);
                    Dc nb = dsc(dmk(np.x - nm.x, np.y - nm.y), 0.5 * g.norm);
                    Dc rb = dmk(-nb.x, -nb.y);
                    keepb[a] = rk_store<R>(stage, sc.dt, ehb, e1b, rb, s0[3 + a], a0[3 + a],
                                           at_byte(S + (3 + a) * g.cS + mu, bk),
                                           at_byte(Vs + (3 + a) * g.cS + mu, bk),
                                           at_byte(acc + (3 + a) * g.cS + mu, bk), wvs);
                }
            }
        }
        if constexpr (CHAIN) {
            // rows outside the kept modes: 0
            if constexpr (MODE == 0) {
#pragma unroll
                for (int a = 0; a < 3; a++) Nacc[a][i] = keep[a];
            } else {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    Nacc[a][i] = keep[a] + keepb[a];
                    Nacc[3 + a][i] = keep[a] - keepb[a];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // one register at a time: bounds the live state values
    }
    if constexpr (CHAIN) {
        // the next stage's x-inverse (k_fx_inv MODE 0 / 1): register i holds row F::out_pos(t, i), which is
        // FI::in_pos(t, i) since the schedule's first and last radix agree
        using FI = LineFFT<R, L, P, +1>;
        cplx<R>* Aout = const_cast<cplx<R>*>(D);
#pragma unroll
        for (int a = 0; a < NA; a++) {   // unrolled: Nacc[a] must stay a compile-time register
            FI::template run<T, false, NoPad, true>(Nacc[a], lds + c, t, twl + opaque_zero());
            if (ok) {
                // row FI::out_pos(t, i) = t + FI::out_pos(0, i) of column mu + c: a uniform row base plus lane_b
                cplx<R>* out = Aout + a * g.cA + mu;
#pragma unroll
                for (int i = 0; i < E; i++) *at_byte(out + (long)FI::out_pos(0, i) * g.sx, lane_b) = Nacc[a][i];
            }
            __builtin_amdgcn_sched_barrier(0);   // one line at a time
        }
    }
}

// ------------------------------------------------------------------------------------------
// SCFT (Rayleigh-Benard) x passes.  Along x the reference uses FFTW's RODFT10/REDFT10 forward and
// RODFT01/REDFT01 inverse (scft_tr.cc:49-639) with sine coefficients shifted right by one
// (ArrayShiftRight/Left_SCFT :153-219) and the 1/(2 N1 N2 N3) norm.  Here each r2r line of
// length N costs one N-point complex FFT (the complex data are two real sequences, and both maps
// are complex-linear):
//   inverse DCT-III:  V_k = conj(w_k) (X_k - i X_{N-k}) (X_N = 0), v = sum_k V_k e^{+2 pi i jk/N}
//   forward DCT-II:   V = FFT(v), X_k = w_k V_k + conj(w_k) V_{N-k},      w_k = e^{-i pi k / 2N}
// where v_j is the real-space sample x_{2j} (j < N/2) or x_{2N-2j-1} (j >= N/2).  Real space is
// only used point-wise (products in the z pass), so rows stay in that order and the permutation
// cancels between the passes.  Sine lines: y_j = (-1)^j DCT-III(c_{N-l})_j inverse and
// A_l = DCT-II((-1)^j x_j)_{N-l} = i (w_l V_l - conj(w_l) V_{N-l}) forward; (-1)^j of the sample in
// row j is +1 for j < N/2 and -1 above.
// ------------------------------------------------------------------------------------------
template <typename R, int L, int P, int T>
__global__ __launch_bounds__(T * P) void k_sx_inv(FG g, const cplx<R>* __restrict__ S, cplx<R>* __restrict__ A,
                                                  const cplx<R>* __restrict__ tw, const cplx<R>* __restrict__ tw4)
{
    using F = LineFFT<R, L, P, +1>;
    __shared__ cplx<R> lds[L * T];
    __shared__ cplx<R> twl[L];
    stage_tw<R, L>(twl, tw);
    __shared__ cplx<R> tw4l[L];
    stage_tw<R, L>(tw4l, tw4);
    __syncthreads();
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int ntz = (g.nkz + T - 1) / T;
    const int cy = blockIdx.x / ntz;
    const int cz = (blockIdx.x % ntz) * T + c;
    const bool ok = cz < g.nkz;
    const long m = (long)cy * g.KZP + cz;
    const int oc = blockIdx.y;                    // V1 (sin), V2, V3 (cos), T (sin)
    const bool sn = (oc == 0 || oc == 3);
    S += blockIdx.z * g.bS;
    A += blockIdx.z * g.bA;
    const cplx<R>* col = S + oc * g.cS + m;
    const int nk = g.mx.nk;
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int k = F::in_pos(t, i);
        cplx<R> ck = (ok && k < nk && !(sn && k == 0)) ? col[(long)k * g.sx] : czero<R>();
        cplx<R> cm = (ok && k > 0 && L - k < nk) ? col[(long)(L - k) * g.sx] : czero<R>();
        const cplx<R> a = sn ? cm : ck, b = sn ? ck : cm;          // X_k - i X_{N-k}
        const cplx<R> x = mk<R>(a.x + b.y, a.y - b.x);
        v[i] = cmul(conjg(tw4l[k]), x);
    }
    F::template run<T, false, NoPad, false, 0>(v, lds + c, t, twl);   // lds unused before: no first pre-write barrier
    if (ok) {
        cplx<R>* out = A + oc * g.cA + m;
#pragma unroll
        for (int i = 0; i < F::E; i++) {
            const int j = F::out_pos(t, i);
            const R sg = (sn && j >= L / 2) ? R(-1) : R(1);
            st_nt(out + (long)j * g.sx, cscale(v[i], sg));
        }
    }
}

// registers holding only positions l >= N1 - N1/3 (dropped by the SCFT 2/3 rule) for every lane
template <int L, int P, int E, int RL> __host__ __device__ constexpr bool sdrop(int i)
{
    const int m = i / RL + (i % RL) * (E / RL);
    return m * P >= 2 * L / 3;
}
// first register after (i) that sdrop keeps, or -1
template <int L, int P, int E, int RL> constexpr int snext(int i)
{
    for (int j = i + 1; j < E; j++)
        if (!sdrop<L, P, E, RL>(j)) return j;
    return -1;
}

// RB x-forward + epilogue: forward DCT-II / DST-II of the 9 product lines, N_i = sum_j d_j(V_i V_j)
// and N_T = sum_j d_j(V_j T) with Xderiv_Sin (+k f1) / Xderiv_Cos (-k f1) along x (scft_tr.cc:
// 649-672), the RB force (Force1 = T, T.Force = V1; compute_force_RB.cc:51-105, add_force.cc:57-63),
// pressure projection (compute_pressure.cc:54-62 with the SCFT divergence), Compute_rhs and the
// FP64 IF-RK4 update with nu for V and kappa for T (time_advance_SF.cc:105-222).
template <typename R, int L, int P, int T>
__global__ __launch_bounds__(T * P) void k_sx_fwd_rb(FG g, const cplx<R>* __restrict__ D, cplx<R>* S, cplx<R>* Vs,
                                                     cplx<R>* acc, const cplx<R>* __restrict__ tw,
                                                     const cplx<R>* __restrict__ tw4,
                                                     const StageCoef* __restrict__ coefs)
{
    using F = LineFFT<R, L, P, -1>;
    constexpr int E = F::E;
    // two exchange buffers (LineFFT DBUF); the output table of each product goes to the one the FFT's
    // last exchange left free, and the next product's first exchange waits for its mirror reads (PRE0 = 1)
    __shared__ cplx<R> lds[2 * L * T];
    cplx<R>* const ob = lds + F::FREEBUF * (L * T);
    __shared__ cplx<R> twl[L];
    // exp(c Kx^2) per row for c = fh[0], f1[0] (velocity) and fh[1], f1[1] (temperature); the
    // (Ky, Kz) factor of exp(c K^2) is computed once per thread
    __shared__ double exl[4][L];
    stage_tw<R, L>(twl, tw);
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int ntz = (g.nkz + T - 1) / T;
    const int lin = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, box = lin / gridDim.x;
    const int cy = tile / ntz;
    const int cz = (tile % ntz) * T + c;
    const bool ok = cz < g.nkz;
    const long m = (long)cy * g.KZP + cz;
    const StageCoef sc = coefs[box];
    for (int l = threadIdx.x; l < L; l += T * P) {
        const double kx = l * g.f1, kx2 = kx * kx;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            exl[2 * q][l] = exp(sc.fh[q] * kx2);
            exl[2 * q + 1][l] = exp(sc.f1[q] * kx2);
        }
    }
    __syncthreads();
    D += box * g.bA;
    S += box * g.bS;
    Vs += box * g.bS;
    acc += box * g.bS;
    const R Ky = (R)(ky_of(ll(g.my, cy), g.N2) * g.f2);
    const R Kz = (R)(cz * g.f3);
    const int nk = g.mx.nk;
    cplx<R> Nacc[4][E];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < E; i++) Nacc[a][i] = czero<R>();
    const long mu = (long)cy * g.KZP + (long)(tile % ntz) * T;
    const int lane_off = c + t * (int)g.sx;
#pragma unroll 1
    for (int p = 0; p < 9; p++) {
        // forward kind of product p: sine for V1V2 (3), V3V1 (5), V2T (7), V3T (8)
        const bool psin = ((0x1A8u >> p) & 1u) != 0;
        cplx<R> v[E];
        const cplx<R>* src = D + p * g.cA + mu;
#pragma unroll
        for (int i = 0; i < E; i++) {
            const int j0 = F::in_pos(0, i);
            const cplx<R>* row = src + (long)j0 * g.sx;
            const cplx<R> x = ok ? row[lane_off] : czero<R>();
            const R sg = (psin && j0 + t >= L / 2) ? R(-1) : R(1);
            v[i] = cscale(x, sg);
        }
        F::template run<T, false, NoPad, true, 1>(v, lds + c, t, twl + opaque_zero());
        if constexpr (F::NST < 2) __syncthreads();    // no exchange: the previous product's mirror reads
#pragma unroll
        for (int i = 0; i < E; i++) ob[F::out_pos(t, i) * T + c] = v[i];
        __syncthreads();
        // product p adds d_j F(p) to N_a: Xderiv_Cos (-Kx) / Xderiv_Sin (+Kx) along x, i Ky, i Kz
        const int a1 = (p == 0 || p == 3 || p == 5) ? 0 : (p == 1 || p == 4) ? 1 : (p == 2) ? 2 : 3;
        const R al1 = (p == 0 || p == 6) ? R(-1) : R(0);
        const R be1 = (p == 1 || p == 3 || p == 7) ? Ky : (p == 2 || p == 4 || p == 5 || p == 8) ? Kz : R(0);
        const int a2 = (p == 3) ? 1 : (p == 4 || p == 5) ? 2 : -1;      // second target
        const R al2 = (p == 3 || p == 5) ? R(1) : R(0);
        const R be2 = (p == 4) ? Ky : R(0);
        const cplx<R>* w4 = tw4 + opaque_zero();   // read per product (no stores in the loop)
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (sdrop<L, P, E, F::RL>(i)) continue;
            const int l = F::out_pos(t, i);
            const cplx<R> w = w4[l];
            const cplx<R> Pv = cmul(w, v[i]);
            const cplx<R> Qv = cmul(conjg(w), ob[((L - l) % L) * T + c]);
            const cplx<R> dv = Pv - Qv, sv = Pv + Qv;
            cplx<R> Av = psin ? mk<R>(-dv.y, dv.x) : sv;                        // i (P - Q) or P + Q
            if (psin && l == 0) Av = czero<R>();
            const R Kx = (R)(l * g.f1);
            const cplx<R> c1 = mk<R>(al1 * Kx * Av.x - be1 * Av.y, al1 * Kx * Av.y + be1 * Av.x);
            const cplx<R> c2 = mk<R>(al2 * Kx * Av.x - be2 * Av.y, al2 * Kx * Av.y + be2 * Av.x);
#pragma unroll
            for (int a = 0; a < 4; a++) {
                if (a == a1) Nacc[a][i] = Nacc[a][i] + c1;
                if (a == a2) Nacc[a][i] = Nacc[a][i] + c2;
            }
        }
    }
    if (!ok) return;
    const double dKy = ky_of(ll(g.my, cy), g.N2) * g.f2, dKz = cz * g.f3;
    const int stage = sc.stage;
    const cplx<R>* F0 = (stage == 1) ? S : Vs;     // field the stage's rhs is evaluated on
    // Epilogue software-pipelined like the FOUR x-forward: the next kept register's S, acc and
    // stage-field values are loaded (unconditionally; lanes past the kept rows read row 0, unused)
    // before this register's results are stored.
    const cplx<R>* accr = stage > 1 ? acc : S;         // stage 1: no accumulator yet (value unused)
    auto fetch = [&](int i, cplx<R> (&s0)[4], cplx<R> (&a0)[4], cplx<R> (&f0)[2]) {
        const int l = F::out_pos(t, i);
        const long idx = (l < nk ? (long)l * g.sx : 0) + m;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            s0[a] = S[a * g.cS + idx];
            a0[a] = accr[a * g.cS + idx];
        }
        f0[0] = F0[idx];
        f0[1] = F0[3 * g.cS + idx];
    };
    cplx<R> nS[4], nA[4], nF[2];
    fetch(snext<L, P, E, F::RL>(-1), nS, nA, nF);
    const double kyz2 = dKy * dKy + dKz * dKz;
    const double eyh = exp(sc.fh[0] * kyz2), ey1 = exp(sc.f1[0] * kyz2);
    const double eyhT = exp(sc.fh[1] * kyz2), ey1T = exp(sc.f1[1] * kyz2);
#pragma unroll
    for (int i = 0; i < E; i++) {
        if (sdrop<L, P, E, F::RL>(i)) continue;
        cplx<R> s0[4], a0[4], f0[2];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            s0[a] = nS[a];
            a0[a] = nA[a];
        }
        f0[0] = nF[0];
        f0[1] = nF[1];
        if (snext<L, P, E, F::RL>(i) >= 0) fetch(snext<L, P, E, F::RL>(i), nS, nA, nF);
        const int l = F::out_pos(t, i);
        if (l < nk) {
            const double dKx = l * g.f1;
            const double K2 = ksq_f(dKx, dKy, dKz);
            const long idx = (long)l * g.sx + m;
            Dc Nn[4];
#pragma unroll
            for (int a = 0; a < 4; a++) Nn[a] = dsc(dl(Nacc[a][i]), g.norm);
            const Dc V1 = dl(f0[0]), Tf = dl(f0[1]);
            Nn[0] = dmk(Nn[0].x - Tf.x, Nn[0].y - Tf.y);            // nlin - Force1, Force1 = T
            Nn[3] = dmk(Nn[3].x - V1.x, Nn[3].y - V1.y);            // T.nlin - T.Force, T.Force = V1
            Dc pr = dmk(0.0, 0.0);
            if (K2 > 0) {
                // div N = Xderiv_Sin(N1) + i Ky N2 + i Kz N3, p = div / K^2
                const double iK2 = 1.0 / K2;
                const double dr = dKx * Nn[0].x - dKy * Nn[1].y - dKz * Nn[2].y;
                const double di = dKx * Nn[0].y + dKy * Nn[1].x + dKz * Nn[2].x;
                pr = dmk(dr * iK2, di * iK2);
            }
            // rhs = -(N + grad p): Xderiv_Cos(p) = -Kx p, d_y p = i Ky p, d_z p = i Kz p
            Dc rhs[4];
            rhs[0] = dmk(-(Nn[0].x - dKx * pr.x), -(Nn[0].y - dKx * pr.y));
            rhs[1] = dmk(-(Nn[1].x - dKy * pr.y), -(Nn[1].y + dKy * pr.x));
            rhs[2] = dmk(-(Nn[2].x - dKz * pr.y), -(Nn[2].y + dKz * pr.x));
            rhs[3] = dmk(-Nn[3].x, -Nn[3].y);
            const double eh = exl[0][l] * eyh, e1 = exl[1][l] * ey1;
            const double ehT = exl[2][l] * eyhT, e1T = exl[3][l] * ey1T;
#pragma unroll
            for (int a = 0; a < 4; a++)
                rk_store<R>(stage, sc.dt, a < 3 ? eh : ehT, a < 3 ? e1 : e1T, rhs[a], s0[a], a0[a], S + a * g.cS + idx,
                            Vs + a * g.cS + idx, acc + a * g.cS + idx);
        }
        __builtin_amdgcn_sched_barrier(0);
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
    __shared__ cplx<R> lds[2 * L * T];   // two exchange buffers (LineFFT DBUF)
    __shared__ cplx<R> twl[L];
    stage_tw<R, L>(twl, tw);
    __syncthreads();
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
        F::template run<T, false, NoPad, true>(v, lds + c, t, twl + opaque_zero());
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

template <typename R> constexpr int tileT() { return 128 / (int)sizeof(cplx<R>); }

// kernel dispatch on line length
#define TH_FLENS(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#define TH_ZLENS(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

// launch shape of variant V of an x/y line pass
constexpr int alt_pts(int L) { return (L % 3 == 0) ? 24 : (pts_per_thread(L) == 16 ? 8 : 16); }
template <int L, int V> constexpr int var_E()
{
    return ((V & 2) && L >= 128 && L % alt_pts(L) == 0) ? alt_pts(L) : pts_per_thread(L);
}
template <typename R, int L, int V> constexpr int var_T() { return ((V & 1) && L >= 128) ? tileT<R>() / 2 : tileT<R>(); }

template <typename R, int L, int MODE, int V>
static int launch_xinv(const FG& g, const cplx<R>* S, cplx<R>* A, int ncomp, const cplx<R>* tw, ShellSel sel, hipStream_t st)
{
    constexpr int E = var_E<L, V>(), P = L / E, T = var_T<R, L, V>();
    const int ntz = (g.nkz + T - 1) / T;
    // MHD: one workgroup per component pair (U_c, B_c) writes both Elsasser components
    const int ny = (MODE == 1) ? ncomp / 2 : ncomp;
    hipLaunchKernelGGL((k_fx_inv<R, L, P, T, MODE>), dim3(g.my.nk * ntz, ny, g.nbox), dim3(T * P), 0, st, g, S, A, tw, sel);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int L, int V>
static int launch_yinv(const FG& g, const cplx<R>* A, cplx<R>* B, int ncomp, const cplx<R>* tw, hipStream_t st)
{
    constexpr int E = var_E<L, V>(), P = L / E, T = var_T<R, L, V>();
    hipLaunchKernelGGL((k_fy_inv<R, L, P, T>), dim3((g.nkz + T - 1) / T, g.N1, ncomp * g.nbox), dim3(T * P), 0, st, g, A, B,
                       tw, ncomp);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int LZ, int MODE, int V>
static int launch_z(const FG& g, const cplx<R>* B, cplx<R>* Cc, const cplx<R>* twL, const cplx<R>* twN, hipStream_t st)
{
    constexpr int E = (V == 3) ? 16 : pts_per_thread(LZ), P = LZ / E;
    constexpr int NT = (V == 1) ? 128 : (V == 2) ? 512 : 256;
    constexpr int TZ = (NT / P) < 1 ? 1 : NT / P;
    if (g.nkz != (2 * LZ) / 3) {   // k_fz takes the kept kz < N3/3 of a dealiased state at compile time
        set_error("fused z pass: %d kept kz for N3 = %d (a dealiased state keeps %d)", g.nkz, 2 * LZ, (2 * LZ) / 3);
        return E_UNSUPPORTED;
    }
    const long lines = (long)g.N1 * g.N2;
    hipLaunchKernelGGL((k_fz<R, LZ, P, TZ, MODE>), dim3((unsigned)((lines + TZ - 1) / TZ), g.nbox), dim3(TZ * P), 0, st, g, B, Cc,
                       twL, twN);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R, int L, int V>
static int launch_yfwd(const FG& g, const cplx<R>* Cc, cplx<R>* D, int ncomp, const cplx<R>* tw, hipStream_t st)
{
    constexpr int E = var_E<L, V>(), P = L / E, T = var_T<R, L, V>();
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

template <typename R, int L, int MODE, int V, bool CHAIN = false>
static int launch_xfwd(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw,
                       const StageCoef* sc, hipStream_t st)
{
    constexpr int E = xfwd_E(L), P = L / E, T0 = xfwd_T<R, L, MODE>();
    // variant 1: half-width tiles; variant 2: double-width tiles (full 128-byte rows at 512 lines)
    constexpr int T = (V == 1 && L >= 128 && T0 > 1) ? T0 / 2 : (V == 2) ? 2 * T0 : T0;
    const int ntz = (g.nkz + T - 1) / T;
    // the kernel's unconditional product loads (tiles that divide the row padding) need the last tile inside a row
    if ((sizeof(R) == 4 ? 16 : 8) % T == 0 && ntz * T > g.KZP) {
        set_error("x-forward: %d columns of %d-column tiles exceed the row padding %d", ntz * T, T, g.KZP);
        return E_UNSUPPORTED;
    }
    if constexpr (CHAIN) {
        if constexpr (MODE <= 1 && LineFFT<R, L, P, -1>::R0 == LineFFT<R, L, P, -1>::RL) {
            hipLaunchKernelGGL((k_fx_fwd<R, L, P, T, MODE, true>), dim3(g.my.nk * ntz, g.nbox), dim3(T * P), 0, st, g, D,
                               S, Vs, acc, tw, sc);
        } else {
            set_error("chained x-forward: no chained kernel for N1 = %d", L);
            return E_UNSUPPORTED;
        }
    } else {
        hipLaunchKernelGGL((k_fx_fwd<R, L, P, T, MODE>), dim3(g.my.nk * ntz, g.nbox), dim3(T * P), 0, st, g, D, S, Vs, acc,
                           tw, sc);
    }
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

template <typename R, int L>
static int launch_sxinv(const FG& g, const cplx<R>* S, cplx<R>* A, const cplx<R>* tw, const cplx<R>* tw4, hipStream_t st)
{
    constexpr int E = pts_per_thread(L), P = L / E, T = tileT<R>();
    const int ntz = (g.nkz + T - 1) / T;
    hipLaunchKernelGGL((k_sx_inv<R, L, P, T>), dim3(g.my.nk * ntz, 4, g.nbox), dim3(T * P), 0, st, g, S, A, tw, tw4);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
// variant 1: half-width tiles (more, smaller workgroups: less idle tail at 256^3, where full tiles leave
// 1026 workgroups for 256 CUs holding one each)
template <typename R, int L, int V>
static int launch_sxfwd(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw,
                        const cplx<R>* tw4, const StageCoef* sc, hipStream_t st)
{
    constexpr int E = xfwd_E(L), P = L / E, T0 = xfwd_T<R, L, 0>();
    constexpr int T = (V == 1 && L >= 128 && T0 > 1) ? T0 / 2 : T0;
    const int ntz = (g.nkz + T - 1) / T;
    hipLaunchKernelGGL((k_sx_fwd_rb<R, L, P, T>), dim3(g.my.nk * ntz, g.nbox), dim3(T * P), 0, st, g, D, S, Vs, acc, tw,
                       tw4, sc);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

// variant dispatch per line length (short lines have one shape)
template <typename R, int MODE, int L>
static int xinv_v(const FG& g, const cplx<R>* S, cplx<R>* A, int nc, const cplx<R>* tw, ShellSel sel, int var, hipStream_t st)
{
    if constexpr (L >= 128) {
        if (var == 1) return launch_xinv<R, L, MODE, 1>(g, S, A, nc, tw, sel, st);
        if (var == 2) return launch_xinv<R, L, MODE, 2>(g, S, A, nc, tw, sel, st);
        if (var == 3) return launch_xinv<R, L, MODE, 3>(g, S, A, nc, tw, sel, st);
    }
    return launch_xinv<R, L, MODE, 0>(g, S, A, nc, tw, sel, st);
}
template <typename R, int L>
static int yinv_v(const FG& g, const cplx<R>* A, cplx<R>* B, int nc, const cplx<R>* tw, int var, hipStream_t st)
{
    if constexpr (L >= 128) {
        if (var == 1) return launch_yinv<R, L, 1>(g, A, B, nc, tw, st);
        if (var == 2) return launch_yinv<R, L, 2>(g, A, B, nc, tw, st);
        if (var == 3) return launch_yinv<R, L, 3>(g, A, B, nc, tw, st);
    }
    return launch_yinv<R, L, 0>(g, A, B, nc, tw, st);
}
template <typename R, int MODE, int LZ>
static int z_v(const FG& g, const cplx<R>* B, cplx<R>* Cc, const cplx<R>* twL, const cplx<R>* twN, int var, hipStream_t st)
{
    if constexpr (LZ >= 64) {
        if (var == 1) return launch_z<R, LZ, MODE, 1>(g, B, Cc, twL, twN, st);
        // 512-thread workgroups cap a wave at 256 VGPRs and 16 points per thread hold a line in 32
        // VGPRs: without spills only for FP32 NS power-of-two lines
        if constexpr (sizeof(R) == 4 && MODE == 0 && LZ % 3 != 0) {
            if (var == 2) return launch_z<R, LZ, MODE, 2>(g, B, Cc, twL, twN, st);
            if (var == 3) return launch_z<R, LZ, MODE, 3>(g, B, Cc, twL, twN, st);
        }
    }
    return launch_z<R, LZ, MODE, 0>(g, B, Cc, twL, twN, st);
}
template <typename R, int L>
static int yfwd_v(const FG& g, const cplx<R>* Cc, cplx<R>* D, int nc, const cplx<R>* tw, int var, hipStream_t st)
{
    if constexpr (L >= 128) {
        if (var == 1) return launch_yfwd<R, L, 1>(g, Cc, D, nc, tw, st);
        if (var == 2) return launch_yfwd<R, L, 2>(g, Cc, D, nc, tw, st);
        if (var == 3) return launch_yfwd<R, L, 3>(g, Cc, D, nc, tw, st);
    }
    return launch_yfwd<R, L, 0>(g, Cc, D, nc, tw, st);
}
template <typename R, int MODE, int L>
static int xfwd_v(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw, const StageCoef* sc,
                  int var, hipStream_t st, bool chain)
{
    if (chain) {
        if constexpr (L >= 128)
            if (var == 1) return launch_xfwd<R, L, MODE, 1, true>(g, D, S, Vs, acc, tw, sc, st);
        if constexpr (L >= 128 && sizeof(R) == 4 && MODE == 0 && L % 3 != 0 && 2 * xfwd_T<R, L, MODE>() * (L / xfwd_E(L)) <= 1024)
            if (var == 2) return launch_xfwd<R, L, MODE, 2, true>(g, D, S, Vs, acc, tw, sc, st);
        return launch_xfwd<R, L, MODE, 0, true>(g, D, S, Vs, acc, tw, sc, st);
    }
    if constexpr (L >= 128)
        if (var == 1) return launch_xfwd<R, L, MODE, 1>(g, D, S, Vs, acc, tw, sc, st);
    // 1024-thread workgroups hold a wave to 128 VGPRs: only the FP32 NS kernels fit (124)
    if constexpr (L >= 128 && sizeof(R) == 4 && MODE == 0 && L % 3 != 0 && 2 * xfwd_T<R, L, MODE>() * (L / xfwd_E(L)) <= 1024)
        if (var == 2) return launch_xfwd<R, L, MODE, 2>(g, D, S, Vs, acc, tw, sc, st);
    return launch_xfwd<R, L, MODE, 0>(g, D, S, Vs, acc, tw, sc, st);
}
// whether the chained x-forward exists for this line length and variant (NS or MHD, first radix = last radix)
template <typename R, int L, int MODE, int V> constexpr bool xfwd_chainable()
{
    constexpr int E = xfwd_E(L), P = L / E;
    return MODE <= 1 && LineFFT<R, L, P, -1>::R0 == LineFFT<R, L, P, -1>::RL;
}

template <typename R, int MODE>
int d_xinv(const FG& g, const cplx<R>* S, cplx<R>* A, int nc, const cplx<R>* tw, ShellSel sel, int var, hipStream_t st)
{
#define C_(n) case n: return xinv_v<R, MODE, n>(g, S, A, nc, tw, sel, var, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
int d_yinv(const FG& g, const cplx<R>* A, cplx<R>* B, int nc, const cplx<R>* tw, int var, hipStream_t st)
{
#define C_(n) case n: return yinv_v<R, n>(g, A, B, nc, tw, var, st);
    switch (g.N2) { TH_FLENS(C_) default: set_error("fused: N2 %d unsupported", g.N2); return E_UNSUPPORTED; }
#undef C_
}
template <typename R, int MODE>
int d_z(const FG& g, const cplx<R>* B, cplx<R>* Cc, const cplx<R>* twL, const cplx<R>* twN, int var, hipStream_t st)
{
#define C_(n) case n: return z_v<R, MODE, n>(g, B, Cc, twL, twN, var, st);
    switch (g.LZ) { TH_ZLENS(C_) default: set_error("fused: N3 %d unsupported", g.N3); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
int d_yfwd(const FG& g, const cplx<R>* Cc, cplx<R>* D, int nc, const cplx<R>* tw, int var, hipStream_t st)
{
#define C_(n) case n: return yfwd_v<R, n>(g, Cc, D, nc, tw, var, st);
    switch (g.N2) { TH_FLENS(C_) default: set_error("fused: N2 %d unsupported", g.N2); return E_UNSUPPORTED; }
#undef C_
}
template <typename R, int MODE>
int d_xfwd(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw, const StageCoef* sc,
           int var, hipStream_t st, bool chain)
{
#define C_(n) case n: return xfwd_v<R, MODE, n>(g, D, S, Vs, acc, tw, sc, var, st, chain);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
bool d_xfwd_chainable(const FG& g)
{
    // every variant of a line length, NS or MHD, uses the same points per thread, hence the same radix schedule
#define C_(n) case n: return xfwd_chainable<R, n, 0, 0>();
    switch (g.N1) { TH_FLENS(C_) default: return false; }
#undef C_
}
template <typename R>
int d_xfwd_et(const FG& g, const cplx<R>* D, const cplx<R>* U, const cplx<R>* tw, const double* radii, int nr, int kin,
              double* partial, int* nblk, hipStream_t st)
{
#define C_(n) case n: return launch_xfwd_et<R, n>(g, D, U, tw, radii, nr, kin, partial, nblk, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}

template <typename R>
int d_sx_inv(const FG& g, const cplx<R>* S, cplx<R>* A, const cplx<R>* tw, const cplx<R>* tw4, hipStream_t st)
{
#define C_(n) case n: return launch_sxinv<R, n>(g, S, A, tw, tw4, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}
template <typename R>
int d_sx_fwd_rb(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw,
                const cplx<R>* tw4, const StageCoef* sc, int var, hipStream_t st)
{
#define C_(n)                                                                                   \
    case n:                                                                                     \
        return var == 1 ? launch_sxfwd<R, n, 1>(g, D, S, Vs, acc, tw, tw4, sc, st)             \
                        : launch_sxfwd<R, n, 0>(g, D, S, Vs, acc, tw, tw4, sc, st);
    switch (g.N1) { TH_FLENS(C_) default: set_error("fused: N1 %d unsupported", g.N1); return E_UNSUPPORTED; }
#undef C_
}

static dim3 grid_1d(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return dim3((unsigned)(b ? b : 1));
}
template <typename R> int d_gather(const FG& g, const cplx<R>* ref, cplx<R>* S, hipStream_t st)
{
    hipLaunchKernelGGL((k_gather<R>), grid_1d((size_t)g.mx.nk * g.sx), dim3(256), 0, st, g, ref, S);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R> int d_scatter(const FG& g, const cplx<R>* S, cplx<R>* ref, hipStream_t st)
{
    hipLaunchKernelGGL((k_scatter<R>), grid_1d((size_t)g.N1 * g.N2 * (g.N3 / 2 + 1)), dim3(256), 0, st, g, S, ref);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

#define TH_FUSED_INSTANTIATE(R)                                                                                       \
    template int d_xinv<R, 0>(const FG&, const cplx<R>*, cplx<R>*, int, const cplx<R>*, ShellSel, int, hipStream_t);  \
    template int d_xinv<R, 1>(const FG&, const cplx<R>*, cplx<R>*, int, const cplx<R>*, ShellSel, int, hipStream_t);  \
    template int d_xinv<R, 2>(const FG&, const cplx<R>*, cplx<R>*, int, const cplx<R>*, ShellSel, int, hipStream_t);  \
    template int d_yinv<R>(const FG&, const cplx<R>*, cplx<R>*, int, const cplx<R>*, int, hipStream_t);              \
    template int d_z<R, 0>(const FG&, const cplx<R>*, cplx<R>*, const cplx<R>*, const cplx<R>*, int, hipStream_t);   \
    template int d_z<R, 1>(const FG&, const cplx<R>*, cplx<R>*, const cplx<R>*, const cplx<R>*, int, hipStream_t);   \
    template int d_z<R, 2>(const FG&, const cplx<R>*, cplx<R>*, const cplx<R>*, const cplx<R>*, int, hipStream_t);   \
    template int d_z<R, 3>(const FG&, const cplx<R>*, cplx<R>*, const cplx<R>*, const cplx<R>*, int, hipStream_t);   \
    template int d_sx_inv<R>(const FG&, const cplx<R>*, cplx<R>*, const cplx<R>*, const cplx<R>*, hipStream_t);      \
    template int d_sx_fwd_rb<R>(const FG&, const cplx<R>*, cplx<R>*, cplx<R>*, cplx<R>*, const cplx<R>*,            \
                                const cplx<R>*, const StageCoef*, int, hipStream_t);                                  \
    template int d_yfwd<R>(const FG&, const cplx<R>*, cplx<R>*, int, const cplx<R>*, int, hipStream_t);              \
    template int d_xfwd<R, 0>(const FG&, const cplx<R>*, cplx<R>*, cplx<R>*, cplx<R>*, const cplx<R>*,              \
                              const StageCoef*, int, hipStream_t, bool);                                               \
    template int d_xfwd<R, 1>(const FG&, const cplx<R>*, cplx<R>*, cplx<R>*, cplx<R>*, const cplx<R>*,              \
                              const StageCoef*, int, hipStream_t, bool);                                               \
    template bool d_xfwd_chainable<R>(const FG&);                                                                      \
    template int d_xfwd_et<R>(const FG&, const cplx<R>*, const cplx<R>*, const cplx<R>*, const double*, int, int,    \
                              double*, int*, hipStream_t);                                                             \
    template int d_gather<R>(const FG&, const cplx<R>*, cplx<R>*, hipStream_t);                                        \
    template int d_scatter<R>(const FG&, const cplx<R>*, cplx<R>*, hipStream_t);

}  // namespace th
