// th_fused.hpp -- shared declarations of the dealias-pruned fused pipeline: compact-layout
// geometry (FG), per-box RK stage coefficients, and the pass launchers, whose kernels are compiled
// once per precision (th_fused_f32.hip, th_fused_f64.hip) and called by the host code in
// th_fused.hip.
#pragma once
#include <math.h>
#include <string.h>
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
    int scft;              // sin/cos along x (SCFT basis), else Fourier
    double f1, f2, f3;
    double norm; // 1 / (2 N1 N2 N3): the line transforms' 1 / (N1 N2 N3) and the z pass's unpacking 1/2 (SCFT: 1/2 more)
    int keep3 = 0; // chained x-forward of a forced state: write the stage-3 output (the last stage's input field)
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


// ---- pass launchers (defined in th_fused_kern.hpp, instantiated per precision) ----
// Tunable launch shapes.  A variant picks, per pass, how many lines share a workgroup and how many
// points each thread holds; the autotuner times them on the live buffers (th_tune) and the
// choice is kept per (pass, grid, precision, physics, boxes) in the context's tune cache.
//   x/y passes (PASS_XINV, PASS_YINV, PASS_YFWD): bit 0 halves the lines per workgroup, bit 1
//     switches to the alternative points per thread; lines shorter than 128 have variant 0 only.
//   z pass: 0 = 256 threads per workgroup, 1 = 128, 2 = 512, 3 = 16 points per thread with
//     radix-16 stages (z lines of 64 or more; 2 and 3 only for FP32 NS power-of-two lines, the
//     others would spill).
//   x-forward epilogue: 0 = default lines per workgroup, 1 = half (lines of 128 or more).
enum { PASS_XINV = 0, PASS_YINV = 1, PASS_Z = 2, PASS_YFWD = 3, PASS_XFWD = 4, NPASS = 5 };
constexpr int pass_nvar(int pass) { return pass == PASS_Z ? 4 : pass == PASS_XFWD ? 3 : 4; }

template <typename R, int MODE>
int d_xinv(const FG& g, const cplx<R>* S, cplx<R>* A, int nc, const cplx<R>* tw, ShellSel sel, int var, hipStream_t st);
template <typename R>
int d_yinv(const FG& g, const cplx<R>* A, cplx<R>* B, int nc, const cplx<R>* tw, int var, hipStream_t st);
template <typename R, int MODE>
int d_z(const FG& g, const cplx<R>* B, cplx<R>* Cc, const cplx<R>* twL, const cplx<R>* twN, int var, hipStream_t st);
template <typename R>
int d_yfwd(const FG& g, const cplx<R>* Cc, cplx<R>* D, int nc, const cplx<R>* tw, int var, hipStream_t st);
template <typename R, int MODE>
int d_xfwd(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw, const StageCoef* sc,
           int var, hipStream_t st, bool chain = false);
// chain: the x-forward also transforms its stage output back along x and writes the next stage's A lines over
// D's first components (NS: U in 0-2; MHD: Z+ in 0-2, Z- in 3-5; the stage field is not written);
// d_xfwd_chainable says whether N1 has that kernel
template <typename R>
bool d_xfwd_chainable(const FG& g);
template <typename R>
int d_xfwd_et(const FG& g, const cplx<R>* D, const cplx<R>* U, const cplx<R>* tw, const double* radii, int nr, int kin,
              double* partial, int* nblk, hipStream_t st);
template <typename R>
int d_sx_inv(const FG& g, const cplx<R>* S, cplx<R>* A, const cplx<R>* tw, const cplx<R>* tw4, hipStream_t st);
template <typename R>
int d_sx_fwd_rb(const FG& g, const cplx<R>* D, cplx<R>* S, cplx<R>* Vs, cplx<R>* acc, const cplx<R>* tw,
                const cplx<R>* tw4, const StageCoef* sc, int var, hipStream_t st);
template <typename R> int d_gather(const FG& g, const cplx<R>* ref, cplx<R>* S, hipStream_t st);
template <typename R> int d_scatter(const FG& g, const cplx<R>* S, cplx<R>* ref, hipStream_t st);

FG make_fg(const Geo& geo, int prec);

}  // namespace th
