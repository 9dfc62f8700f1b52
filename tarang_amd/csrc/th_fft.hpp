// th_fft.hpp -- cooperative line FFT for gfx950: P threads of a workgroup own one line of length L
// (E = L/P points each, held in VGPRs).  Stockham autosort stages (radix 8/4/2/3) run in registers;
// between stages the line is exchanged through LDS.  The first stage reads its inputs from the
// registers the caller filled in "input order" and the last stage leaves outputs in natural order
// in registers, so HBM is touched only by the caller's coalesced loads and stores.
//
// Convention (FFTW's, as used by trunk/lib/four/four_tr.cc): SIGN = -1 forward exp(-2 pi i jk/L),
// SIGN = +1 backward, no normalisation.  Twiddles W_L^m = exp(-2 pi i m/L) come from a table that
// the host fills from FP64 libm (rounded to the storage precision).
#pragma once
#include "th_common.hpp"

namespace th {

struct Sched {
    int n;
    int r[16];
};

// Greedy radix schedule: each radix must divide the remaining length and the points per thread.
// Radix 16 only on request (it needs all 16 points of a butterfly live at once).
constexpr Sched make_sched(int L, int E, bool r16 = false)
{
    Sched s{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    int m = L;
    while (m > 1) {
        int R = 0;
        if (r16 && m % 16 == 0 && E % 16 == 0) R = 16;
        else if (m % 8 == 0 && E % 8 == 0) R = 8;
        else if (m % 4 == 0 && E % 4 == 0) R = 4;
        else if (m % 2 == 0 && E % 2 == 0) R = 2;
        else if (m % 3 == 0 && E % 3 == 0) R = 3;
        if (R == 0) { s.n = -1; return s; }
        s.r[s.n++] = R;
        m /= R;
    }
    // Prefer a schedule whose first and last radix agree: then the input order of a forward FFT
    // equals the output order of an inverse one and fused passes skip an LDS shuffle.
    if (s.n >= 3 && s.r[0] != s.r[s.n - 1]) {
        for (int i = 1; i < s.n - 1; i++)
            if (s.r[i] == s.r[0]) {
                int t = s.r[i];
                for (int j = i; j < s.n - 1; j++) s.r[j] = s.r[j + 1];
                s.r[s.n - 1] = t;
                break;
            }
    }
    return s;
}

template <typename R> struct K {
    static constexpr R s12 = R(0.70710678118654752440084436210485);   // sqrt(1/2)
    static constexpr R s32 = R(0.86602540378443864676372317075294);   // sqrt(3)/2
};

template <int SIGN, typename R> __device__ __forceinline__ cplx<R> mul_si(cplx<R> a)   // (SIGN*i)*a
{
    return SIGN < 0 ? mul_mi(a) : mul_i(a);
}

// Orders the LDS accesses of one wavefront (the compiler may not move LDS reads or writes across
// it).  A wavefront's LDS operations execute in issue order, so lines owned by the lanes of one
// wavefront exchange through LDS without a workgroup barrier.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// LDS placement of a line's elements during exchange s: element A lives at A + k (A >> sh).
// The identity suits the interleaved layouts (LS > 1: consecutive lanes own different lines).
// ZPad suits line-contiguous layouts (LS = 1: consecutive lanes own one line, lines `line`
// elements apart), where the Stockham scatter A = (j / Ns) Ns Rr + j % Ns + r Ns of consecutive
// j would otherwise pile onto a few banks.  (sh, k) per exchange and the line stride minimise the
// conflict degree of the gfx950 LDS (64 banks of 4 bytes; a 64-bit access serves 32 lanes per
// pass) over every access of the radix schedules below.
// Radix-16 z schedules ([16, 4], [16, 8], [16, 16]): one padding element every 16 and a line stride
// of 16 mod 32 elements keep every access conflict-free under the same bank model.
template <int L> struct ZPad16 {
    static constexpr int map(int, int A) { return A + (A >> 4); }
    static constexpr int line = L + L / 16;
};
struct NoPad {
    static constexpr int map(int, int A) { return A; }
};
template <int L> struct ZPad {
    static constexpr int code(int s)
    {
        // sh * 16 + k; 0 = identity
        return L == 16 ? (s == 0 ? 3 * 16 + 1 : 0)
             : L == 24 ? (s <= 1 ? 2 * 16 + 1 : 0)
             : L == 32 ? (s == 0 ? 3 * 16 + 1 : 0)
             : L == 48 ? (s <= 1 ? 2 * 16 + 1 : 0)
             : L == 64 ? (s == 0 ? 3 * 16 + 1 : 0)
             : L == 96 ? (s <= 2 ? 3 * 16 + 1 : 0)
             : L == 128 ? (s <= 1 ? 4 * 16 + 1 : 0)
             : L == 192 ? (s == 0 || s == 2 ? 4 * 16 + 1 : s == 1 ? 2 * 16 + 1 : 0)
             : L == 256 ? (s == 0 ? 5 * 16 + 1 : s == 1 ? 3 * 16 + 2 : 0)
             : 0;
    }
    static constexpr int map(int s, int A) { return code(s) ? A + (code(s) & 15) * (A >> (code(s) >> 4)) : A; }
    static constexpr int need()
    {
        int m = L;
        for (int s = 0; s < 8; s++) m = map(s, L - 1) + 1 > m ? map(s, L - 1) + 1 : m;
        return m;
    }
    static constexpr int line = L == 16 ? 18 : L == 24 ? 30 : L == 32 ? 36 : L == 48 ? 60 : L == 64 ? 72 : L == 96 ? 120 : L == 128 ? 144 : L == 192 ? 240 : L == 256 ? 318 : L + 1;
    static_assert(line >= need(), "ZPad line stride too short");
};

// Workgroup barrier that orders LDS accesses only: the fences are scoped to the local address space, so
// the compiler waits for this wave's LDS operations (lgkmcnt) but not for its outstanding global-memory
// operations, which keeps loads into an LDS staging buffer (global_load_lds) in flight across it.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <bool WAVE> __device__ __forceinline__ void line_sync()
{
    if constexpr (WAVE) wave_sync();
    else lds_barrier();
}

template <int RADIX, int SIGN, typename R> struct DFT;

template <int SIGN, typename R> struct DFT<2, SIGN, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        cplx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <int SIGN, typename R> struct DFT<4, SIGN, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        cplx<R> t0 = v[0] + v[2], t1 = v[0] - v[2];
        cplx<R> t2 = v[1] + v[3], t3 = mul_si<SIGN>(v[1] - v[3]);
        v[0] = t0 + t2;
        v[2] = t0 - t2;
        v[1] = t1 + t3;
        v[3] = t1 - t3;
    }
};

template <int SIGN, typename R> struct DFT<8, SIGN, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        const R c = K<R>::s12;
        cplx<R> a[4], b[4];
#pragma unroll
        for (int n = 0; n < 4; n++) { a[n] = v[n] + v[n + 4]; b[n] = v[n] - v[n + 4]; }
        // b[n] *= W8^n, W8 = exp(SIGN i pi/4)
        b[1] = mk<R>(c * (b[1].x - SIGN * b[1].y), c * (b[1].y + SIGN * b[1].x));
        b[2] = mul_si<SIGN>(b[2]);
        b[3] = mk<R>(-c * (b[3].x + SIGN * b[3].y), c * (SIGN * b[3].x - b[3].y));
        DFT<4, SIGN, R>::run(a);
        DFT<4, SIGN, R>::run(b);
#pragma unroll
        for (int k = 0; k < 4; k++) { v[2 * k] = a[k]; v[2 * k + 1] = b[k]; }
    }
};

// 16 points as 4 x 4: DFT4 over n1 of x[4 n1 + n2], twiddles W16^(n2 k1), DFT4 over n2;
// X[k1 + 4 k2] lands in v[k1 + 4 k2].
template <int SIGN, typename R> struct DFT<16, SIGN, R> {
    static __device__ __forceinline__ cplx<R> wmul(const cplx<R>& a, R c, R s)   // a * (c + SIGN i s)
    {
        return mk<R>(a.x * c - R(SIGN) * a.y * s, a.y * c + R(SIGN) * a.x * s);
    }
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        const R c1 = R(0.92387953251128675612818318939678828682);   // cos(pi/8)
        const R s1 = R(0.38268343236508977172845998403039886676);   // sin(pi/8)
        const R h = K<R>::s12;
        cplx<R> a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            cplx<R> q[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            DFT<4, SIGN, R>::run(q);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) a[n2][k1] = q[k1];
        }
        a[1][1] = wmul(a[1][1], c1, s1);    // W^1
        a[1][2] = wmul(a[1][2], h, h);      // W^2
        a[1][3] = wmul(a[1][3], s1, c1);    // W^3
        a[2][1] = wmul(a[2][1], h, h);      // W^2
        a[2][2] = mul_si<SIGN>(a[2][2]);    // W^4
        a[2][3] = wmul(a[2][3], -h, h);     // W^6
        a[3][1] = wmul(a[3][1], s1, c1);    // W^3
        a[3][2] = wmul(a[3][2], -h, h);     // W^6
        a[3][3] = wmul(a[3][3], -c1, -s1);  // W^9
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            cplx<R> q[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            DFT<4, SIGN, R>::run(q);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = q[k2];
        }
    }
};

template <int SIGN, typename R> struct DFT<3, SIGN, R> {
    static __device__ __forceinline__ void run(cplx<R>* v)
    {
        cplx<R> t1 = v[1] + v[2];
        cplx<R> t2 = v[0] - cscale(t1, R(0.5));
        cplx<R> t3 = cscale(v[1] - v[2], R(SIGN) * K<R>::s32);
        v[0] = v[0] + t1;
        v[1] = t2 + mul_i(t3);
        v[2] = t2 - mul_i(t3);
    }
};

// One line of length L, P cooperating threads, E = L/P points per thread.
//   before run(): v[i] = x[in_pos(t, i)]       after run(): v[i] = X[out_pos(t, i)]
// lds points at element 0 of this line; element m lives at lds[m * LS].
template <typename R, int L, int P, int SIGN, bool R16 = false> struct LineFFT {
    static constexpr int E = L / P;
    static constexpr Sched S = make_sched(L, E, R16);
    static_assert(L % P == 0, "P must divide L");
    static_assert(S.n > 0, "no radix schedule for (L, E)");
    static constexpr int NST = S.n;
    static constexpr int R0 = S.r[0];
    static constexpr int RL = S.r[S.n - 1];

    static __device__ __forceinline__ int in_pos(int t, int i) { return t + P * (i / R0) + (i % R0) * (L / R0); }
    static __device__ __forceinline__ int out_pos(int t, int i) { return t + P * (i / RL) + (i % RL) * (L / RL); }

    // WAVE: the P threads of every line sit in one wavefront (P divides 64 and each line's
    // threads are consecutive lanes), so LDS exchanges need no workgroup barrier.
    // DBUF: exchange s uses LDS buffer s % 2 (lds and lds + L * LS; the caller provides both), so an exchange
    // never overwrites the buffer the previous exchange is still being read from, and only the barrier after
    // each exchange's writes remains.  Across consecutive FFTs on the same buffers: with an even number of
    // exchanges the next FFT's first exchange writes buffer 0 after the barrier that follows this FFT's last
    // writes (to buffer 1), which every thread reaches after its reads of buffer 0; with an odd number both
    // would use buffer 0 back to back, so the first exchange keeps its barrier before writing.
    // PRE0 overrides that first barrier: -1 as above (always taken without DBUF), 0 never (the caller
    // guarantees that no thread still reads the first exchange's buffer, e.g. nothing read it before this
    // FFT), 1 always (the caller read buffer 0 after the previous FFT).
    template <int LS, bool WAVE = false, class Lay = NoPad, bool DBUF = false, int PRE0 = -1>
    static __device__ __forceinline__ void run(cplx<R> (&v)[E], cplx<R>* lds, int t, const cplx<R>* __restrict__ tw)
    {
        stage<0, 1, LS, WAVE, Lay, DBUF, PRE0>(v, lds, t, tw);
    }
    // With DBUF, the buffer (0 or 1) that the last exchange did not use; after run() returns every thread
    // has finished reading it (the barrier after the last exchange's writes ensures that), so a caller may
    // write it at once.  Without exchanges (NST == 1) there is no such barrier and this is buffer 0.
    static constexpr int FREEBUF = (NST >= 2) ? 1 - ((NST - 2) & 1) : 0;

    template <int s, int Ns, int LS, bool WAVE, class Lay, bool DBUF = false, int PRE0 = -1>
    static __device__ __forceinline__ void stage(cplx<R> (&v)[E], cplx<R>* lds, int t, const cplx<R>* __restrict__ tw)
    {
        constexpr int Rr = S.r[s];
        constexpr int B = E / Rr;
#pragma unroll
        for (int b = 0; b < B; b++) {
            if constexpr (Ns > 1) {
                const int j = t + P * b;
                const int k = j % Ns;
#pragma unroll
                for (int r = 1; r < Rr; r++) {
                    cplx<R> w = tw[r * k * (L / (Ns * Rr))];
                    if (SIGN > 0) w.y = -w.y;
                    v[b * Rr + r] = cmul(v[b * Rr + r], w);
                }
            }
            DFT<Rr, SIGN, R>::run(&v[b * Rr]);
        }
        if constexpr (s + 1 < NST) {
            cplx<R>* buf = DBUF ? lds + (s & 1) * (L * LS) : lds;
            constexpr bool pre = (s == 0 && PRE0 >= 0) ? (PRE0 == 1) : (!DBUF || (s == 0 && (NST - 1) % 2 == 1));
            if constexpr (pre) line_sync<WAVE>();
#pragma unroll
            for (int b = 0; b < B; b++) {
                const int j = t + P * b;
                const int base = (j / Ns) * Ns * Rr + (j % Ns);
#pragma unroll
                for (int r = 0; r < Rr; r++) buf[Lay::map(s, base + r * Ns) * LS] = v[b * Rr + r];
            }
            line_sync<WAVE>();
            constexpr int Rn = S.r[s + 1];
#pragma unroll
            for (int i = 0; i < E; i++) v[i] = buf[Lay::map(s, t + P * (i / Rn) + (i % Rn) * (L / Rn)) * LS];
            stage<s + 1, Ns * Rr, LS, WAVE, Lay, DBUF, PRE0>(v, lds, t, tw);
        }
    }
};

// Points per thread used by the generic passes for a line of length L.
constexpr int pts_per_thread(int L)
{
    return (L % 3 == 0) ? (L >= 12 ? 12 : L) : (L >= 512 ? 16 : (L >= 8 ? 8 : L));
}

}  // namespace th
