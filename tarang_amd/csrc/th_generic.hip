// th_generic.hip -- full (unpruned) transform passes and spectral element-wise kernels on the
// reference layout (N1, N2, N3/2+1).  These back the boundary calls th_forward / th_inverse /
// th_deriv ... and the "plain" GPU step that keeps the reference's operation sequence
// (trunk/lib/IncFlow/nlin_calc/compute_nlin.cc:58-108): every pass reads and writes a whole array.
#include <math.h>
#include "th_fft.hpp"
#include "th_internal.hpp"

namespace th {

// ------------------------------------------------------------------------------------------
// Strided in-place c2c pass: data[o*sO + l*sL + m], FFT over l (length L) for all (o, m).
// A workgroup owns T consecutive columns m (coalesced 8*T / 16*T byte rows) x one whole line.
// ------------------------------------------------------------------------------------------
template <typename R, int L, int P, int T, int SIGN>
__global__ __launch_bounds__(T * P) void k_fft_strided(cplx<R>* __restrict__ data, long sO, long sL, int M, R scale,
                                                       const cplx<R>* __restrict__ tw)
{
    using F = LineFFT<R, L, P, SIGN>;
    __shared__ cplx<R> lds[L * T];
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int m = blockIdx.x * T + c;
    const bool ok = m < M;
    cplx<R>* base = data + (long)blockIdx.y * sO + m;
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) v[i] = ok ? base[(long)F::in_pos(t, i) * sL] : czero<R>();
    F::template run<T>(v, lds + c, t, tw);
    if (F::NST == 1) __syncthreads();
    if (ok) {
#pragma unroll
        for (int i = 0; i < F::E; i++) base[(long)F::out_pos(t, i) * sL] = cscale(v[i], scale);
    }
}

template <typename R> constexpr int tile_cols(int L, int P)
{
    int tb = 128 / (int)sizeof(cplx<R>);
    int t = 256 / P > tb ? 256 / P : tb;
    int cap = 65536 / (L * (int)sizeof(cplx<R>));
    if (t > cap) t = cap;
    if (t * P > 1024) t = 1024 / P;
    return t < 1 ? 1 : t;
}

template <typename R, int L, int SIGN>
static int launch_strided(cplx<R>* data, long sO, long sL, int O, int M, R scale, const cplx<R>* tw, hipStream_t st)
{
    constexpr int E = pts_per_thread(L);
    constexpr int P = L / E;
    constexpr int T = tile_cols<R>(L, P);
    dim3 grid((M + T - 1) / T, O);
    hipLaunchKernelGGL((k_fft_strided<R, L, P, T, SIGN>), grid, dim3(T * P), 0, st, data, sO, sL, M, scale, tw);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

#define TH_LENGTHS(X) X(2) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

template <typename R>
int fft_strided(cplx<R>* data, long sO, long sL, int O, int L, int M, int sign, R scale, const cplx<R>* tw,
                hipStream_t st)
{
    if (L == 1) return OK;
#define TH_CASE(n)                                                                            \
    case n:                                                                                   \
        return sign < 0 ? launch_strided<R, n, -1>(data, sO, sL, O, M, scale, tw, st)         \
                        : launch_strided<R, n, +1>(data, sO, sL, O, M, scale, tw, st);
    switch (L) {
        TH_LENGTHS(TH_CASE)
        default: set_error("fft_strided: unsupported length %d", L); return E_UNSUPPORTED;
    }
#undef TH_CASE
}

// ------------------------------------------------------------------------------------------
// z-pass: packed real line of N3 = 2*LZ samples <-> LZ+1 complex (FFTW in-place r2c/c2r layout),
// via one LZ-point complex FFT.  c2r ignores Im of the kz = 0 and kz = N3/2 entries (FFTW c2r).
// ------------------------------------------------------------------------------------------
template <typename R, int LZ, int P, int TZ>
__global__ __launch_bounds__(TZ * P) void k_z_c2r(cplx<R>* __restrict__ data, long nlines, const cplx<R>* __restrict__ twL,
                                                  const cplx<R>* __restrict__ twN)
{
    using F = LineFFT<R, LZ, P, +1>;
    __shared__ cplx<R> lds[TZ * LZ];
    const int t = threadIdx.x % P, c = threadIdx.x / P;
    const long q = (long)blockIdx.x * TZ + c;
    const bool ok = q < nlines;
    cplx<R>* X = data + q * (LZ + 1);
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int k = F::in_pos(t, i);
        if (ok) {
            cplx<R> a = X[k], b = X[LZ - k];
            if (k == 0) { a.y = R(0); b.y = R(0); }
            b.y = -b.y;                                   // conj(X[LZ-k])
            cplx<R> s = a + b, d = a - b;
            cplx<R> w = twN[k];
            w.y = -w.y;                                   // W_N^{-k}
            v[i] = s + mul_i(cmul(w, d));
        } else {
            v[i] = czero<R>();
        }
    }
    F::template run<1>(v, lds + c * LZ, t, twL);
    __syncthreads();
    if (ok) {
#pragma unroll
        for (int i = 0; i < F::E; i++) X[F::out_pos(t, i)] = v[i];
        if (t == 0) X[LZ] = czero<R>();                   // Zero_pad_last_plane_FOUR
    }
}

template <typename R, int LZ, int P, int TZ>
__global__ __launch_bounds__(TZ * P) void k_z_r2c(cplx<R>* __restrict__ data, long nlines, R scale,
                                                  const cplx<R>* __restrict__ twL, const cplx<R>* __restrict__ twN)
{
    using F = LineFFT<R, LZ, P, -1>;
    __shared__ cplx<R> lds[TZ * LZ];
    const int t = threadIdx.x % P, c = threadIdx.x / P;
    const long q = (long)blockIdx.x * TZ + c;
    const bool ok = q < nlines;
    cplx<R>* X = data + q * (LZ + 1);
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) v[i] = ok ? X[F::in_pos(t, i)] : czero<R>();
    F::template run<1>(v, lds + c * LZ, t, twL);
    __syncthreads();
    cplx<R>* Z = lds + c * LZ;
#pragma unroll
    for (int i = 0; i < F::E; i++) Z[F::out_pos(t, i)] = v[i];
    __syncthreads();
    if (ok) {
        for (int k = t; k <= LZ; k += P) {
            cplx<R> zk = Z[k % LZ], zc = conjg(Z[(LZ - k) % LZ]);
            cplx<R> e = cscale(zk + zc, R(0.5));
            cplx<R> tt = cmul(twN[k], zk - zc);
            X[k] = cscale(mk<R>(e.x + R(0.5) * tt.y, e.y - R(0.5) * tt.x), scale);
        }
    }
}

template <typename R, int LZ>
static int launch_z(cplx<R>* data, long nlines, int dir, R scale, const cplx<R>* twL, const cplx<R>* twN,
                    hipStream_t st)
{
    constexpr int E = pts_per_thread(LZ);
    constexpr int P = LZ / E;
    constexpr int TZ0 = (256 / P) < 1 ? 1 : (256 / P);
    constexpr int TZ = (TZ0 * LZ * (int)sizeof(cplx<R>) > 65536) ? (65536 / (LZ * (int)sizeof(cplx<R>))) : TZ0;
    dim3 grid((unsigned)((nlines + TZ - 1) / TZ));
    if (dir < 0)
        hipLaunchKernelGGL((k_z_r2c<R, LZ, P, TZ>), grid, dim3(TZ * P), 0, st, data, nlines, scale, twL, twN);
    else
        hipLaunchKernelGGL((k_z_c2r<R, LZ, P, TZ>), grid, dim3(TZ * P), 0, st, data, nlines, twL, twN);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
static int fft_z(cplx<R>* data, long nlines, int N3, int dir, R scale, const cplx<R>* twL, const cplx<R>* twN,
                 hipStream_t st)
{
    const int LZ = N3 / 2;
#define TH_CASE(n) \
    case n: return launch_z<R, n>(data, nlines, dir, scale, twL, twN, st);
    switch (LZ) {
        TH_LENGTHS(TH_CASE)
        default: set_error("fft_z: unsupported N3 %d", N3); return E_UNSUPPORTED;
    }
#undef TH_CASE
}

template <typename R>
int fft_z_r2c(cplx<R>* data, long nlines, int N3, R scale, const cplx<R>* twL, const cplx<R>* twN, hipStream_t st)
{
    return fft_z<R>(data, nlines, N3, -1, scale, twL, twN, st);
}
template <typename R>
int fft_z_c2r(cplx<R>* data, long nlines, int N3, const cplx<R>* twL, const cplx<R>* twN, hipStream_t st)
{
    return fft_z<R>(data, nlines, N3, +1, R(1), twL, twN, st);
}

// ------------------------------------------------------------------------------------------
// SCFT x-transforms (FFTW RODFT10 / REDFT10 / REDFT01 / RODFT01 of length N = N1, applied to the
// complex array, i.e. to real and imaginary parts separately) through one 2N-point complex FFT
// of the mirrored sequence.  kind: 0 DCT-II, 1 DST-II, 2 DCT-III, 3 DST-III.
// ------------------------------------------------------------------------------------------
template <typename R, int N, int KIND, int T>
__global__ __launch_bounds__(T * ((2 * N) / pts_per_thread(2 * N))) void k_r2r_x(cplx<R>* __restrict__ data, long sL,
                                                                                 int M, R scale,
                                                                                 const cplx<R>* __restrict__ tw2N,
                                                                                 const cplx<R>* __restrict__ tw4N)
{
    constexpr int L2 = 2 * N;
    constexpr int SIGN = (KIND < 2) ? -1 : +1;
    using F = LineFFT<R, L2, L2 / pts_per_thread(L2), SIGN>;
    constexpr int P = L2 / F::E;
    __shared__ cplx<R> lds[L2 * T];
    const int c = threadIdx.x % T, t = threadIdx.x / T;
    const int m = blockIdx.x * T + c;
    const bool ok = m < M;
    cplx<R>* base = data + m;
    cplx<R> v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int p = F::in_pos(t, i);
        cplx<R> a = czero<R>();
        if (ok) {
            if (KIND == 0 || KIND == 1) {
                if (p < N) a = base[(long)p * sL];
                else { a = base[(long)(L2 - 1 - p) * sL]; if (KIND == 1) { a.x = -a.x; a.y = -a.y; } }
            } else if (KIND == 2) {
                if (p == 0) a = base[0];
                else if (p < N) a = cmul(base[(long)p * sL], conjg(tw4N[p]));            // x_j e^{+i pi j/2N}
                else if (p > N) a = cmul(base[(long)(L2 - p) * sL], tw4N[L2 - p]);       // x_j e^{-i pi j/2N}
            } else {
                if (p == N) a = base[(long)(N - 1) * sL];
                else if (p > 0 && p < N) a = mul_mi(cmul(base[(long)(p - 1) * sL], conjg(tw4N[p])));
                else if (p > N) a = mul_i(cmul(base[(long)(L2 - p - 1) * sL], tw4N[L2 - p]));
            }
        }
        v[i] = a;
    }
    F::template run<T>(v, lds + c, t, tw2N);
    __syncthreads();
    if (ok) {
#pragma unroll
        for (int i = 0; i < F::E; i++) {
            const int q = F::out_pos(t, i);
            if (KIND == 0) {
                if (q < N) base[(long)q * sL] = cscale(cmul(v[i], tw4N[q]), scale);
            } else if (KIND == 1) {
                if (q >= 1 && q <= N) base[(long)(q - 1) * sL] = cscale(mul_i(cmul(v[i], tw4N[q])), scale);
            } else {
                if (q < N) base[(long)q * sL] = cscale(v[i], scale);
            }
        }
    }
    (void)P;
}

template <typename R, int N>
static int launch_r2r(cplx<R>* data, long sL, int M, int kind, R scale, const cplx<R>* tw2N, const cplx<R>* tw4N,
                      hipStream_t st)
{
    constexpr int L2 = 2 * N;
    constexpr int P = L2 / pts_per_thread(L2);
    constexpr int T = tile_cols<R>(L2, P);
    dim3 grid((M + T - 1) / T);
    switch (kind) {
        case 0: hipLaunchKernelGGL((k_r2r_x<R, N, 0, T>), grid, dim3(T * P), 0, st, data, sL, M, scale, tw2N, tw4N); break;
        case 1: hipLaunchKernelGGL((k_r2r_x<R, N, 1, T>), grid, dim3(T * P), 0, st, data, sL, M, scale, tw2N, tw4N); break;
        case 2: hipLaunchKernelGGL((k_r2r_x<R, N, 2, T>), grid, dim3(T * P), 0, st, data, sL, M, scale, tw2N, tw4N); break;
        default: hipLaunchKernelGGL((k_r2r_x<R, N, 3, T>), grid, dim3(T * P), 0, st, data, sL, M, scale, tw2N, tw4N); break;
    }
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
int r2r_x(cplx<R>* data, const Geo& g, int kind, cplx<R>* work, const cplx<R>* tw2N, const cplx<R>* tw4N,
          hipStream_t st)
{
    (void)work;
    const long sL = (long)g.N2 * g.H;
    const int M = g.N2 * g.H;
    const R scale = R(1);
#define TH_CASE(n) \
    case n: return launch_r2r<R, n>(data, sL, M, kind, scale, tw2N, tw4N, st);
    switch (g.N1) {
        TH_CASE(4) TH_CASE(8) TH_CASE(12) TH_CASE(16) TH_CASE(24) TH_CASE(32) TH_CASE(48)
        TH_CASE(64) TH_CASE(96) TH_CASE(128) TH_CASE(192) TH_CASE(256) TH_CASE(512)
        default: set_error("r2r_x: unsupported N1 %d", g.N1); return E_UNSUPPORTED;
    }
#undef TH_CASE
}

// Sine-coefficient shift along x (ArrayShiftRight/Left_SCFT, trunk/lib/scft/scft_tr.cc:153-219):
// dir > 0: A(i+1) = A(i), A(0) = 0;  dir < 0: A(i-1) = A(i), A(N1-1) = 0.
template <typename R>
__global__ void k_shift_x(cplx<R>* __restrict__ A, int N1, long plane, int dir)
{
    long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= plane) return;
    if (dir > 0) {
        for (int i = N1 - 1; i >= 1; i--) A[(long)i * plane + m] = A[(long)(i - 1) * plane + m];
        A[m] = czero<R>();
    } else {
        for (int i = 0; i < N1 - 1; i++) A[(long)i * plane + m] = A[(long)(i + 1) * plane + m];
        A[(long)(N1 - 1) * plane + m] = czero<R>();
    }
}

template <typename R>
int shift_x(cplx<R>* data, const Geo& g, int dir, hipStream_t st)
{
    long plane = (long)g.N2 * g.H;
    hipLaunchKernelGGL((k_shift_x<R>), dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, st, data, g.N1, plane, dir);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

// ------------------------------------------------------------------------------------------
// Element-wise spectral ops (four_tr.cc:314-353, scft_tr.cc:649-707, four_basic.cc:220-337,
// universal_misc.cc:118-159, array_basic.cc:62-71)
// ------------------------------------------------------------------------------------------
struct GeoK {
    int basis, N1, N2, N3, H, dealias;
    double f1, f2, f3;
};
static GeoK gk(const Geo& g) { return GeoK{g.basis, g.N1, g.N2, g.N3, g.H, g.dealias, g.kf[0], g.kf[1], g.kf[2]}; }

__device__ __forceinline__ void decode(const GeoK& g, long i, int& l1, int& l2, int& l3)
{
    l3 = (int)(i % g.H);
    long r = i / g.H;
    l2 = (int)(r % g.N2);
    l1 = (int)(r / g.N2);
}
__device__ __forceinline__ int kx_g(const GeoK& g, int l1) { return g.basis == 1 ? l1 : kx_four(l1, g.N1); }

__device__ __forceinline__ bool kept_g(const GeoK& g, int l1, int l2, int l3)
{
    return kept_mode(g.basis, g.dealias, g.N1, g.N2, g.N3, l1, l2, l3);
}

__device__ __forceinline__ double ksqr_g(const GeoK& g, int l1, int l2, int l3)
{
#pragma clang fp contract(off)
    double a = kx_g(g, l1) * g.f1, b = ky_of(l2, g.N2) * g.f2, c = l3 * g.f3;
    return a * a + b * b + c * c;
}

__device__ __forceinline__ double kmag_g(const GeoK& g, int l1, int l2, int l3) { return sqrt(ksqr_g(g, l1, l2, l3)); }

#define GRID_LOOP(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)(n); i += (long)gridDim.x * blockDim.x)

static dim3 ew_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return dim3((unsigned)(b ? b : 1));
}

template <typename R, bool ADD>
__global__ void k_deriv_kernel(GeoK g, int axis, int parity, const cplx<R>* __restrict__ A, cplx<R>* __restrict__ B)
{
    long n = (long)g.N1 * g.N2 * g.H;
    GRID_LOOP(i, n) {
        int l1, l2, l3;
        decode(g, i, l1, l2, l3);
        cplx<R> a = A[i];
        double ar = (double)a.x, ai = (double)a.y, br, bi;
        if (axis == 1 && g.basis == 1) {
            double fk = g.f1 * (1.0 * l1);
            if (parity == 0) fk = -fk;
            br = fk * ar;
            bi = fk * ai;
        } else {
            int k = (axis == 1) ? kx_four(l1, g.N1) : (axis == 2) ? ky_of(l2, g.N2) : l3;
            double f = (axis == 1) ? g.f1 : (axis == 2) ? g.f2 : g.f3;
            double fk = f * (1.0 * k);
            br = -fk * ai;
            bi = fk * ar;
        }
        if (ADD) {
            cplx<R> b = B[i];
            B[i] = mk<R>((R)((double)b.x + br), (R)((double)b.y + bi));
        } else {
            B[i] = mk<R>((R)br, (R)bi);
        }
    }
}

template <typename R>
int k_deriv(const Geo& g, int axis, int parity, const cplx<R>* A, cplx<R>* B, hipStream_t st)
{
    hipLaunchKernelGGL((k_deriv_kernel<R, false>), ew_grid(g.ncplx()), dim3(256), 0, st, gk(g), axis, parity, A, B);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
template <typename R>
int k_add_deriv(const Geo& g, int axis, int parity, const cplx<R>* A, cplx<R>* B, hipStream_t st)
{
    hipLaunchKernelGGL((k_deriv_kernel<R, true>), ew_grid(g.ncplx()), dim3(256), 0, st, gk(g), axis, parity, A, B);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_dealias_kernel(GeoK g, cplx<R>* __restrict__ A)
{
    long n = (long)g.N1 * g.N2 * g.H;
    GRID_LOOP(i, n) {
        int l1, l2, l3;
        decode(g, i, l1, l2, l3);
        if (!kept_g(g, l1, l2, l3)) A[i] = czero<R>();
    }
}
template <typename R>
int k_dealias(const Geo& g, cplx<R>* A, hipStream_t st)
{
    if (!g.dealias) return OK;
    hipLaunchKernelGGL((k_dealias_kernel<R>), ew_grid(g.ncplx()), dim3(256), 0, st, gk(g), A);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_real_mult_kernel(long n, const cplx<R>* __restrict__ A, const cplx<R>* __restrict__ B, cplx<R>* __restrict__ C)
{
    GRID_LOOP(i, n) {
        cplx<R> a = A[i], b = B[i];
        C[i] = mk<R>(a.x * b.x, a.y * b.y);
    }
}
template <typename R>
int k_real_mult(size_t n, const cplx<R>* A, const cplx<R>* B, cplx<R>* C, hipStream_t st)
{
    hipLaunchKernelGGL((k_real_mult_kernel<R>), ew_grid(n), dim3(256), 0, st, (long)n, A, B, C);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

// UB_to_Elsasser_field / Elsasser_to_UB_field elsasser.cc:53-75: (U, B) -> (U + B, U - B) and back with 1/2
template <typename R>
__global__ void k_elsasser_kernel(long n, cplx<R>* __restrict__ U, cplx<R>* __restrict__ B, int back)
{
    GRID_LOOP(i, n) {
        const cplx<R> u = U[i], b = B[i];
        cplx<R> p = mk<R>(u.x + b.x, u.y + b.y), m = mk<R>(u.x - b.x, u.y - b.y);
        if (back) {
            p = mk<R>(R(0.5) * p.x, R(0.5) * p.y);
            m = mk<R>(R(0.5) * m.x, R(0.5) * m.y);
        }
        U[i] = p;
        B[i] = m;
    }
}
template <typename R>
int k_elsasser(size_t n, cplx<R>* U, cplx<R>* B, int back, hipStream_t st)
{
    hipLaunchKernelGGL((k_elsasser_kernel<R>), ew_grid(n), dim3(256), 0, st, (long)n, U, B, back);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_divide_ksqr_kernel(GeoK g, cplx<R>* __restrict__ A)
{
    long n = (long)g.N1 * g.N2 * g.H;
    GRID_LOOP(i, n) {
        int l1, l2, l3;
        decode(g, i, l1, l2, l3);
        if (i == 0) { A[0] = czero<R>(); continue; }
        double k2 = ksqr_g(g, l1, l2, l3);
        cplx<R> a = A[i];
        A[i] = mk<R>((R)((double)a.x / k2), (R)((double)a.y / k2));
    }
}
template <typename R>
int k_divide_ksqr(const Geo& g, cplx<R>* A, hipStream_t st)
{
    hipLaunchKernelGGL((k_divide_ksqr_kernel<R>), ew_grid(g.ncplx()), dim3(256), 0, st, gk(g), A);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_exp_ksqr_kernel(GeoK g, cplx<R>* __restrict__ A, double factor)
{
    long n = (long)g.N1 * g.N2 * g.H;
    GRID_LOOP(i, n) {
        int l1, l2, l3;
        decode(g, i, l1, l2, l3);
        double e = exp(factor * ksqr_g(g, l1, l2, l3));
        cplx<R> a = A[i];
        A[i] = mk<R>((R)((double)a.x * e), (R)((double)a.y * e));
    }
}
template <typename R>
int k_exp_ksqr(const Geo& g, cplx<R>* A, double factor, hipStream_t st)
{
    hipLaunchKernelGGL((k_exp_ksqr_kernel<R>), ew_grid(g.ncplx()), dim3(256), 0, st, gk(g), A, factor);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_axpy_kernel(long n, cplx<R>* __restrict__ Y, double a, const cplx<R>* __restrict__ X)
{
    GRID_LOOP(i, n) {
        cplx<R> y = Y[i], x = X[i];
        Y[i] = mk<R>((R)((double)y.x + a * (double)x.x), (R)((double)y.y + a * (double)x.y));
    }
}
template <typename R>
int k_axpy(size_t n, cplx<R>* Y, double a, const cplx<R>* X, hipStream_t st)
{
    hipLaunchKernelGGL((k_axpy_kernel<R>), ew_grid(n), dim3(256), 0, st, (long)n, Y, a, X);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_axpby_kernel(long n, cplx<R>* __restrict__ Y, double a, const cplx<R>* __restrict__ X, double b)
{
    GRID_LOOP(i, n) {
        cplx<R> y = Y[i], x = X[i];
        Y[i] = mk<R>((R)(a * (double)x.x + b * (double)y.x), (R)(a * (double)x.y + b * (double)y.y));
    }
}
template <typename R>
int k_axpby(size_t n, cplx<R>* Y, double a, const cplx<R>* X, double b, hipStream_t st)
{
    hipLaunchKernelGGL((k_axpby_kernel<R>), ew_grid(n), dim3(256), 0, st, (long)n, Y, a, X, b);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_scale_kernel(long n, cplx<R>* __restrict__ Y, double a)
{
    GRID_LOOP(i, n) {
        cplx<R> y = Y[i];
        Y[i] = mk<R>((R)(a * (double)y.x), (R)(a * (double)y.y));
    }
}
template <typename R>
int k_scale(size_t n, cplx<R>* Y, double a, hipStream_t st)
{
    hipLaunchKernelGGL((k_scale_kernel<R>), ew_grid(n), dim3(256), 0, st, (long)n, Y, a);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

// Fill_array_shell (trunk/lib/fields/universal/universal_ET.cc:45-65): B = A on inner < |K| <= outer
template <typename R>
__global__ void k_fill_shell_kernel(GeoK g, const cplx<R>* __restrict__ A, cplx<R>* __restrict__ B, double inner, double outer)
{
    long n = (long)g.N1 * g.N2 * g.H;
    GRID_LOOP(i, n) {
        int l1, l2, l3;
        decode(g, i, l1, l2, l3);
        double km = kmag_g(g, l1, l2, l3);
        B[i] = ((km > inner) && (km <= outer)) ? A[i] : czero<R>();
    }
}
template <typename R>
int k_fill_shell(const Geo& g, const cplx<R>* A, cplx<R>* B, double inner, double outer, hipStream_t st)
{
    hipLaunchKernelGGL((k_fill_shell_kernel<R>), ew_grid(g.ncplx()), dim3(256), 0, st, gk(g), A, B, inner, outer);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_zero_padding_kernel(GeoK g, cplx<R>* __restrict__ A)
{
    long n = (long)g.N1 * g.N2;
    GRID_LOOP(i, n) A[i * g.H + g.N3 / 2] = czero<R>();
}
template <typename R>
int k_zero_padding(const Geo& g, cplx<R>* A, hipStream_t st)
{
    hipLaunchKernelGGL((k_zero_padding_kernel<R>), ew_grid((size_t)g.N1 * g.N2), dim3(256), 0, st, gk(g), A);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

template <typename R>
__global__ void k_zero_x0_kernel(long plane, cplx<R>* __restrict__ A)
{
    GRID_LOOP(i, plane) A[i] = czero<R>();
}
template <typename R>
int k_zero_plane_x0(const Geo& g, cplx<R>* A, hipStream_t st)
{
    long plane = (long)g.N2 * g.H;
    hipLaunchKernelGGL((k_zero_x0_kernel<R>), ew_grid(plane), dim3(256), 0, st, plane, A);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

#define TH_INST(R)                                                                                                   \
    template int fft_strided<R>(cplx<R>*, long, long, int, int, int, int, R, const cplx<R>*, hipStream_t);           \
    template int fft_z_r2c<R>(cplx<R>*, long, int, R, const cplx<R>*, const cplx<R>*, hipStream_t);                 \
    template int fft_z_c2r<R>(cplx<R>*, long, int, const cplx<R>*, const cplx<R>*, hipStream_t);                    \
    template int r2r_x<R>(cplx<R>*, const Geo&, int, cplx<R>*, const cplx<R>*, const cplx<R>*, hipStream_t);        \
    template int shift_x<R>(cplx<R>*, const Geo&, int, hipStream_t);                                                 \
    template int k_deriv<R>(const Geo&, int, int, const cplx<R>*, cplx<R>*, hipStream_t);                           \
    template int k_add_deriv<R>(const Geo&, int, int, const cplx<R>*, cplx<R>*, hipStream_t);                       \
    template int k_dealias<R>(const Geo&, cplx<R>*, hipStream_t);                                                    \
    template int k_real_mult<R>(size_t, const cplx<R>*, const cplx<R>*, cplx<R>*, hipStream_t);                     \
    template int k_elsasser<R>(size_t, cplx<R>*, cplx<R>*, int, hipStream_t);                                         \
    template int k_divide_ksqr<R>(const Geo&, cplx<R>*, hipStream_t);                                                \
    template int k_exp_ksqr<R>(const Geo&, cplx<R>*, double, hipStream_t);                                           \
    template int k_axpy<R>(size_t, cplx<R>*, double, const cplx<R>*, hipStream_t);                                   \
    template int k_axpby<R>(size_t, cplx<R>*, double, const cplx<R>*, double, hipStream_t);                          \
    template int k_scale<R>(size_t, cplx<R>*, double, hipStream_t);                                                  \
    template int k_fill_shell<R>(const Geo&, const cplx<R>*, cplx<R>*, double, double, hipStream_t);                 \
    template int k_zero_plane_x0<R>(const Geo&, cplx<R>*, hipStream_t);                                              \
    template int k_zero_padding<R>(const Geo&, cplx<R>*, hipStream_t);

TH_INST(float)
TH_INST(double)

}  // namespace th
