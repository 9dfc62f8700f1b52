// th_probe.hip -- measured comparison of the two ways to run the 16-point DFTs of the line FFTs'
// radix-16 stages on gfx950.  No counterpart in the reference (FFTW plans its own codelets,
// four_tr.cc:44-98); this is the measurement behind the choice of the vector-ALU codelet.
//
//   method 0 (VALU): DFT<16> of th_fft.hpp, one 16-point vector per lane: 4 x DFT4, 9 twiddle
//            products, 4 x DFT4 (about 160 FP32 VALU instructions per vector).
//   method 1 (MFMA): the DFT as a real 32 x 32 matrix [[Wr, -Wi], [Wi, Wr]] times 32 vectors
//            (rows: 16 real parts, then 16 imaginary parts) on the FP32 matrix cores,
//            16 v_mfma_f32_32x32x2f32 per 32 vectors (4096 flops each).
//
// Vectors are stored point-major, x[j * nvec + v] = point j of vector v, so both kernels read and
// write whole 64-byte segments.  `reps` applies the DFT repeatedly in registers (compute-bound
// timing); the result is (W / 4)^reps x, the 4^-reps applied once at the end (exact in binary).
#include <math.h>
#include "th_internal.hpp"
#include "th_fft.hpp"

namespace th {

typedef float v16f __attribute__((ext_vector_type(16)));

template <int SIGN>
__global__ __launch_bounds__(256) void k_dft16_valu(const cplx<float>* __restrict__ x, cplx<float>* __restrict__ y,
                                                    long nvec, int reps, float scale)
{
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    if (v >= nvec) return;
    cplx<float> a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = x[j * nvec + v];
#pragma unroll 1
    for (int r = 0; r < reps; r++) DFT<16, SIGN, float>::run(a);
#pragma unroll
    for (int j = 0; j < 16; j++) y[j * nvec + v] = cscale(a[j], scale);
}

// Lane l of a wave works on vector l % 32 of the wave's 32 and half h = l / 32.  The 32 x 32 x 2
// MFMA takes A[i][k] from lane i + 32 k, B[k][j] from lane j + 32 k, and returns D[i][j] in
// register q of lane j + 32 h for i = 8 (q / 4) + 4 h + q % 4.  The reduction index k of
// instruction kb is ordered so that lane (j, h) feeds row mrow(h, kb) of vector j: the rows it holds
// in D, so one transform's result is the next one's B operand without data movement.
__device__ __forceinline__ int mrow(int h, int q) { return 8 * (q / 4) + 4 * h + q % 4; }

template <int SIGN>
__global__ __launch_bounds__(256) void k_dft16_mfma(const cplx<float>* __restrict__ x, cplx<float>* __restrict__ y,
                                                    long nvec, int reps, float scale)
{
    const int lane = threadIdx.x % 64, j = lane % 32, h = lane / 32;
    const long v = ((long)blockIdx.x * 4 + threadIdx.x / 64) * 32 + j;
    // A operand of instruction kb: M[j][mrow(h, kb)], M = [[Wr, -Wi], [Wi, Wr]], W[k][n] = exp(SIGN 2 pi i k n / 16)
    float am[16];
#pragma unroll
    for (int kb = 0; kb < 16; kb++) {
        const int c = mrow(h, kb), r = j;
        const int k = r % 16, n = c % 16;
        const double ang = (double)SIGN * (double)((k * n) % 16) / 8.0;   // in units of pi
        const float wr = (float)cospi(ang), wi = (float)sinpi(ang);
        am[kb] = (r < 16) ? (c < 16 ? wr : -wi) : (c < 16 ? wi : wr);
    }
    // points p(h, q) = 8 (q / 4) + 4 h + q % 4, q < 8: real parts in registers q, imaginary in 8 + q
    v16f b;
    const bool ok = v < nvec;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const cplx<float> e = ok ? x[(long)mrow(h, q) * nvec + v] : czero<float>();
        b[q] = e.x;
        b[8 + q] = e.y;
    }
#pragma unroll 1
    for (int r = 0; r < reps; r++) {
        v16f d = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 16; kb++) d = __builtin_amdgcn_mfma_f32_32x32x2f32(am[kb], b[kb], d, 0, 0, 0);
        b = d;
    }
    if (!ok) return;
#pragma unroll
    for (int q = 0; q < 8; q++) y[(long)mrow(h, q) * nvec + v] = mk<float>(b[q] * scale, b[8 + q] * scale);
}

}  // namespace th

using namespace th;

extern "C" int th_dft16_probe(th_ctx* ctx, int method, int sign, const void* x, void* y, long nvec, int reps,
                              int ntimed, double* ms)
{
    if (!ctx || !x || !y || nvec <= 0 || nvec % 256 != 0 || reps < 1 || reps > 32 || ntimed < 0 ||
        (ntimed > 0 && !ms) || (method != 0 && method != 1) || (sign != 1 && sign != -1)) {
        set_error("th_dft16_probe: bad arguments (nvec a positive multiple of 256, 1 <= reps <= 32, method 0|1, sign +-1)");
        return E_ARG;
    }
    TH_CHECK_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const float scale = ldexpf(1.0f, -2 * reps);
    const cplx<float>* xi = static_cast<const cplx<float>*>(x);
    cplx<float>* yo = static_cast<cplx<float>*>(y);
    auto launch = [&]() -> int {
        const dim3 blk(256), grd((unsigned)(nvec / 256));
        if (method == 0) {
            if (sign < 0) hipLaunchKernelGGL(k_dft16_valu<-1>, grd, blk, 0, st, xi, yo, nvec, reps, scale);
            else hipLaunchKernelGGL(k_dft16_valu<+1>, grd, blk, 0, st, xi, yo, nvec, reps, scale);
        } else {
            const dim3 grm((unsigned)(nvec / 128));   // 4 waves x 32 vectors per workgroup
            if (sign < 0) hipLaunchKernelGGL(k_dft16_mfma<-1>, grm, blk, 0, st, xi, yo, nvec, reps, scale);
            else hipLaunchKernelGGL(k_dft16_mfma<+1>, grm, blk, 0, st, xi, yo, nvec, reps, scale);
        }
        TH_CHECK_HIP(hipGetLastError());
        return OK;
    };
    TRY(launch());
    if (ntimed == 0) return OK;
    hipEvent_t e0, e1;
    TH_CHECK_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); set_error("hipEventCreate"); return E_HIP; }
    int r = OK;
    if (hipEventRecord(e0, st) != hipSuccess) r = E_HIP;
    for (int i = 0; i < ntimed && r == OK; i++) r = launch();
    float t = 0;
    if (r == OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        r = E_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r == E_HIP) set_error("th_dft16_probe: event timing failed");
    if (r == OK) *ms = (double)t / ntimed;
    return r;
}
