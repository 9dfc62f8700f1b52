// th_ic.hip -- initial conditions on reference-layout device fields.
//   Taylor-Green: IncVF::Setup_Taylor_Green_field trunk/lib/IncFlow/misc/ic_model.cc:58-107.
//   Random-phase spectrum: the reference seeds Blitz's generator with time(0) (src/main.cc:100), so
//   fields are re-created from a counter-based generator keyed by (seed, component, l1, l2, l3):
//   a splitmix64 chain gives two 53-bit uniforms, Box-Muller gives a complex Gaussian; then the
//   projection P(k), k = 0 zeroed, Satisfy_reality_array (universal_misc.cc:51-109), Dealias_array,
//   and each unit shell k = ceil|K| rescaled to E(k) = C k^4 exp(-2 (k/kp)^2), sum_k E(k) = E0.
#include <math.h>
#include <stdint.h>
#include <vector>
#include "../../include/tarang_hip.h"
#include "th_internal.hpp"
#include "th_plan.hpp"

using namespace th;

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ void gauss_pair(uint64_t seed, int comp, int l1, int l2, int l3, double& re, double& im)
{
    uint64_t h = splitmix64(seed);
    h = splitmix64(h ^ (uint64_t)(unsigned)comp);
    h = splitmix64(h ^ (uint64_t)(unsigned)l1);
    h = splitmix64(h ^ (uint64_t)(unsigned)l2);
    h = splitmix64(h ^ (uint64_t)(unsigned)l3);
    uint64_t h2 = splitmix64(h);
    double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(u1));
    double th = 2.0 * M_PI * u2;
    re = r * cos(th);
    im = r * sin(th);
}

struct GI {
    int basis, N1, N2, N3, H;
    double f1, f2, f3;
};

__device__ __forceinline__ void dec(const GI& g, long i, int& l1, int& l2, int& l3)
{
    l3 = (int)(i % g.H);
    long r = i / g.H;
    l2 = (int)(r % g.N2);
    l1 = (int)(r / g.N2);
}
__device__ __forceinline__ int kxg(const GI& g, int l1) { return g.basis == 1 ? l1 : kx_four(l1, g.N1); }

template <typename R>
__global__ void k_ic_gauss(GI g, cplx<R>* V0, cplx<R>* V1, cplx<R>* V2, uint64_t seed, int project)
{
    long n = (long)g.N1 * g.N2 * g.H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int l1, l2, l3;
        dec(g, i, l1, l2, l3);
        double re[3], im[3];
        for (int c = 0; c < 3; c++) gauss_pair(seed, c, l1, l2, l3, re[c], im[c]);
        if (project) {
            double K[3] = {kxg(g, l1) * g.f1, ky_of(l2, g.N2) * g.f2, l3 * g.f3};
            double k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
            if (k2 != 0) {
                double dr = K[0] * re[0] + K[1] * re[1] + K[2] * re[2];
                double di = K[0] * im[0] + K[1] * im[1] + K[2] * im[2];
                for (int c = 0; c < 3; c++) {
                    re[c] = re[c] - K[c] * dr / k2;
                    im[c] = im[c] - K[c] * di / k2;
                }
            }
        }
        if (i == 0) re[0] = re[1] = re[2] = im[0] = im[1] = im[2] = 0.0;
        V0[i] = mk<R>((R)re[0], (R)im[0]);
        V1[i] = mk<R>((R)re[1], (R)im[1]);
        V2[i] = mk<R>((R)re[2], (R)im[2]);
    }
}

template <typename R>
__global__ void k_ic_gauss_scalar(GI g, cplx<R>* T, uint64_t seed, double amp)
{
    long n = (long)g.N1 * g.N2 * g.H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int l1, l2, l3;
        dec(g, i, l1, l2, l3);
        double re, im;
        gauss_pair(seed, 7, l1, l2, l3, re, im);
        T[i] = (l1 == 0) ? czero<R>() : mk<R>((R)(amp * re), (R)(amp * im));
    }
}

// Satisfy_reality_array: the written entries (ly > N2/2; ly = 0 with kx < 0) and the entries they
// are copied from (0 < ky < N2/2; ly = 0 with kx > 0) are disjoint, so one in-place pass suffices.
template <typename R>
__global__ void k_reality(GI g, cplx<R>* A)
{
    long n = 2L * g.N1 * g.N2;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        int z = (q & 1) ? g.N3 / 2 : 0;
        long r = q >> 1;
        int ly = (int)(r % g.N2), lx = (int)(r / g.N2);
        auto at = [&](int a, int b) -> cplx<R>& { return A[((long)a * g.N2 + b) * g.H + z]; };
        if (g.basis == 0) {
            int kx = kx_four(lx, g.N1);
            if (ly > g.N2 / 2) {
                int cnx = (-kx >= 0) ? -kx : -kx + g.N1;
                int piy = -ky_of(ly, g.N2);
                at(lx, ly) = conjg(at(cnx, piy));
            } else if (ly == 0 && kx < 0) {
                at(lx, 0) = conjg(at(-kx, 0));
            }
        } else if (ly > g.N2 / 2) {
            at(lx, ly) = conjg(at(lx, -ky_of(ly, g.N2)));
        }
    }
}

template <typename R>
__global__ void k_shell_scale(GI g, cplx<R>* V0, cplx<R>* V1, cplx<R>* V2, const double* __restrict__ scale, int kkmax)
{
    long n = (long)g.N1 * g.N2 * g.H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int l1, l2, l3;
        dec(g, i, l1, l2, l3);
        double a = kxg(g, l1) * g.f1, b = ky_of(l2, g.N2) * g.f2, c = l3 * g.f3;
        double km;
        {
#pragma clang fp contract(off)
            km = sqrt(a * a + b * b + c * c);
        }
        int idx = (int)ceil(km);
        double s = (idx <= kkmax) ? scale[idx] : 0.0;
        cplx<R>* V[3] = {V0, V1, V2};
        for (int k = 0; k < 3; k++) {
            cplx<R> v = V[k][i];
            V[k][i] = mk<R>((R)((double)v.x * s), (R)((double)v.y * s));
        }
    }
}

template <typename R>
__global__ void k_set_modes(cplx<R>* V0, cplx<R>* V1, int n, const long* __restrict__ idx, const int* __restrict__ comp,
                            const double* __restrict__ val)
{
    int t = threadIdx.x;
    if (t < n) {
        cplx<R>* A = comp[t] == 0 ? V0 : V1;
        A[idx[t]] = mk<R>((R)val[2 * t], (R)val[2 * t + 1]);
    }
}

GI gi(const Geo& g) { return GI{g.basis, g.N1, g.N2, g.N3, g.H, g.kf[0], g.kf[1], g.kf[2]}; }

dim3 grid_of(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return dim3((unsigned)(b ? b : 1));
}

template <typename R> int ic_random(th_plan* p, void* Vv[3], uint64_t seed, double E0, double kp)
{
    const Geo& g = p->g;
    hipStream_t st = p->ctx->stream;
    cplx<R>* V[3] = {(cplx<R>*)Vv[0], (cplx<R>*)Vv[1], (cplx<R>*)Vv[2]};
    size_t n = g.ncplx();
    hipLaunchKernelGGL((k_ic_gauss<R>), grid_of(n), dim3(256), 0, st, gi(g), V[0], V[1], V[2], seed, 1);
    TH_CHECK_HIP(hipGetLastError());
    for (int c = 0; c < 3; c++) {
        hipLaunchKernelGGL((k_reality<R>), grid_of(2 * (size_t)g.N1 * g.N2), dim3(256), 0, st, gi(g), V[c]);
        TH_CHECK_HIP(hipGetLastError());
        TRY(k_dealias<R>(g, V[c], st));
    }
    int kkmax = min_radius_outside(g);
    int nb = kkmax + 1;
    std::vector<double> Ek(nb, 0.0), tmp(nb);
    for (int c = 0; c < 3; c++) {
        TRY(reduce_shell<R>(g, V[c], nullptr, 0.0, nb, tmp.data(), nullptr, p->rs, st, false));
        for (int k = 0; k < nb; k++) Ek[k] += tmp[k];
    }
    double norm = 0;
    for (int k = 1; k <= kkmax; k++)
        if (Ek[k] > 0) norm += pow((double)k, 4.0) * exp(-2.0 * (k / kp) * (k / kp));
    double C = E0 / norm;
    std::vector<double> scale(nb, 0.0);
    for (int k = 1; k <= kkmax; k++)
        if (Ek[k] > 0) scale[k] = sqrt(C * pow((double)k, 4.0) * exp(-2.0 * (k / kp) * (k / kp)) / Ek[k]);
    double* dscale = nullptr;
    TH_CHECK_HIP(hipMalloc(&dscale, nb * sizeof(double)));
    TH_CHECK_HIP(hipMemcpy(dscale, scale.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL((k_shell_scale<R>), grid_of(n), dim3(256), 0, st, gi(g), V[0], V[1], V[2], (const double*)dscale, kkmax);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(st);
    (void)hipFree(dscale);
    if (e != hipSuccess) { set_error("k_shell_scale: %s", hipGetErrorString(e)); return E_HIP; }
    return OK;
}

template <typename R> int ic_scalar(th_plan* p, void* Tv, uint64_t seed, double amp)
{
    const Geo& g = p->g;
    hipStream_t st = p->ctx->stream;
    cplx<R>* T = (cplx<R>*)Tv;
    hipLaunchKernelGGL((k_ic_gauss_scalar<R>), grid_of(g.ncplx()), dim3(256), 0, st, gi(g), T, seed, amp);
    TH_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_reality<R>), grid_of(2 * (size_t)g.N1 * g.N2), dim3(256), 0, st, gi(g), T);
    TH_CHECK_HIP(hipGetLastError());
    TRY(k_dealias<R>(g, T, st));
    return OK;
}

template <typename R> int ic_tg(th_plan* p, void* Vv[3], int k0, double amp)
{
    const Geo& g = p->g;
    hipStream_t st = p->ctx->stream;
    size_t bytes = g.ncplx() * sizeof(cplx<R>);
    for (int c = 0; c < 3; c++) TH_CHECK_HIP(hipMemsetAsync(Vv[c], 0, bytes, st));
    auto id = [&](int a, int b, int c) { return ((long)a * g.N2 + b) * g.H + c; };
    int lxp = k0, lxm = (g.basis == 0) ? g.N1 - k0 : k0, lyp = k0, lym = g.N2 - k0;
    double a = amp / 8;
    std::vector<long> idx;
    std::vector<int> comp;
    std::vector<double> val;
    auto put = [&](int c, long i, double re, double im) { comp.push_back(c); idx.push_back(i); val.push_back(re); val.push_back(im); };
    if (g.basis == 0) {
        put(0, id(lxp, lyp, k0), 0, -a); put(0, id(lxp, lym, k0), 0, -a);
        put(1, id(lxp, lyp, k0), 0, a);  put(1, id(lxp, lym, k0), 0, -a);
        put(0, id(lxm, lyp, k0), 0, a);  put(0, id(lxm, lym, k0), 0, a);
        put(1, id(lxm, lyp, k0), 0, a);  put(1, id(lxm, lym, k0), 0, -a);
    } else {
        put(0, id(k0, lyp, k0), a, 0); put(0, id(k0, lym, k0), a, 0);
        put(1, id(k0, lyp, k0), 0, a); put(1, id(k0, lym, k0), 0, -a);
    }
    int m = (int)idx.size();
    long* didx;
    int* dcomp;
    double* dval;
    TH_CHECK_HIP(hipMalloc(&didx, m * sizeof(long)));
    TH_CHECK_HIP(hipMalloc(&dcomp, m * sizeof(int)));
    TH_CHECK_HIP(hipMalloc(&dval, 2 * m * sizeof(double)));
    TH_CHECK_HIP(hipMemcpy(didx, idx.data(), m * sizeof(long), hipMemcpyHostToDevice));
    TH_CHECK_HIP(hipMemcpy(dcomp, comp.data(), m * sizeof(int), hipMemcpyHostToDevice));
    TH_CHECK_HIP(hipMemcpy(dval, val.data(), 2 * m * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL((k_set_modes<R>), dim3(1), dim3(64), 0, st, (cplx<R>*)Vv[0], (cplx<R>*)Vv[1], m, (const long*)didx,
                       (const int*)dcomp, (const double*)dval);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(st);
    (void)hipFree(didx);
    (void)hipFree(dcomp);
    (void)hipFree(dval);
    if (e != hipSuccess) { set_error("k_set_modes: %s", hipGetErrorString(e)); return E_HIP; }
    return OK;
}

}  // namespace

extern "C" int th_ic_taylor_green(th_plan* p, void* V[3], int k0, double amp)
{
    if (!p || !V) { set_error("th_ic_taylor_green: null"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    if (k0 < 1 || k0 >= p->g.N1 / 2 || k0 >= p->g.N2 / 2 || k0 > p->g.N3 / 2) { set_error("th_ic_taylor_green: k0 %d", k0); return E_ARG; }
    return p->prec ? ic_tg<float>(p, V, k0, amp) : ic_tg<double>(p, V, k0, amp);
}
extern "C" int th_ic_random(th_plan* p, void* V[3], unsigned long long seed, double E0, double kp)
{
    if (!p || !V) { set_error("th_ic_random: null"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? ic_random<float>(p, V, seed, E0, kp) : ic_random<double>(p, V, seed, E0, kp);
}
extern "C" int th_ic_random_scalar(th_plan* p, void* T, unsigned long long seed, double amp)
{
    if (!p || !T) { set_error("th_ic_random_scalar: null"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? ic_scalar<float>(p, T, seed, amp) : ic_scalar<double>(p, T, seed, amp);
}

template <typename R> static int reality(th_plan* p, void* A)
{
    hipLaunchKernelGGL((k_reality<R>), grid_of(2 * (size_t)p->g.N1 * p->g.N2), dim3(256), 0, p->ctx->stream, gi(p->g),
                       (cplx<R>*)A);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}
extern "C" int th_satisfy_reality(th_plan* p, void* A)
{
    if (!p || !A) { set_error("th_satisfy_reality: null"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? reality<float>(p, A) : reality<double>(p, A);
}
