// th_api.hip -- the C-ABI (include/tarang_hip.h): contexts, plans, the boundary transforms and
// spectral operators, the plain GPU step (the reference's operation sequence), diagnostics and ICs.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <new>
#include <vector>
#include "../../include/tarang_hip.h"
#include "th_internal.hpp"
#include "th_plan.hpp"

using namespace th;

static thread_local char g_err[1024] = "";

void th::set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char* th_last_error(void) { return g_err; }
extern "C" const char* th_version(void) { return "tarang_hip 0.1 (gfx950)"; }

// ------------------------------------------------------------------------------------------
// twiddles: W_L^m = exp(-2 pi i m/L) from 80-bit libm, rounded once to the storage precision
// ------------------------------------------------------------------------------------------
const void* TwiddleCache::get(int L, int prec, hipStream_t s)
{
    auto& mp = prec ? f32 : f64;
    auto it = mp.find(L);
    if (it != mp.end()) return it->second;
    std::vector<double> t(2 * (size_t)L);
    const long double PI_L = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < L; m++) {
        long double a = 2.0L * PI_L * (long double)m / (long double)L;
        double c = (double)cosl(a), sn = (double)sinl(a);
        if ((4LL * m) % L == 0) {
            long long q = (4LL * m) / L;
            c = (q == 0) ? 1.0 : (q == 2) ? -1.0 : 0.0;
            sn = (q == 1) ? 1.0 : (q == 3) ? -1.0 : 0.0;
        }
        t[2 * m] = c;
        t[2 * m + 1] = -sn;
    }
    void* d = nullptr;
    if (prec) {
        std::vector<float> f(t.begin(), t.end());
        if (hipMalloc(&d, f.size() * sizeof(float)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    } else {
        if (hipMalloc(&d, t.size() * sizeof(double)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    }
    (void)s;
    mp[L] = d;
    return d;
}
void TwiddleCache::release()
{
    for (auto& kv : f32) (void)hipFree(kv.second);
    for (auto& kv : f64) (void)hipFree(kv.second);
    f32.clear();
    f64.clear();
}

// ------------------------------------------------------------------------------------------
// context / plan
// ------------------------------------------------------------------------------------------
namespace th {
static void tune_load(th_ctx* c)
{
    if (c->tune_path.empty()) return;
    FILE* f = fopen(c->tune_path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char key[256];
        int n = 0, off = 0;
        if (line[0] == '#' || sscanf(line, "%255s%n", key, &off) != 1) continue;
        std::vector<int> v;
        const char* q = line + off;
        int x;
        while (sscanf(q, "%d%n", &x, &n) == 1) {
            v.push_back(x);
            q += n;
        }
        if (!v.empty()) c->tune[key] = v;
    }
    fclose(f);
}
bool tune_lookup(th_ctx* c, const std::string& key, std::vector<int>& v)
{
    std::lock_guard<std::mutex> lk(c->tune_mu);
    auto it = c->tune.find(key);
    if (it == c->tune.end()) return false;
    v = it->second;
    return true;
}
void tune_store(th_ctx* c, const std::string& key, const std::vector<int>& v)
{
    std::lock_guard<std::mutex> lk(c->tune_mu);
    c->tune[key] = v;
    if (c->tune_path.empty()) return;
    std::string tmp = c->tune_path + ".tmp";
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return;
    fprintf(f, "# tarang_hip launch-shape cache: key, then one variant per pass\n");
    for (auto& kv : c->tune) {
        fprintf(f, "%s", kv.first.c_str());
        for (int x : kv.second) fprintf(f, " %d", x);
        fprintf(f, "\n");
    }
    fclose(f);
    (void)rename(tmp.c_str(), c->tune_path.c_str());
}
}  // namespace th

extern "C" int th_ctx_create(th_ctx** ctx, int device, const char* tune_cache_path)
{
    if (!ctx) { set_error("th_ctx_create: null"); return E_ARG; }
    int n = 0;
    TH_CHECK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) { set_error("th_ctx_create: device %d of %d", device, n); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(device));
    th_ctx* c = new (std::nothrow) th_ctx();
    if (!c) return E_NOMEM;
    c->device = device;
    if (tune_cache_path) c->tune_path = tune_cache_path;
    tune_load(c);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; set_error("hipStreamCreate: %s", hipGetErrorString(e)); return E_HIP; }
    *ctx = c;
    return OK;
}

extern "C" int th_ctx_destroy(th_ctx* ctx)
{
    if (!ctx) return OK;
    if (ctx->nplans > 0) { set_error("th_ctx_destroy: %d plans still alive", (int)ctx->nplans); return E_STATE; }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return OK;
}

static bool supported_len(int n)
{
    const int ok[] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};
    for (int v : ok)
        if (v == n) return true;
    return false;
}

extern "C" int th_plan_create(th_ctx* ctx, th_plan** plan, int basis, int N1, int N2, int N3, const double kfactor[3],
                              int precision, int dealias)
{
    if (!ctx || !plan) { set_error("th_plan_create: null"); return E_ARG; }
    if (basis != TH_FOUR && basis != TH_SCFT) { set_error("th_plan_create: basis %d", basis); return E_ARG; }
    if (precision != TH_FP64 && precision != TH_FP32_MIXED) { set_error("th_plan_create: precision %d", precision); return E_ARG; }
    if (!supported_len(N1) || !supported_len(N2) || N3 % 2 || !supported_len(N3 / 2) || !supported_len(N3) ||
        N1 < 4 || N2 < 4 || N3 < 4) {
        set_error("th_plan_create: unsupported grid %dx%dx%d (lengths 2^a 3^b, 4..512; N3 even)", N1, N2, N3);
        return E_UNSUPPORTED;
    }
    if (basis == TH_SCFT && N1 > 256 && false) return E_UNSUPPORTED;
    th_plan* p = new (std::nothrow) th_plan();
    if (!p) return E_NOMEM;
    p->ctx = ctx;
    p->g.basis = basis;
    p->g.N1 = N1;
    p->g.N2 = N2;
    p->g.N3 = N3;
    p->g.H = N3 / 2 + 1;
    for (int i = 0; i < 3; i++) p->g.kf[i] = kfactor ? kfactor[i] : 1.0;
    p->g.dealias = dealias ? 1 : 0;
    p->prec = precision;
    ctx->nplans++;
    *plan = p;
    return OK;
}

th_plan::~th_plan()
{
    for (void* s : scratch)
        if (s) (void)hipFree(s);
    tw.release();
    rs.release();
    if (radii_dev) (void)hipFree(radii_dev);
    release_fused(this);
}

void th::plan_free(th_plan* p)
{
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    th_ctx* c = p->ctx;
    delete p;
    c->nplans--;
}

// A plan that still has fused states is released by the destruction of its last state, so states
// and plans may be destroyed in either order.
extern "C" int th_plan_destroy(th_plan* p)
{
    if (!p || p->dead) return OK;
    if (p->nstates > 0) {
        p->dead = true;
        return OK;
    }
    th::plan_free(p);
    return OK;
}

extern "C" size_t th_plan_field_elems(const th_plan* p) { return p ? p->g.ncplx() : 0; }
extern "C" size_t th_plan_elem_bytes(const th_plan* p) { return p ? (p->prec ? 8 : 16) : 0; }
extern "C" int th_shell_ek_size(const th_plan* p) { return p ? min_radius_outside(p->g) + 1 : 0; }
extern "C" int th_max_radius_inside(const th_plan* p) { return p ? max_radius_inside(p->g) : 0; }
extern "C" int th_sync(th_plan* p)
{
    if (!p) return E_ARG;
    TH_CHECK_HIP(hipStreamSynchronize(p->ctx->stream));
    return OK;
}

void* th_plan::scr(int i)
{
    if ((int)scratch.size() <= i) scratch.resize(i + 1, nullptr);
    if (!scratch[i]) {
        if (hipMalloc(&scratch[i], g.ncplx() * (prec ? 8 : 16)) != hipSuccess) {
            scratch[i] = nullptr;
            set_error("scratch allocation of %zu bytes failed", g.ncplx() * (size_t)(prec ? 8 : 16));
        }
    }
    return scratch[i];
}

template <typename R> const cplx<R>* th_plan::twd(int L)
{
    return (const cplx<R>*)tw.get(L, sizeof(R) == 4 ? 1 : 0, ctx->stream);
}
template const cplx<float>* th_plan::twd<float>(int);
template const cplx<double>* th_plan::twd<double>(int);

// ------------------------------------------------------------------------------------------
// basis transforms
// ------------------------------------------------------------------------------------------
namespace th {

template <typename R> int plain_forward(th_plan* p, cplx<R>* A, int parity)
{
    const Geo& g = p->g;
    hipStream_t st = p->ctx->stream;
    const long plane = (long)g.N2 * g.H;
    if (g.basis == 0) {
        TRY(fft_z_r2c<R>(A, (long)g.N1 * g.N2, g.N3, R(1), p->twd<R>(g.N3 / 2), p->twd<R>(g.N3), st));
        TRY(fft_strided<R>(A, 0, plane, 1, g.N1, (int)plane, -1, R(1), p->twd<R>(g.N1), st));
        const R norm = (R)(1.0 / ((double)g.N1 * (double)g.N2 * (double)g.N3));
        TRY(fft_strided<R>(A, plane, g.H, g.N1, g.N2, g.H, -1, norm, p->twd<R>(g.N2), st));
        return OK;
    }
    TRY(r2r_x<R>(A, g, parity ? 1 : 0, nullptr, p->twd<R>(2 * g.N1), p->twd<R>(4 * g.N1), st));
    TRY(k_zero_padding<R>(g, A, st));
    TRY(fft_z_r2c<R>(A, (long)g.N1 * g.N2, g.N3, R(1), p->twd<R>(g.N3 / 2), p->twd<R>(g.N3), st));
    const R norm = (R)(1.0 / (2.0 * (double)g.N1 * (double)g.N2 * (double)g.N3));
    TRY(fft_strided<R>(A, plane, g.H, g.N1, g.N2, g.H, -1, norm, p->twd<R>(g.N2), st));
    if (parity) TRY(shift_x<R>(A, g, +1, st));
    return OK;
}

template <typename R> int plain_inverse(th_plan* p, cplx<R>* A, int parity)
{
    const Geo& g = p->g;
    hipStream_t st = p->ctx->stream;
    const long plane = (long)g.N2 * g.H;
    if (g.basis == 0) {
        TRY(fft_strided<R>(A, plane, g.H, g.N1, g.N2, g.H, +1, R(1), p->twd<R>(g.N2), st));
        TRY(fft_strided<R>(A, 0, plane, 1, g.N1, (int)plane, +1, R(1), p->twd<R>(g.N1), st));
        TRY(fft_z_c2r<R>(A, (long)g.N1 * g.N2, g.N3, p->twd<R>(g.N3 / 2), p->twd<R>(g.N3), st));
        return OK;
    }
    if (parity) TRY(shift_x<R>(A, g, -1, st));
    TRY(fft_strided<R>(A, plane, g.H, g.N1, g.N2, g.H, +1, R(1), p->twd<R>(g.N2), st));
    TRY(fft_z_c2r<R>(A, (long)g.N1 * g.N2, g.N3, p->twd<R>(g.N3 / 2), p->twd<R>(g.N3), st));
    TRY(r2r_x<R>(A, g, parity ? 3 : 2, nullptr, p->twd<R>(2 * g.N1), p->twd<R>(4 * g.N1), st));
    TRY(k_zero_padding<R>(g, A, st));
    return OK;
}

template int plain_forward<float>(th_plan*, cplx<float>*, int);
template int plain_forward<double>(th_plan*, cplx<double>*, int);
template int plain_inverse<float>(th_plan*, cplx<float>*, int);
template int plain_inverse<double>(th_plan*, cplx<double>*, int);

}  // namespace th

#define DISPATCH(p, fn, ...) ((p)->prec ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))
#define CP(R, x) ((cplx<R>*)(x))
#define CCP(R, x) ((const cplx<R>*)(x))

template <typename R> static int do_forward(th_plan* p, void* A, int parity) { return plain_forward<R>(p, CP(R, A), parity); }
template <typename R> static int do_inverse(th_plan* p, void* A, int parity) { return plain_inverse<R>(p, CP(R, A), parity); }

static int check_plan(th_plan* p)
{
    if (!p) { set_error("null plan"); return E_ARG; }
    hipError_t e = hipSetDevice(p->ctx->device);
    if (e != hipSuccess) { set_error("hipSetDevice: %s", hipGetErrorString(e)); return E_HIP; }
    return OK;
}

extern "C" int th_forward(th_plan* p, void* A, int parity)
{
    TRY(check_plan(p));
    if (!A) { set_error("th_forward: null array"); return E_ARG; }
    return DISPATCH(p, do_forward, p, A, parity);
}
extern "C" int th_inverse(th_plan* p, void* A, int parity)
{
    TRY(check_plan(p));
    if (!A) { set_error("th_inverse: null array"); return E_ARG; }
    return DISPATCH(p, do_inverse, p, A, parity);
}

template <typename R> static int do_deriv(th_plan* p, int axis, const void* A, void* B, int parity)
{
    return k_deriv<R>(p->g, axis, parity, CCP(R, A), CP(R, B), p->ctx->stream);
}
extern "C" int th_deriv(th_plan* p, int axis, const void* A, void* B, int parity)
{
    TRY(check_plan(p));
    if (axis < 1 || axis > 3 || !A || !B) { set_error("th_deriv: bad arguments"); return E_ARG; }
    return DISPATCH(p, do_deriv, p, axis, A, B, parity);
}
template <typename R> static int do_dealias(th_plan* p, void* A) { return k_dealias<R>(p->g, CP(R, A), p->ctx->stream); }
extern "C" int th_dealias(th_plan* p, void* A)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_dealias, p, A);
}
template <typename R> static int do_rmult(th_plan* p, const void* A, const void* B, void* C)
{
    return k_real_mult<R>(p->g.ncplx(), CCP(R, A), CCP(R, B), CP(R, C), p->ctx->stream);
}
extern "C" int th_real_mult(th_plan* p, const void* A, const void* B, void* C)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_rmult, p, A, B, C);
}
template <typename R> static int do_divk(th_plan* p, void* A) { return k_divide_ksqr<R>(p->g, CP(R, A), p->ctx->stream); }
extern "C" int th_divide_ksqr(th_plan* p, void* A)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_divk, p, A);
}
template <typename R> static int do_expk(th_plan* p, void* A, double f) { return k_exp_ksqr<R>(p->g, CP(R, A), f, p->ctx->stream); }
extern "C" int th_exp_ksqr(th_plan* p, void* A, double factor)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_expk, p, A, factor);
}

// ------------------------------------------------------------------------------------------
// plain nonlinear terms (reference operation sequence, see th_plan.hpp for scratch slots)
// ------------------------------------------------------------------------------------------
namespace th {

template <typename R> struct Ops {
    th_plan* p;
    hipStream_t st;
    const Geo& g;
    size_t n;
    explicit Ops(th_plan* pp) : p(pp), st(pp->ctx->stream), g(pp->g), n(pp->g.ncplx()) {}
    cplx<R>* S(int i) { return (cplx<R>*)p->scr(i); }
    int copy(cplx<R>* d, const cplx<R>* s)
    {
        TH_CHECK_HIP(hipMemcpyAsync(d, s, n * sizeof(cplx<R>), hipMemcpyDeviceToDevice, st));
        return OK;
    }
    int fwd(cplx<R>* A, int par) { return plain_forward<R>(p, A, par); }
    int inv(cplx<R>* A, int par) { return plain_inverse<R>(p, A, par); }
    int rmul(const cplx<R>* A, const cplx<R>* B, cplx<R>* C) { return k_real_mult<R>(n, A, B, C, st); }
    int der(int ax, const cplx<R>* A, cplx<R>* B, int par) { return k_deriv<R>(g, ax, par, A, B, st); }
    int add_der(cplx<R>* B, int ax, const cplx<R>* A, int par) { return k_add_deriv<R>(g, ax, par, A, B, st); }
    int dealias(cplx<R>* A) { return k_dealias<R>(g, A, st); }
    int axpy(cplx<R>* Y, double a, const cplx<R>* X) { return k_axpy<R>(n, Y, a, X, st); }
    int axpby(cplx<R>* Y, double a, const cplx<R>* X, double b) { return k_axpby<R>(n, Y, a, X, b, st); }
    int scale(cplx<R>* Y, double a) { return k_scale<R>(n, Y, a, st); }
    int expk(cplx<R>* A, double f) { return k_exp_ksqr<R>(g, A, f, st); }

    // IncVF::Compute_nlin() compute_nlin.cc:58-108
    int nlin_ns(cplx<R>* const V[3], cplx<R>* nl[3])
    {
        cplx<R>* Vr[3] = {S(SC_VR0), S(SC_VR0 + 1), S(SC_VR0 + 2)};
        cplx<R>* VFt = S(SC_VFT);
        if (!Vr[0] || !Vr[1] || !Vr[2] || !VFt) return E_NOMEM;
        for (int i = 0; i < 3; i++) TRY(copy(Vr[i], V[i]));
        TRY(inv(Vr[0], 1));
        TRY(inv(Vr[1], 0));
        TRY(inv(Vr[2], 0));
        for (int i = 0; i < 3; i++) TRY(rmul(Vr[i], Vr[i], nl[i]));
        for (int i = 0; i < 3; i++) TRY(fwd(nl[i], 0));
        TRY(der(1, nl[0], nl[0], 0));
        TRY(der(2, nl[1], nl[1], 0));
        TRY(der(3, nl[2], nl[2], 0));
        TRY(copy(VFt, Vr[0]));
        TRY(rmul(Vr[0], Vr[1], Vr[0]));
        TRY(rmul(Vr[1], Vr[2], Vr[1]));
        TRY(rmul(Vr[2], VFt, Vr[2]));
        TRY(fwd(Vr[0], 1));
        TRY(fwd(Vr[1], 0));
        TRY(fwd(Vr[2], 1));
        TRY(add_der(nl[0], 2, Vr[0], 0));
        TRY(add_der(nl[0], 3, Vr[2], 0));
        TRY(add_der(nl[1], 1, Vr[0], 1));
        TRY(add_der(nl[1], 3, Vr[1], 0));
        TRY(add_der(nl[2], 1, Vr[2], 1));
        TRY(add_der(nl[2], 2, Vr[1], 0));
        for (int i = 0; i < 3; i++) TRY(dealias(nl[i]));
        return OK;
    }

    // IncVF::Compute_pressure() compute_pressure.cc:54-62
    int pressure(cplx<R>* const nl[3], cplx<R>* pr)
    {
        TRY(der(1, nl[0], pr, 1));
        TRY(add_der(pr, 2, nl[1], 0));
        TRY(add_der(pr, 3, nl[2], 0));
        TRY(k_divide_ksqr<R>(g, pr, st));
        TRY(dealias(pr));
        return OK;
    }

    // IncFluid::Compute_rhs() compute_rhs.cc:52-60
    int rhs(cplx<R>* nl[3], const cplx<R>* pr)
    {
        TRY(add_der(nl[0], 1, pr, 0));
        TRY(add_der(nl[1], 2, pr, 0));
        TRY(add_der(nl[2], 3, pr, 0));
        for (int i = 0; i < 3; i++) TRY(scale(nl[i], -1.0));
        return OK;
    }

    // IncVF::Compute_nlin(IncVF&) compute_nlin_VF_MHD.cc:72-164
    int nlin_mhd(cplx<R>* U[3], cplx<R>* B[3], cplx<R>* nU[3], cplx<R>* nB[3])
    {
        cplx<R>* Vr[3] = {S(SC_VR0), S(SC_VR0 + 1), S(SC_VR0 + 2)};
        cplx<R>* Wr[3] = {S(SC_WR0), S(SC_WR0 + 1), S(SC_WR0 + 2)};
        cplx<R>* VFt = S(SC_VFT);
        cplx<R>* WFt = S(SC_WFT);
        for (int i = 0; i < 3; i++) {
            TRY(copy(VFt, U[i]));
            TRY(axpy(U[i], 1.0, B[i]));            // U = U + B
            TRY(axpby(B[i], 1.0, VFt, -1.0));      // B = VFt - B
        }
        for (int i = 0; i < 3; i++) { TRY(copy(Vr[i], U[i])); TRY(copy(Wr[i], B[i])); }
        TRY(inv(Vr[0], 1)); TRY(inv(Vr[1], 0)); TRY(inv(Vr[2], 0));
        TRY(inv(Wr[0], 1)); TRY(inv(Wr[1], 0)); TRY(inv(Wr[2], 0));
        for (int i = 0; i < 3; i++) TRY(rmul(Vr[i], Wr[i], nU[i]));
        for (int i = 0; i < 3; i++) TRY(fwd(nU[i], 0));
        TRY(der(1, nU[0], nU[0], 0));
        TRY(der(2, nU[1], nU[1], 0));
        TRY(der(3, nU[2], nU[2], 0));
        for (int i = 0; i < 3; i++) TRY(copy(nB[i], nU[i]));
        TRY(copy(VFt, Vr[0]));
        TRY(copy(WFt, Wr[0]));
        TRY(rmul(Vr[0], Wr[1], Vr[0]));
        TRY(rmul(Wr[0], Vr[1], Wr[0]));
        TRY(rmul(Vr[1], Wr[2], Vr[1]));
        TRY(rmul(Wr[1], Vr[2], Wr[1]));
        TRY(rmul(Vr[2], WFt, Vr[2]));
        TRY(rmul(Wr[2], VFt, Wr[2]));
        TRY(fwd(Vr[0], 1)); TRY(fwd(Vr[1], 0)); TRY(fwd(Vr[2], 1));
        TRY(fwd(Wr[0], 1)); TRY(fwd(Wr[1], 0)); TRY(fwd(Wr[2], 1));
        TRY(add_der(nU[0], 2, Vr[0], 0));
        TRY(add_der(nU[0], 3, Wr[2], 0));
        TRY(add_der(nU[1], 1, Wr[0], 1));
        TRY(add_der(nU[1], 3, Vr[1], 0));
        TRY(add_der(nU[2], 1, Vr[2], 1));
        TRY(add_der(nU[2], 2, Wr[1], 0));
        TRY(add_der(nB[0], 2, Wr[0], 0));
        TRY(add_der(nB[0], 3, Vr[2], 0));
        TRY(add_der(nB[1], 1, Vr[0], 1));
        TRY(add_der(nB[1], 3, Wr[1], 0));
        TRY(add_der(nB[2], 1, Wr[2], 1));
        TRY(add_der(nB[2], 2, Vr[1], 0));
        for (int i = 0; i < 3; i++) {
            // Elsasser_to_UB_nlin / Elsasser_to_UB_field: x = 0.5 (x + y), y = 0.5 (x_old - y)
            TRY(copy(VFt, nU[i]));
            TRY(axpby(nU[i], 0.5, nB[i], 0.5));
            TRY(axpby(nB[i], 0.5, VFt, -0.5));
            TRY(copy(VFt, U[i]));
            TRY(axpby(U[i], 0.5, B[i], 0.5));
            TRY(axpby(B[i], 0.5, VFt, -0.5));
        }
        for (int i = 0; i < 3; i++) { TRY(dealias(nU[i])); TRY(dealias(nB[i])); }
        return OK;
    }

    // IncVF::Compute_nlin_scalar(IncSF&) compute_nlin_SF.cc:80-177
    int nlin_rb(cplx<R>* const V[3], const cplx<R>* T, cplx<R>* nV[3], cplx<R>* nT)
    {
        cplx<R>* Vr[3] = {S(SC_VR0), S(SC_VR0 + 1), S(SC_VR0 + 2)};
        cplx<R>* Tr = S(SC_TR);
        cplx<R>* SFt = S(SC_SFT);
        cplx<R>* VFt = S(SC_VFT);
        for (int i = 0; i < 3; i++) TRY(copy(Vr[i], V[i]));
        TRY(copy(Tr, T));
        TRY(inv(Vr[0], 1)); TRY(inv(Vr[1], 0)); TRY(inv(Vr[2], 0));
        TRY(inv(Tr, 1));
        TRY(rmul(Vr[0], Tr, nT));
        TRY(rmul(Vr[1], Tr, SFt));
        TRY(fwd(nT, 0));
        TRY(fwd(SFt, 1));
        TRY(der(1, nT, nT, 0));
        TRY(der(2, SFt, SFt, 0));
        TRY(axpy(nT, 1.0, SFt));
        TRY(rmul(Vr[2], Tr, SFt));
        TRY(fwd(SFt, 1));
        TRY(der(3, SFt, SFt, 0));
        TRY(axpy(nT, 1.0, SFt));
        for (int i = 0; i < 3; i++) TRY(rmul(Vr[i], Vr[i], nV[i]));
        for (int i = 0; i < 3; i++) TRY(fwd(nV[i], 0));
        TRY(der(1, nV[0], nV[0], 0));
        TRY(der(2, nV[1], nV[1], 0));
        TRY(der(3, nV[2], nV[2], 0));
        TRY(copy(VFt, Vr[0]));
        TRY(rmul(Vr[0], Vr[1], Vr[0]));
        TRY(rmul(Vr[1], Vr[2], Vr[1]));
        TRY(rmul(Vr[2], VFt, Vr[2]));
        TRY(fwd(Vr[0], 1)); TRY(fwd(Vr[1], 0)); TRY(fwd(Vr[2], 1));
        TRY(add_der(nV[0], 2, Vr[0], 0));
        TRY(add_der(nV[0], 3, Vr[2], 0));
        TRY(add_der(nV[1], 1, Vr[0], 1));
        TRY(add_der(nV[1], 3, Vr[1], 0));
        TRY(add_der(nV[2], 1, Vr[2], 1));
        TRY(add_der(nV[2], 2, Vr[1], 0));
        for (int i = 0; i < 3; i++) TRY(dealias(nV[i]));
        TRY(dealias(nT));
        return OK;
    }

    // IncVF::EnergyTr_Compute_nlin() ET_compute_nlin.cc:85-120 (helper H, giver G; not dealiased)
    int et_nlin(cplx<R>* const Hh[3], cplx<R>* const G[3], cplx<R>* nl[3])
    {
        cplx<R>* Vr[3] = {S(SC_VR0), S(SC_VR0 + 1), S(SC_VR0 + 2)};
        cplx<R>* Gr[3] = {S(SC_WR0), S(SC_WR0 + 1), S(SC_WR0 + 2)};
        cplx<R>* VFt = S(SC_VFT);
        cplx<R>* VFt2 = S(SC_WFT);
        for (int i = 0; i < 3; i++) { TRY(copy(Vr[i], Hh[i])); TRY(copy(Gr[i], G[i])); }
        TRY(inv(Vr[0], 1)); TRY(inv(Vr[1], 0)); TRY(inv(Vr[2], 0));
        TRY(inv(Gr[0], 1)); TRY(inv(Gr[1], 0)); TRY(inv(Gr[2], 0));
        for (int i = 0; i < 3; i++) TRY(rmul(Vr[i], Gr[i], nl[i]));
        for (int i = 0; i < 3; i++) TRY(fwd(nl[i], 0));
        TRY(der(1, nl[0], nl[0], 0));
        TRY(der(2, nl[1], nl[1], 0));
        TRY(der(3, nl[2], nl[2], 0));
        TRY(copy(VFt, Vr[0]));
        TRY(copy(VFt2, Gr[0]));
        TRY(rmul(Vr[0], Gr[1], Vr[0]));
        TRY(rmul(Gr[0], Vr[1], Gr[0]));
        TRY(rmul(Vr[1], Gr[2], Vr[1]));
        TRY(rmul(Gr[1], Vr[2], Gr[1]));
        TRY(rmul(Vr[2], VFt2, Vr[2]));
        TRY(rmul(Gr[2], VFt, Gr[2]));
        TRY(fwd(Vr[0], 1)); TRY(fwd(Vr[1], 0)); TRY(fwd(Vr[2], 1));
        TRY(fwd(Gr[0], 1)); TRY(fwd(Gr[1], 0)); TRY(fwd(Gr[2], 1));
        TRY(add_der(nl[0], 2, Gr[0], 0));
        TRY(add_der(nl[0], 3, Vr[2], 0));
        TRY(add_der(nl[1], 1, Vr[0], 1));
        TRY(add_der(nl[1], 3, Gr[1], 0));
        TRY(add_der(nl[2], 1, Gr[2], 1));
        TRY(add_der(nl[2], 2, Vr[1], 0));
        return OK;
    }

    // IncVF::EnergyTr_Compute_nlin(IncSF& T) ET_compute_nlin.cc:190-234 (helper H, scalar giver G of
    // sine parity): nl = D_x F(H1 G) (cos parity) + D_y F(H2 G) + D_z F(H3 G) (sin parity); not dealiased
    int et_nlin_scalar(cplx<R>* const Hh[3], const cplx<R>* G, cplx<R>* nl)
    {
        cplx<R>* Vr[3] = {S(SC_VR0), S(SC_VR0 + 1), S(SC_VR0 + 2)};
        cplx<R>* Gr = S(SC_TR);
        cplx<R>* tmp = S(SC_SFT);
        for (int i = 0; i < 3; i++) TRY(copy(Vr[i], Hh[i]));
        TRY(copy(Gr, G));
        TRY(inv(Vr[0], 1)); TRY(inv(Vr[1], 0)); TRY(inv(Vr[2], 0));
        TRY(inv(Gr, 1));
        TRY(rmul(Vr[0], Gr, nl));
        TRY(fwd(nl, 0));
        TRY(der(1, nl, nl, 0));
        TRY(rmul(Vr[1], Gr, tmp));
        TRY(fwd(tmp, 1));
        TRY(der(2, tmp, tmp, 0));
        TRY(axpy(nl, 1.0, tmp));
        TRY(rmul(Vr[2], Gr, tmp));
        TRY(fwd(tmp, 1));
        TRY(der(3, tmp, tmp, 0));
        TRY(axpy(nl, 1.0, tmp));
        return OK;
    }

    // ---------------- steppers (time_advance.cc, single_time_step.cc) ----------------
    int nc = 0;
    cplx<R>* F[7];
    cplx<R>* N[7];
    double diss[7];
    int physics = 0;
    double Ra = 0, Pr = 0;
    const ForceBand* force = nullptr;    // NS forcing (force_field_proc 1 / 2), else none
    const ForceBand* force_w = nullptr;  // MHD: the band of W's supply (force of U in `force`)
    cplx<R>* keep[6] = {};   // receives each forced stage's fields (NS: V, MHD: U then B; the last stays)

    int compute_rhs()
    {
        cplx<R>* pr = S(SC_P);
        if (physics == TH_NS) {
            TRY(nlin_ns(F, N));
            // Compute_force_TO_rhs single_time_step.cc:351-364: the force of the stage field (which
            // Compute_nlin leaves unchanged) is subtracted from nlin before the pressure
            if (force) TRY(force_subtract<R>(*force, F, N, st));
            if (force && keep[0])
                for (int c = 0; c < 3; c++) TRY(copy(keep[c], F[c]));
        } else if (physics == TH_MHD) {
            // Compute_force(IncVF& W) precedes Compute_nlin, which round-trips U and B through the Elsasser
            // variables: the forces are taken from U and B first and subtracted afterwards (Add_force(W))
            cplx<R>* FU[3] = {S(SC_FU0), S(SC_FU0 + 1), S(SC_FU0 + 2)};
            cplx<R>* FW[3] = {S(SC_FW0), S(SC_FW0 + 1), S(SC_FW0 + 2)};
            const bool fu = force && (force->proc == 3 || force->per_mode > 1e-15),
                       fw = force_w && (force_w->proc == 3 || force_w->per_mode > 1e-15);
            if ((fu || fw) && (!FU[0] || !FU[1] || !FU[2] || !FW[0] || !FW[1] || !FW[2])) return E_NOMEM;
            if (fu) TRY(force_write<R>(*force, F, FU, n, st));
            if (fw) TRY(force_write<R>(*force_w, F + 3, FW, n, st));
            if (keep[0])
                for (int c = 0; c < 6; c++) TRY(copy(keep[c], F[c]));
            TRY(nlin_mhd(F, F + 3, N, N + 3));
            for (int c = 0; c < 3; c++) {
                if (fu) TRY(axpy(N[c], -1.0, FU[c]));
                if (fw) TRY(axpy(N[3 + c], -1.0, FW[c]));
            }
        } else {
            // Compute_force_RB compute_force_RB.cc:51-105 (PRLARGE, ULARGE): Force1 = T, T.Force = V1
            cplx<R>* F1 = S(SC_F1);
            cplx<R>* FT = S(SC_FT);
            TRY(copy(F1, F[3]));
            TRY(copy(FT, F[0]));
            TRY(dealias(F1));
            TRY(dealias(FT));
            TRY(nlin_rb(F, F[3], N, N[3]));
            TRY(axpy(N[0], -1.0, F1));           // Add_force add_force.cc:57-63
            TRY(axpy(N[3], -1.0, FT));
        }
        TRY(pressure(N, pr));
        TRY(rhs(N, pr));
        if (physics == TH_MHD)
            for (int i = 3; i < 6; i++) TRY(scale(N[i], -1.0));
        if (physics == TH_RB) TRY(scale(N[3], -1.0));
        return OK;
    }

    int mult_exp(cplx<R>* A, double d, double dt, double a)
    {
        if (fabs(a) > 1e-15) return expk(A, -d * a * dt);
        return OK;
    }

    int single_time_step(double dt, double a, double b, double c)
    {
        const double EPS = 1e-15;
        for (int q = 0; q < nc; q++) {
            cplx<R>*V = F[q], *Nq = N[q];
            double d = diss[q];
            if ((fabs(a) < EPS) && (fabs(b) < EPS)) {
                TRY(axpy(V, c * dt, Nq));
            } else if ((fabs(a) < EPS) && (fabs(b) > EPS)) {
                TRY(mult_exp(Nq, d, b, dt));
                TRY(axpy(V, c * dt, Nq));
            } else if ((fabs(b) < EPS) && (fabs(a) > EPS)) {
                TRY(mult_exp(V, d, a, dt));
                TRY(axpy(V, c * dt, Nq));
            } else if ((fabs(a - b) < EPS) && (fabs(a) > EPS)) {
                TRY(axpy(V, c * dt, Nq));
                TRY(mult_exp(V, d, a, dt));
            } else if ((fabs(a) > EPS) && (fabs(b) > EPS)) {
                TRY(mult_exp(V, d, (a - b), dt));
                TRY(axpy(V, c * dt, Nq));
                TRY(mult_exp(V, d, b, dt));
            }
        }
        return OK;
    }

    int rk4_parts(cplx<R>** tot, double dt, double b, double factor)
    {
        for (int q = 0; q < nc; q++) {
            TRY(mult_exp(N[q], diss[q], dt, b));
            TRY(axpy(tot[q], factor, N[q]));
        }
        return OK;
    }

    int step(int scheme, double Tdt)
    {
        cplx<R>* copyf[7];
        cplx<R>* tot[7];
        for (int q = 0; q < nc; q++) {
            copyf[q] = S(SC_COPY0 + q);
            tot[q] = S(SC_TOT0 + q);
            if (!copyf[q] || !tot[q]) return E_NOMEM;
        }
        if (scheme == 0) {                                  // EULER
            TRY(compute_rhs());
            TRY(single_time_step(Tdt, 1, 1, 1));
        } else if (scheme == 1) {                           // RK2
            for (int q = 0; q < nc; q++) TRY(copy(copyf[q], F[q]));
            TRY(compute_rhs());
            TRY(single_time_step(Tdt, 0.5, 0.5, 0.5));
            TRY(compute_rhs());
            for (int q = 0; q < nc; q++) TRY(copy(F[q], copyf[q]));
            TRY(single_time_step(Tdt, 1, 0.5, 1));
        } else {                                            // RK4 time_advance.cc:100-155
            for (int q = 0; q < nc; q++) {
                TRY(copy(copyf[q], F[q]));
                TH_CHECK_HIP(hipMemsetAsync(tot[q], 0, n * sizeof(cplx<R>), st));
            }
            TRY(compute_rhs());
            TRY(single_time_step(Tdt, 0.5, 0.5, 0.5));
            TRY(rk4_parts(tot, Tdt, 1.0, Tdt / 6));
            TRY(compute_rhs());
            for (int q = 0; q < nc; q++) TRY(copy(F[q], copyf[q]));
            TRY(single_time_step(Tdt, 0.5, 0, 0.5));
            TRY(rk4_parts(tot, Tdt, 0.5, Tdt / 3));
            TRY(compute_rhs());
            for (int q = 0; q < nc; q++) TRY(copy(F[q], copyf[q]));
            TRY(single_time_step(Tdt, 1, 0.5, 1));
            TRY(rk4_parts(tot, Tdt, 0.5, Tdt / 3));
            TRY(compute_rhs());
            TRY(rk4_parts(tot, Tdt, 0, Tdt / 6));
            for (int q = 0; q < nc; q++) {
                TRY(copy(F[q], copyf[q]));
                TRY(mult_exp(F[q], diss[q], Tdt, 1.0));
                TRY(axpy(F[q], 1.0, tot[q]));
            }
        }
        for (int q = 0; q < nc; q++) TRY(dealias(F[q]));
        return OK;
    }
};

}  // namespace th

template <typename R> static int do_nlin_ns(th_plan* p, void* const V[3], void* nl[3])
{
    Ops<R> o(p);
    cplx<R>* v[3] = {CP(R, V[0]), CP(R, V[1]), CP(R, V[2])};
    cplx<R>* n[3] = {CP(R, nl[0]), CP(R, nl[1]), CP(R, nl[2])};
    return o.nlin_ns(v, n);
}
extern "C" int th_nlin_ns(th_plan* p, void* const V[3], void* nlin[3])
{
    TRY(check_plan(p));
    return DISPATCH(p, do_nlin_ns, p, V, nlin);
}
template <typename R> static int do_pressure(th_plan* p, void* const nl[3], void* pr)
{
    Ops<R> o(p);
    cplx<R>* n[3] = {CP(R, nl[0]), CP(R, nl[1]), CP(R, nl[2])};
    return o.pressure(n, CP(R, pr));
}
extern "C" int th_pressure(th_plan* p, void* const nlin[3], void* pr)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_pressure, p, nlin, pr);
}
template <typename R> static int do_rhs(th_plan* p, void* nl[3], const void* pr)
{
    Ops<R> o(p);
    cplx<R>* n[3] = {CP(R, nl[0]), CP(R, nl[1]), CP(R, nl[2])};
    return o.rhs(n, CCP(R, pr));
}
extern "C" int th_rhs(th_plan* p, void* nlin[3], const void* pr)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_rhs, p, nlin, pr);
}
template <typename R> static int do_nlin_mhd(th_plan* p, void* U[3], void* B[3], void* nU[3], void* nB[3])
{
    Ops<R> o(p);
    cplx<R>* u[3] = {CP(R, U[0]), CP(R, U[1]), CP(R, U[2])};
    cplx<R>* b[3] = {CP(R, B[0]), CP(R, B[1]), CP(R, B[2])};
    cplx<R>* nu[3] = {CP(R, nU[0]), CP(R, nU[1]), CP(R, nU[2])};
    cplx<R>* nb[3] = {CP(R, nB[0]), CP(R, nB[1]), CP(R, nB[2])};
    return o.nlin_mhd(u, b, nu, nb);
}
extern "C" int th_nlin_mhd(th_plan* p, void* U[3], void* B[3], void* nU[3], void* nB[3])
{
    TRY(check_plan(p));
    return DISPATCH(p, do_nlin_mhd, p, U, B, nU, nB);
}
template <typename R> static int do_nlin_rb(th_plan* p, void* const V[3], const void* T, void* nV[3], void* nT)
{
    Ops<R> o(p);
    cplx<R>* v[3] = {CP(R, V[0]), CP(R, V[1]), CP(R, V[2])};
    cplx<R>* nv[3] = {CP(R, nV[0]), CP(R, nV[1]), CP(R, nV[2])};
    return o.nlin_rb(v, CCP(R, T), nv, CP(R, nT));
}
extern "C" int th_nlin_rb(th_plan* p, void* const V[3], const void* T, void* nV[3], void* nT)
{
    TRY(check_plan(p));
    if (p->g.basis != TH_SCFT) { set_error("th_nlin_rb: needs an SCFT plan"); return E_ARG; }
    return DISPATCH(p, do_nlin_rb, p, V, T, nV, nT);
}
template <typename R> static int do_et_nlin(th_plan* p, void* const H[3], void* const G[3], void* nl[3])
{
    Ops<R> o(p);
    cplx<R>* h[3] = {CP(R, H[0]), CP(R, H[1]), CP(R, H[2])};
    cplx<R>* gg[3] = {CP(R, G[0]), CP(R, G[1]), CP(R, G[2])};
    cplx<R>* n[3] = {CP(R, nl[0]), CP(R, nl[1]), CP(R, nl[2])};
    return o.et_nlin(h, gg, n);
}
extern "C" int th_et_nlin(th_plan* p, void* const H[3], void* const G[3], void* nlin[3])
{
    TRY(check_plan(p));
    return DISPATCH(p, do_et_nlin, p, H, G, nlin);
}

template <typename R> static int do_step_plain(th_plan* p, int physics, int scheme, void* const fields[], double dt,
                                               const th_phys* ph)
{
    Ops<R> o(p);
    o.physics = physics;
    o.nc = (physics == TH_NS) ? 3 : (physics == TH_MHD) ? 6 : 4;
    double nu = ph->nu, eta = ph->eta, kap = ph->kappa;
    if (physics == TH_RB && ph->Ra > 0 && ph->Pr > 0) {
        nu = sqrt(ph->Pr / ph->Ra);
        kap = 1 / sqrt(ph->Pr * ph->Ra);
    }
    for (int q = 0; q < o.nc; q++) {
        o.F[q] = CP(R, fields[q]);
        o.N[q] = o.S(SC_N0 + q);
        if (!o.N[q]) return E_NOMEM;
        o.diss[q] = (q < 3) ? nu : (physics == TH_MHD ? eta : kap);
    }
    return o.step(scheme, dt);
}
extern "C" int th_step_plain(th_plan* p, int physics, int scheme, void* const fields[], double dt, const th_phys* ph)
{
    TRY(check_plan(p));
    if (!fields || !ph || physics < 0 || physics > 2 || scheme < 0 || scheme > 2) { set_error("th_step_plain: bad args"); return E_ARG; }
    if (physics == TH_RB && p->g.basis != TH_SCFT) { set_error("th_step_plain: RB needs SCFT"); return E_ARG; }
    return DISPATCH(p, do_step_plain, p, physics, scheme, fields, dt, ph);
}

template <typename R> static int do_step_plain_forced(th_plan* p, int scheme, void* const V[3], double dt,
                                                      const th_phys* ph, const ForceBand* fb, void* const last[3])
{
    Ops<R> o(p);
    o.physics = TH_NS;
    o.nc = 3;
    o.force = fb;
    if (last)
        for (int c = 0; c < 3; c++) o.keep[c] = CP(R, last[c]);
    for (int q = 0; q < 3; q++) {
        o.F[q] = CP(R, V[q]);
        o.N[q] = o.S(SC_N0 + q);
        if (!o.N[q]) return E_NOMEM;
        o.diss[q] = ph->nu;
    }
    return o.step(scheme, dt);
}
template <typename R> static int do_step_plain_forced_mhd(th_plan* p, int scheme, void* const U[3], void* const B[3],
                                                          double dt, const th_phys* ph, const ForceBand* fu,
                                                          const ForceBand* fw, void* const last[6])
{
    Ops<R> o(p);
    o.physics = TH_MHD;
    o.nc = 6;
    o.force = fu;
    o.force_w = fw;
    if (last)
        for (int c = 0; c < 6; c++) o.keep[c] = CP(R, last[c]);
    for (int q = 0; q < 6; q++) {
        o.F[q] = CP(R, q < 3 ? U[q] : B[q - 3]);
        o.N[q] = o.S(SC_N0 + q);
        if (!o.N[q]) return E_NOMEM;
        o.diss[q] = q < 3 ? ph->nu : ph->eta;
    }
    return o.step(scheme, dt);
}
extern "C" int th_step_plain_forced_mhd(th_plan* p, int scheme, void* const U[3], void* const B[3], double dt,
                                        const th_phys* ph, const th_force* force, void* const last_stage[6])
{
    TRY(check_plan(p));
    if (!U || !B || !U[0] || !U[1] || !U[2] || !B[0] || !B[1] || !B[2] || !ph || scheme < 0 || scheme > 2) {
        set_error("th_step_plain_forced_mhd: bad args");
        return E_ARG;
    }
    if (last_stage)
        for (int c = 0; c < 6; c++)
            if (!last_stage[c]) { set_error("th_step_plain_forced_mhd: null last_stage field"); return E_ARG; }
    ForceBand fu, fw;
    TRY(force_setup_mhd(p->g, force, fu, fw));
    return DISPATCH(p, do_step_plain_forced_mhd, p, scheme, U, B, dt, ph, &fu, &fw, last_stage);
}

extern "C" int th_step_plain_forced(th_plan* p, int scheme, void* const V[3], double dt, const th_phys* ph,
                                    const th_force* force, void* const last_stage[3])
{
    TRY(check_plan(p));
    if (!V || !V[0] || !V[1] || !V[2] || !ph || scheme < 0 || scheme > 2 ||
        (last_stage && (!last_stage[0] || !last_stage[1] || !last_stage[2]))) {
        set_error("th_step_plain_forced: bad args");
        return E_ARG;
    }
    ForceBand fb;
    TRY(force_setup(p->g, force, fb));
    return DISPATCH(p, do_step_plain_forced, p, scheme, V, dt, ph, &fb, last_stage);
}

// ------------------------------------------------------------------------------------------
// diagnostics
// ------------------------------------------------------------------------------------------
template <typename R> static int do_energy(th_plan* p, const void* A, const void* B, double* out)
{
    return reduce_energy<R>(p->g, CCP(R, A), CCP(R, B), out, p->rs, p->ctx->stream);
}
extern "C" int th_energy(th_plan* p, const void* A, const void* B, double* out)
{
    TRY(check_plan(p));
    if (!A || !out) { set_error("th_energy: null"); return E_ARG; }
    return DISPATCH(p, do_energy, p, A, B, out);
}
template <typename R> static int do_Sn(th_plan* p, const void* A, double n, double* out)
{
    return reduce_Sn<R>(p->g, CCP(R, A), n, out, p->rs, p->ctx->stream);
}
extern "C" int th_Sn(th_plan* p, const void* A, double n, double* out)
{
    TRY(check_plan(p));
    if (!A || !out || !(n == 0 || n == 2 || n == 4)) { set_error("th_Sn: bad args"); return E_ARG; }
    return DISPATCH(p, do_Sn, p, A, n, out);
}
template <typename R> static int do_spec(th_plan* p, const void* A, const void* B, int n, double* Sk, double* cnt)
{
    return reduce_shell<R>(p->g, CCP(R, A), CCP(R, B), (double)n, min_radius_outside(p->g) + 1, Sk, cnt, p->rs,
                           p->ctx->stream);
}
extern "C" int th_shell_spectrum(th_plan* p, const void* A, const void* B, int n, double* Sk, double* count)
{
    TRY(check_plan(p));
    if (!A || !Sk || !(n == 0 || n == 2 || n == 4)) { set_error("th_shell_spectrum: bad args"); return E_ARG; }
    return DISPATCH(p, do_spec, p, A, B, n, Sk, count);
}
template <typename R> static int do_spec_noalloc(th_plan* p, const void* A, const void* B, int n, double* Sk, double* cnt)
{
    return reduce_shell_noalloc<R>(p->g, CCP(R, A), CCP(R, B), (double)n, min_radius_outside(p->g) + 1, Sk, cnt,
                                   p->ctx->stream);
}
extern "C" int th_shell_spectrum_noalloc(th_plan* p, const void* A, const void* B, int n, double* Sk, double* count)
{
    TRY(check_plan(p));
    if (!A || !Sk || !(n == 0 || n == 2 || n == 4)) { set_error("th_shell_spectrum_noalloc: bad args"); return E_ARG; }
    return DISPATCH(p, do_spec_noalloc, p, A, B, n, Sk, count);
}
template <typename R> static int do_helicity(th_plan* p, void* const V[3], double* out)
{
    const cplx<R>* v[3] = {CCP(R, V[0]), CCP(R, V[1]), CCP(R, V[2])};
    return reduce_helicity<R>(p->g, v, out, p->rs, p->ctx->stream);
}
extern "C" int th_helicity(th_plan* p, void* const V[3], double* out)
{
    TRY(check_plan(p));
    if (!V || !out) { set_error("th_helicity: null"); return E_ARG; }
    return DISPATCH(p, do_helicity, p, V, out);
}
template <typename R> static int do_entropy(th_plan* p, void* const V[3], double* out)
{
    const cplx<R>* v[3] = {CCP(R, V[0]), CCP(R, V[1]), CCP(R, V[2])};
    return reduce_entropy<R>(p->g, v, out, p->rs, p->ctx->stream);
}
extern "C" int th_entropy(th_plan* p, void* const V[3], double* out)
{
    TRY(check_plan(p));
    if (!V || !out) { set_error("th_entropy: null"); return E_ARG; }
    return DISPATCH(p, do_entropy, p, V, out);
}
template <typename R> static int do_entropy_scalar(th_plan* p, const void* A, double* out)
{
    return reduce_entropy_scalar<R>(p->g, CCP(R, A), out, p->rs, p->ctx->stream);
}
extern "C" int th_entropy_scalar(th_plan* p, const void* A, double* out)
{
    TRY(check_plan(p));
    if (!A || !out) { set_error("th_entropy_scalar: null"); return E_ARG; }
    return DISPATCH(p, do_entropy_scalar, p, A, out);
}
template <typename R> static int do_shell_hel(th_plan* p, void* const V[3], double* H1, double* H2, double* H3)
{
    const cplx<R>* v[3] = {CCP(R, V[0]), CCP(R, V[1]), CCP(R, V[2])};
    return reduce_shell_helicity<R>(p->g, v, min_radius_outside(p->g) + 1, H1, H2, H3, p->rs, p->ctx->stream);
}
extern "C" int th_shell_spectrum_helicity(th_plan* p, void* const V[3], double* H1k1, double* H1k2, double* H1k3)
{
    TRY(check_plan(p));
    if (!V || !H1k1 || !H1k2 || !H1k3) { set_error("th_shell_spectrum_helicity: null"); return E_ARG; }
    return DISPATCH(p, do_shell_hel, p, V, H1k1, H1k2, H1k3);
}
template <typename R> static int do_aniso(th_plan* p, void* const V[3], int mode, int nbin, int n, int axis, double* S1,
                                          double* S2)
{
    const cplx<R>* v[3] = {CCP(R, V[0]), CCP(R, V[1]), CCP(R, V[2])};
    return reduce_aniso<R>(p->g, v, mode, nbin, n, axis, S1, S2, p->rs, p->ctx->stream);
}
extern "C" int th_ring_spectrum(th_plan* p, void* const V[3], int nsect, int n, int axis, double* S1k, double* S2k)
{
    TRY(check_plan(p));
    if (!V || !V[0] || !V[1] || !V[2] || !S1k || !S2k) { set_error("th_ring_spectrum: null"); return E_ARG; }
    return DISPATCH(p, do_aniso, p, V, 0, nsect, n, axis, S1k, S2k);
}
extern "C" int th_cylinder_ring_spectrum(th_plan* p, void* const V[3], int nslab, int n, int axis, double* S1k,
                                         double* S2k)
{
    TRY(check_plan(p));
    if (!V || !V[0] || !V[1] || !V[2] || !S1k || !S2k) { set_error("th_cylinder_ring_spectrum: null"); return E_ARG; }
    return DISPATCH(p, do_aniso, p, V, 1, nslab, n, axis, S1k, S2k);
}
extern "C" int th_ring_spectrum_helicity(th_plan* p, void* const V[3], int nsect, int axis, double* H1k)
{
    TRY(check_plan(p));
    if (!V || !V[0] || !V[1] || !V[2] || !H1k) { set_error("th_ring_spectrum_helicity: null"); return E_ARG; }
    return DISPATCH(p, do_aniso, p, V, 2, nsect, 0, axis, H1k, (double*)nullptr);
}
extern "C" int th_cylinder_ring_spectrum_helicity(th_plan* p, void* const V[3], int nslab, int axis, double* H1k)
{
    TRY(check_plan(p));
    if (!V || !V[0] || !V[1] || !V[2] || !H1k) { set_error("th_cylinder_ring_spectrum_helicity: null"); return E_ARG; }
    return DISPATCH(p, do_aniso, p, V, 3, nslab, 0, axis, H1k, (double*)nullptr);
}
extern "C" int th_cylinder_shell_ek_size(const th_plan* p, int axis)
{
    if (!p || axis < 1 || axis > 3) { set_error("th_cylinder_shell_ek_size: bad args"); return E_ARG; }
    return cylinder_shell_ek_size(p->g, axis);
}
template <typename R> static int do_fill(th_plan* p, const void* A, void* B, double in, double out)
{
    return k_fill_shell<R>(p->g, CCP(R, A), CP(R, B), in, out, p->ctx->stream);
}
extern "C" int th_fill_array_shell(th_plan* p, const void* A, void* B, double inner, double outer)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_fill, p, A, B, inner, outer);
}
template <typename R> static int do_sms(th_plan* p, void* const A[3], void* const B[3], double in, double out, double* res)
{
    const cplx<R>* a[3] = {CCP(R, A[0]), CCP(R, A[1]), CCP(R, A[2])};
    const cplx<R>* b[3] = {CCP(R, B[0]), CCP(R, B[1]), CCP(R, B[2])};
    return reduce_shell_mult_single<R>(p->g, a, b, in, out, res, p->rs, p->ctx->stream);
}
extern "C" int th_shell_mult_single(th_plan* p, void* const A[3], void* const B[3], double inner, double outer, double* out)
{
    TRY(check_plan(p));
    return DISPATCH(p, do_sms, p, A, B, inner, outer, out);
}

static int upload_radii(th_plan* p, const double* radii, int nr)
{
    if (nr > p->radii_cap) {
        if (p->radii_dev) (void)hipFree(p->radii_dev);
        p->radii_dev = nullptr;
        TH_CHECK_HIP(hipMalloc(&p->radii_dev, nr * sizeof(double)));
        p->radii_cap = nr;
    }
    TH_CHECK_HIP(hipMemcpyAsync(p->radii_dev, radii, nr * sizeof(double), hipMemcpyHostToDevice, p->ctx->stream));
    TH_CHECK_HIP(hipStreamSynchronize(p->ctx->stream));
    return OK;
}

template <typename R> static int do_sma(th_plan* p, void* const A[3], void* const B[3], int nr, double* res)
{
    const cplx<R>* a[3] = {CCP(R, A[0]), CCP(R, A[1]), CCP(R, A[2])};
    const cplx<R>* b[3] = {CCP(R, B[0]), CCP(R, B[1]), CCP(R, B[2])};
    return reduce_shell_mult_all<R>(p->g, a, b, p->radii_dev, nr, res, p->rs, p->ctx->stream);
}
extern "C" int th_shell_mult_all(th_plan* p, void* const A[3], void* const B[3], const double* radii, int nr, double* out)
{
    TRY(check_plan(p));
    if (!radii || nr < 1 || !out) { set_error("th_shell_mult_all: bad args"); return E_ARG; }
    TRY(upload_radii(p, radii, nr));
    return DISPATCH(p, do_sma, p, A, B, nr, out);
}

template <typename R> static int do_flux_ns(th_plan* p, void* const V[3], const double* radii, int nsph, double* flux)
{
    Ops<R> o(p);
    cplx<R>* v[3] = {CP(R, V[0]), CP(R, V[1]), CP(R, V[2])};
    cplx<R>* G[3] = {o.S(SC_G0), o.S(SC_G0 + 1), o.S(SC_G0 + 2)};
    cplx<R>* Nn[3] = {o.S(SC_NE0), o.S(SC_NE0 + 1), o.S(SC_NE0 + 2)};
    flux[0] = 0;
    for (int s = 1; s <= nsph; s++) {
        for (int i = 0; i < 3; i++) TRY(k_fill_shell<R>(p->g, v[i], G[i], 0.0, radii[s], o.st));
        TRY(o.et_nlin(v, G, Nn));
        const cplx<R>* a[3] = {Nn[0], Nn[1], Nn[2]};
        const cplx<R>* b[3] = {v[0], v[1], v[2]};
        double r;
        TRY(reduce_shell_mult_single<R>(p->g, a, b, radii[s], 10000.0, &r, p->rs, o.st));
        flux[s] = -r;
    }
    return OK;
}
extern "C" int th_flux_ns(th_plan* p, void* const V[3], const double* radii, int nsph, double* flux)
{
    TRY(check_plan(p));
    if (!radii || nsph < 1 || !flux) { set_error("th_flux_ns: bad args"); return E_ARG; }
    return DISPATCH(p, do_flux_ns, p, V, radii, nsph, flux);
}

// IncVF::Compute_flux(IncVF& W) compute_flux.cc:148-262: for each sphere s, G = U<, B<, U> (or the
// Elsasser Zp<, Zm<), N = EnergyTr_Compute_nlin with the helper field named there, and the products
// over |K| > R (Prod_out_sphere_nlinV) or 0 < |K| <= R (Prod_in_sphere_nlinV).
template <typename R> static int do_flux_mhd(th_plan* p, void* const U[3], void* const B[3], const double* radii,
                                             int nsph, double* out)
{
    Ops<R> o(p);
    const double INF = 10000.0;
    cplx<R>* u[3] = {CP(R, U[0]), CP(R, U[1]), CP(R, U[2])};
    cplx<R>* b[3] = {CP(R, B[0]), CP(R, B[1]), CP(R, B[2])};
    cplx<R>* G[3] = {o.S(SC_G0), o.S(SC_G0 + 1), o.S(SC_G0 + 2)};
    cplx<R>* Nn[3] = {o.S(SC_NE0), o.S(SC_NE0 + 1), o.S(SC_NE0 + 2)};
    cplx<R>* zp[3] = {o.S(SC_COPY0), o.S(SC_COPY0 + 1), o.S(SC_COPY0 + 2)};
    cplx<R>* zm[3] = {o.S(SC_COPY0 + 3), o.S(SC_COPY0 + 4), o.S(SC_COPY0 + 5)};
    const int W = nsph + 1;
    for (int i = 0; i < 9 * W; i++) out[i] = 0.0;
    auto fill = [&](cplx<R>* const* F, double in, double outr) -> int {
        for (int i = 0; i < 3; i++) TRY(k_fill_shell<R>(p->g, F[i], G[i], in, outr, o.st));
        return OK;
    };
    auto prod = [&](cplx<R>* const* F, double in, double outr, double* r) -> int {
        const cplx<R>* a[3] = {Nn[0], Nn[1], Nn[2]};
        const cplx<R>* f[3] = {F[0], F[1], F[2]};
        return reduce_shell_mult_single<R>(p->g, a, f, in, outr, r, p->rs, o.st);
    };
    for (int s = 1; s <= nsph; s++) {
        const double Rs = radii[s];
        double r;
        TRY(fill(u, 0.0, Rs));                                    // G = U<
        TRY(o.et_nlin(u, G, Nn));                                 // U.grad U<
        TRY(prod(u, Rs, INF, &r)); out[0 * W + s] = -r;           // flux_self
        TRY(o.et_nlin(b, G, Nn));                                 // B.grad U<
        TRY(prod(b, Rs, INF, &r)); out[2 * W + s] = r;            // flux_VF_in_out
        TRY(prod(b, 0.0, Rs, &r)); out[3 * W + s] = r;            // flux_VF_in_in
        TRY(fill(b, 0.0, Rs));                                    // G = B<
        TRY(o.et_nlin(u, G, Nn));                                 // U.grad B<
        TRY(prod(b, Rs, INF, &r)); out[1 * W + s] = -r;           // W.flux_self
        TRY(prod(u, Rs, INF, &r)); out[4 * W + s] = r;            // W.flux_VF_in_out
        TRY(prod(u, 0.0, Rs, &r)); out[5 * W + s] = r;            // W.flux_VF_in_in
        TRY(fill(u, Rs, INF));                                    // G = U>
        TRY(o.et_nlin(b, G, Nn));                                 // B.grad U>
        TRY(prod(b, Rs, INF, &r)); out[6 * W + s] = r;            // flux_VF_out_out
    }
    for (int i = 0; i < 3; i++) {                                 // UB_to_Elsasser_field
        TRY(o.copy(zp[i], u[i])); TRY(o.axpy(zp[i], 1.0, b[i]));
        TRY(o.copy(zm[i], u[i])); TRY(o.axpy(zm[i], -1.0, b[i]));
    }
    for (int s = 1; s <= nsph; s++) {
        const double Rs = radii[s];
        double r;
        TRY(fill(zp, 0.0, Rs));                                   // G = Zp<
        TRY(o.et_nlin(zm, G, Nn));                                // Zm.grad Zp<
        TRY(prod(zp, Rs, INF, &r)); out[7 * W + s] = -r;          // flux_Elsasser
        TRY(fill(zm, 0.0, Rs));                                   // G = Zm<
        TRY(o.et_nlin(zp, G, Nn));                                // Zp.grad Zm<
        TRY(prod(zm, Rs, INF, &r)); out[8 * W + s] = -r;          // W.flux_Elsasser
    }
    return OK;
}
extern "C" int th_flux_mhd(th_plan* p, void* const U[3], void* const B[3], const double* radii, int nsph, double* out)
{
    TRY(check_plan(p));
    if (!U || !B || !radii || nsph < 1 || !out) { set_error("th_flux_mhd: bad args"); return E_ARG; }
    return DISPATCH(p, do_flux_mhd, p, U, B, radii, nsph, out);
}

// Compute_flux_scalar compute_flux.cc:89-109 (what Compute_flux_RB :113-141 runs for the PRLARGE and
// ULARGE scalings): out row 0 = flux_self (Compute_flux), row 1 = flux_SF with, for each sphere s,
// G = T< (Fill_in_sphere(s, T)), N = U.grad G and flux_SF[s] = -Prod_out_sphere_nlinV(s, T).
template <typename R> static int do_flux_scalar(th_plan* p, void* const V[3], void* T, const double* radii, int nsph,
                                                double* out)
{
    TRY(do_flux_ns<R>(p, V, radii, nsph, out));
    Ops<R> o(p);
    cplx<R>* v[3] = {CP(R, V[0]), CP(R, V[1]), CP(R, V[2])};
    cplx<R>* th = CP(R, T);
    cplx<R>* G = o.S(SC_G0);
    cplx<R>* Nn = o.S(SC_NE0);
    double* sf = out + nsph + 1;
    sf[0] = 0;
    for (int s = 1; s <= nsph; s++) {
        TRY(k_fill_shell<R>(p->g, th, G, 0.0, radii[s], o.st));
        TRY(o.et_nlin_scalar(v, G, Nn));
        const cplx<R>* a[3] = {Nn, nullptr, nullptr};
        const cplx<R>* b[3] = {th, nullptr, nullptr};
        double r;
        TRY(reduce_shell_mult_single<R>(p->g, a, b, radii[s], 10000.0, &r, p->rs, o.st));
        sf[s] = -r;
    }
    return OK;
}
extern "C" int th_flux_scalar(th_plan* p, void* const V[3], void* T, const double* radii, int nsph, double* out)
{
    TRY(check_plan(p));
    if (!V || !T || !radii || nsph < 1 || !out) { set_error("th_flux_scalar: bad args"); return E_ARG; }
    return DISPATCH(p, do_flux_scalar, p, V, T, radii, nsph, out);
}

template <typename R> static int do_shell_tr(th_plan* p, void* const V[3], const double* radii, int nsh, double* T)
{
    Ops<R> o(p);
    cplx<R>* v[3] = {CP(R, V[0]), CP(R, V[1]), CP(R, V[2])};
    cplx<R>* G[3] = {o.S(SC_G0), o.S(SC_G0 + 1), o.S(SC_G0 + 2)};
    cplx<R>* Nn[3] = {o.S(SC_NE0), o.S(SC_NE0 + 1), o.S(SC_NE0 + 2)};
    TRY(upload_radii(p, radii, nsh + 1));
    std::vector<double> row(nsh + 1);
    for (int i = 0; i < (nsh + 1) * (nsh + 1); i++) T[i] = 0;
    for (int nn = 1; nn < nsh; nn++) {
        for (int i = 0; i < 3; i++) TRY(k_fill_shell<R>(p->g, v[i], G[i], radii[nn - 1], radii[nn], o.st));
        TRY(o.et_nlin(v, G, Nn));
        const cplx<R>* a[3] = {Nn[0], Nn[1], Nn[2]};
        const cplx<R>* b[3] = {v[0], v[1], v[2]};
        TRY(reduce_shell_mult_all<R>(p->g, a, b, p->radii_dev, nsh + 1, row.data(), p->rs, o.st));
        for (int m = 0; m <= nsh; m++) T[nn * (nsh + 1) + m] = -row[m];
    }
    return OK;
}
extern "C" int th_shell_tr_ns(th_plan* p, void* const V[3], const double* radii, int nsh, double* T)
{
    TRY(check_plan(p));
    if (!radii || nsh < 2 || !T) { set_error("th_shell_tr_ns: bad args"); return E_ARG; }
    return DISPATCH(p, do_shell_tr, p, V, radii, nsh, T);
}

// Compute_shell_tr_scalar compute_shell_tr.cc:98-117 (Compute_shell_tr_RB :120-150 for PRLARGE and
// ULARGE): Tself = Compute_shell_tr(); Tsf[n][m] = -Shell_mult_all(U.grad T_n, T)[m], n = 1..nsh-1.
template <typename R> static int do_shell_tr_scalar(th_plan* p, void* const V[3], void* T, const double* radii,
                                                    int nsh, double* Tself, double* Tsf)
{
    TRY(do_shell_tr<R>(p, V, radii, nsh, Tself));
    Ops<R> o(p);
    cplx<R>* v[3] = {CP(R, V[0]), CP(R, V[1]), CP(R, V[2])};
    cplx<R>* th = CP(R, T);
    cplx<R>* G = o.S(SC_G0);
    cplx<R>* Nn = o.S(SC_NE0);
    std::vector<double> row(nsh + 1);
    for (int i = 0; i < (nsh + 1) * (nsh + 1); i++) Tsf[i] = 0;
    for (int nn = 1; nn < nsh; nn++) {
        TRY(k_fill_shell<R>(p->g, th, G, radii[nn - 1], radii[nn], o.st));
        TRY(o.et_nlin_scalar(v, G, Nn));
        const cplx<R>* a[3] = {Nn, nullptr, nullptr};
        const cplx<R>* b[3] = {th, nullptr, nullptr};
        TRY(reduce_shell_mult_all<R>(p->g, a, b, p->radii_dev, nsh + 1, row.data(), p->rs, o.st));
        for (int m = 0; m <= nsh; m++) Tsf[nn * (nsh + 1) + m] = -row[m];
    }
    return OK;
}
extern "C" int th_shell_tr_scalar(th_plan* p, void* const V[3], void* T, const double* radii, int nsh, double* Tself,
                                  double* Tsf)
{
    TRY(check_plan(p));
    if (!V || !T || !radii || nsh < 2 || !Tself || !Tsf) { set_error("th_shell_tr_scalar: bad args"); return E_ARG; }
    return DISPATCH(p, do_shell_tr_scalar, p, V, T, radii, nsh, Tself, Tsf);
}

// IncVF::Compute_shell_tr(IncVF& W) compute_shell_tr.cc:154-260: out holds five (nsh+1)^2 row-major
// matrices, rows n = 1..nsh-1 the giver shell: 0 U to U (Compute_shell_tr), 1 W to W = -(U.grad W_n).W,
// 2 U to W = +(W.grad U_n).W, 3 Zp to Zp = -(Zm.grad Zp_n).Zp, 4 Zm to Zm = -(Zp.grad Zm_n).Zm,
// Z+- = U +- B (UB_to_Elsasser_field elsasser.cc:53-60).
template <typename R> static int do_shell_tr_mhd(th_plan* p, void* const U[3], void* const B[3], const double* radii,
                                                 int nsh, double* out)
{
    const size_t M = (size_t)(nsh + 1) * (nsh + 1);
    TRY(do_shell_tr<R>(p, U, radii, nsh, out));
    Ops<R> o(p);
    cplx<R>* u[3] = {CP(R, U[0]), CP(R, U[1]), CP(R, U[2])};
    cplx<R>* b[3] = {CP(R, B[0]), CP(R, B[1]), CP(R, B[2])};
    cplx<R>* G[3] = {o.S(SC_G0), o.S(SC_G0 + 1), o.S(SC_G0 + 2)};
    cplx<R>* Nn[3] = {o.S(SC_NE0), o.S(SC_NE0 + 1), o.S(SC_NE0 + 2)};
    cplx<R>* zp[3] = {o.S(SC_COPY0), o.S(SC_COPY0 + 1), o.S(SC_COPY0 + 2)};
    cplx<R>* zm[3] = {o.S(SC_COPY0 + 3), o.S(SC_COPY0 + 4), o.S(SC_COPY0 + 5)};
    for (int i = 0; i < 3; i++) {
        TRY(o.copy(zp[i], u[i])); TRY(o.axpy(zp[i], 1.0, b[i]));
        TRY(o.copy(zm[i], u[i])); TRY(o.axpy(zm[i], -1.0, b[i]));
    }
    TRY(upload_radii(p, radii, nsh + 1));
    // giver, helper, receiver and sign of matrices 1..4
    cplx<R>* const* giv[4] = {b, u, zp, zm};
    cplx<R>* const* hel[4] = {u, b, zm, zp};
    cplx<R>* const* rec[4] = {b, b, zp, zm};
    const double sgn[4] = {-1.0, 1.0, -1.0, -1.0};
    std::vector<double> row(nsh + 1);
    for (int k = 0; k < 4; k++) {
        double* T = out + (k + 1) * M;
        for (size_t i = 0; i < M; i++) T[i] = 0;
        for (int nn = 1; nn < nsh; nn++) {
            for (int i = 0; i < 3; i++) TRY(k_fill_shell<R>(p->g, giv[k][i], G[i], radii[nn - 1], radii[nn], o.st));
            TRY(o.et_nlin(hel[k], G, Nn));
            const cplx<R>* a[3] = {Nn[0], Nn[1], Nn[2]};
            const cplx<R>* r[3] = {rec[k][0], rec[k][1], rec[k][2]};
            TRY(reduce_shell_mult_all<R>(p->g, a, r, p->radii_dev, nsh + 1, row.data(), p->rs, o.st));
            for (int m = 0; m <= nsh; m++) T[nn * (nsh + 1) + m] = sgn[k] * row[m];
        }
    }
    return OK;
}
extern "C" int th_shell_tr_mhd(th_plan* p, void* const U[3], void* const B[3], const double* radii, int nsh, double* out)
{
    TRY(check_plan(p));
    if (!U || !B || !radii || nsh < 2 || !out) { set_error("th_shell_tr_mhd: bad args"); return E_ARG; }
    return DISPATCH(p, do_shell_tr_mhd, p, U, B, radii, nsh, out);
}

// Ring-to-ring / cylinder-ring-to-ring transfer (IncVF::Compute_ring_tr compute_ring_tr.cc:60-291,
// Compute_cylinder_ring_tr compute_cylin_ring_tr.cc:54-133): for giver ring (gs, gq), gs = 1..ngiv,
// gq = 1..nsec, G = F on the ring, N = EnergyTr_Compute_nlin with helper H (ncomp 3) or its scalar form
// (ncomp 1, G = T ring), and row (gs, gq) of out = sign * Ring_mult_all(N, Rv) over the receiver rings.
template <typename R>
static int do_ring_tr(th_plan* p, int cyl, int axis, void* const H[3], void* const F[3], void* const Rv[3], int ncomp,
                      const double* radii, int nr, const double* edges, int nsec, int ngiv, double sign, double* out)
{
    RingP rp;
    TRY(ring_params(p->g, cyl, axis, radii, nr, edges, nsec, &rp));
    if (ngiv < 1 || ngiv > nr) { set_error("th_ring_tr: giver shells must be 1..%d", nr); return E_ARG; }
    Ops<R> o(p);
    cplx<R>* h[3] = {CP(R, H[0]), CP(R, H[1]), CP(R, H[2])};
    cplx<R>* G[3] = {o.S(SC_G0), o.S(SC_G0 + 1), o.S(SC_G0 + 2)};
    cplx<R>* Nn[3] = {o.S(SC_NE0), o.S(SC_NE0 + 1), o.S(SC_NE0 + 2)};
    const size_t M = (size_t)(nr + 1) * (nsec + 1);
    std::vector<double> row(M);
    for (size_t i = 0; i < M * M; i++) out[i] = 0;
    const cplx<R>* a[3] = {Nn[0], ncomp == 3 ? Nn[1] : nullptr, ncomp == 3 ? Nn[2] : nullptr};
    const cplx<R>* b[3] = {CCP(R, Rv[0]), ncomp == 3 ? CCP(R, Rv[1]) : nullptr, ncomp == 3 ? CCP(R, Rv[2]) : nullptr};
    for (int gs = 1; gs <= ngiv; gs++)
        for (int gq = 1; gq <= nsec; gq++) {
            for (int c = 0; c < ncomp; c++) TRY(fill_ring<R>(p->g, CCP(R, F[c]), G[c], rp, gs, gq, o.st));
            if (ncomp == 3) TRY(o.et_nlin(h, G, Nn));
            else TRY(o.et_nlin_scalar(h, G[0], Nn[0]));
            TRY(reduce_ring_mult<R>(p->g, a, b, rp, row.data(), p->rs, o.st));
            double* dst = out + (size_t)(gs * (nsec + 1) + gq) * M;
            for (size_t j = 0; j < M; j++) dst[j] = sign * row[j];
        }
    return OK;
}
extern "C" int th_ring_tr(th_plan* p, int cyl, int axis, void* const H[3], void* const F[3], void* const Rv[3], int ncomp,
                          const double* radii, int nr, const double* edges, int nsec, int ngiv, double sign, double* out)
{
    TRY(check_plan(p));
    if (!H || !F || !Rv || !out || (ncomp != 1 && ncomp != 3)) { set_error("th_ring_tr: bad args"); return E_ARG; }
    return DISPATCH(p, do_ring_tr, p, cyl, axis, H, F, Rv, ncomp, radii, nr, edges, nsec, ngiv, sign, out);
}

template <typename R>
static int do_ring_mult_all(th_plan* p, int cyl, int axis, void* const A[3], void* const B[3], int ncomp,
                            const double* radii, int nr, const double* edges, int nsec, double* out)
{
    RingP rp;
    TRY(ring_params(p->g, cyl, axis, radii, nr, edges, nsec, &rp));
    const cplx<R>* a[3] = {CCP(R, A[0]), ncomp == 3 ? CCP(R, A[1]) : nullptr, ncomp == 3 ? CCP(R, A[2]) : nullptr};
    const cplx<R>* b[3] = {CCP(R, B[0]), ncomp == 3 ? CCP(R, B[1]) : nullptr, ncomp == 3 ? CCP(R, B[2]) : nullptr};
    return reduce_ring_mult<R>(p->g, a, b, rp, out, p->rs, p->ctx->stream);
}
extern "C" int th_ring_mult_all(th_plan* p, int cyl, int axis, void* const A[3], void* const B[3], int ncomp,
                                const double* radii, int nr, const double* edges, int nsec, double* out)
{
    TRY(check_plan(p));
    if (!A || !B || !out || (ncomp != 1 && ncomp != 3)) { set_error("th_ring_mult_all: bad args"); return E_ARG; }
    return DISPATCH(p, do_ring_mult_all, p, cyl, axis, A, B, ncomp, radii, nr, edges, nsec, out);
}

template <typename R> static int do_elsasser(th_plan* p, void* const U[3], void* const B[3], int back)
{
    for (int c = 0; c < 3; c++) TRY(k_elsasser<R>(p->g.ncplx(), CP(R, U[c]), CP(R, B[c]), back, p->ctx->stream));
    return OK;
}
extern "C" int th_ub_to_elsasser(th_plan* p, void* const U[3], void* const B[3], int back)
{
    TRY(check_plan(p));
    if (!U || !B) { set_error("th_ub_to_elsasser: null"); return E_ARG; }
    return DISPATCH(p, do_elsasser, p, U, B, back);
}

extern "C" int th_nusselt(th_plan* p, void* const V[3], const void* T, double Ra, double Pr, double* Nu)
{
    TRY(check_plan(p));
    double e;
    TRY(th_energy(p, V[0], T, &e));
    *Nu = 1 + 2 * sqrt(Ra * Pr) * e;
    return OK;
}
