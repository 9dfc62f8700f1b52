// th_fused.hip -- host side of the dealias-pruned, fused pseudo-spectral RK stage (FOUR basis):
// device-resident states (single boxes and batched boxes of one size), the hipGraph-captured RK4
// step, per-box stage coefficients, energy transfer from shell givers, the per-pass profiler, the
// launch-shape autotuner, the uneven-box ensemble and the forced step.  The five passes and their
// kernels are described in th_fused_kern.hpp; NS and MHD states step chained (four launches per stage: each
// stage's x-forward also performs the next stage's x-inverse).
#include <stdio.h>
#include <array>
#include <string>
#include <vector>
#include "th_force.hpp"
#include "th_fused.hpp"

namespace th {

struct Fused {};

void release_fused(th_plan* p)
{
    delete p->fused;
    p->fused = nullptr;
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
    g.scft = geo.basis == TH_SCFT;
    if (geo.dealias) {
        int hx = geo.N1 / 3, nkx = 2 * (geo.N1 / 3) - 1;
        g.mx = LMap{geo.N1, nkx, hx};
        if (g.scft) g.mx = LMap{geo.N1, 2 * geo.N1 / 3, 2 * geo.N1 / 3};   // kx >= 0, kept kx < 2 N1 / 3
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
    // 1 / (N1 N2 N3) times the 1/2 that the z pass's r2c unpacking leaves out (k_fz writes 2 X_k)
    g.norm = 0.5 / ((double)geo.N1 * (double)geo.N2 * (double)geo.N3);
    if (g.scft) g.norm *= 0.5;                                          // Norm_SCFT: transforms scale by 1 / (2 N1 N2 N3)
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
    int var[NPASS] = {0, 0, 0, 0, 0};   // launch-shape variant per pass (autotuner)
    // Chained step (NS or MHD, unforced): each stage's x-forward also does the next stage's x-inverse, so a
    // step runs 4 passes per stage; X1 must then hold the x-inverse of S when a step starts (chain_ok), which a
    // single x-inverse launch restores after anything else has used the buffers or changed S.
    bool chain_ok = false;
    bool graph_chained = false;
    // the stage-field buffer holds the last step's last-stage input (false after a chained step of an unforced
    // state, which never writes it)
    bool stage_kept = true;
    bool forced = false;         // energy / helicity forcing (th_state_set_force)
    ForceBand fb{};
    double* fbox = nullptr;      // projected force P(F) of the stage input, 3 complex FP64 per band-box mode
    size_t fbox_cap = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    ~th_state()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (fbox) (void)hipFree(fbox);
        for (void* q : {S, Vs, acc, X1, X2, X3}) if (q) (void)hipFree(q);
        if (partial) (void)hipFree(partial);
        if (radii) (void)hipFree(radii);
        if (dcoef) (void)hipFree(dcoef);
        if (hcoef) (void)hipHostFree(hcoef);
        if (cev) (void)hipEventDestroy(cev);
    }
};

static std::string tune_key(const th_state* s)
{
    char k[160];
    snprintf(k, sizeof k, "fused:gfx950:%s:%s:N%dx%dx%d:boxes%d", s->esz == 8 ? "fp32" : "fp64",
             s->physics == TH_MHD ? "mhd" : s->physics == TH_RB ? "rb" : "ns", s->fg.N1, s->fg.N2, s->fg.N3, s->nbox);
    return k;
}

// Fused state for nbox boxes of the plan's size. Single states (nbox == 1) also keep room for the
// energy-transfer pipeline (6 B-layout and 9 C-layout comps); batched states hold only what the
// RK step needs.
static int state_create(th_plan* p, th_state** out, int physics, int nbox)
{
    if (!p || !out || nbox < 1) { set_error("th_state_create: bad args"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    if (physics != TH_NS && physics != TH_MHD && physics != TH_RB) { set_error("th_state_create: physics %d unknown", physics); return E_ARG; }
    if ((physics == TH_RB) != (p->g.basis == TH_SCFT)) {
        set_error("th_state_create: the fused path runs NS/MHD on the FOUR basis and RB on the SCFT basis");
        return E_UNSUPPORTED;
    }
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
    s->ncomp = physics == TH_NS ? 3 : physics == TH_MHD ? 6 : 4;
    s->nin = s->ncomp;
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
    std::vector<int> tv;
    if (tune_lookup(p->ctx, tune_key(s), tv) && (int)tv.size() == NPASS)
        for (int k = 0; k < NPASS; k++) s->var[k] = (tv[k] >= 0 && tv[k] < pass_nvar(k)) ? tv[k] : 0;
    const size_t cb = 4 * nb * sizeof(StageCoef);
    if (hipMalloc((void**)&s->dcoef, cb) != hipSuccess || hipHostMalloc((void**)&s->hcoef, cb, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->cev, hipEventDisableTiming) != hipSuccess) {
        set_error("th_state_create: coefficient buffers");
        delete s;
        return E_NOMEM;
    }
    p->nstates++;
    *out = s;
    return OK;
}

extern "C" int th_state_create(th_plan* p, th_state** out, int physics) { return state_create(p, out, physics, 1); }

extern "C" int th_state_destroy(th_state* s)
{
    if (!s) return OK;
    th_plan* p = s->plan;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    delete s;
    if (--p->nstates == 0 && p->dead) plan_free(p);
    return OK;
}

static int state_upload(th_state* s, int box, int comp, const void* src, int src_on_host, hipStream_t st)
{
    if (!s || !src || comp < 0 || comp >= s->ncomp || box < 0 || box >= s->nbox) { set_error("th_state_upload: bad args"); return E_ARG; }
    s->chain_ok = false;   // S changes
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
    const long off = box * g.bS + comp * g.cS;
    int r = p->prec ? d_gather<float>(g, (const cplx<float>*)dsrc, (cplx<float>*)s->S + off, st)
                    : d_gather<double>(g, (const cplx<double>*)dsrc, (cplx<double>*)s->S + off, st);
    if (tmp) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(tmp);
    }
    return r;
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
    const long off = box * g.bS + comp * g.cS;
    int r = p->prec ? d_scatter<float>(g, (const cplx<float>*)s->S + off, (cplx<float>*)ddst, st)
                    : d_scatter<double>(g, (const cplx<double>*)s->S + off, (cplx<double>*)ddst, st);
    hipError_t e = hipSuccess;
    if (!r && tmp) {
        e = hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (tmp) (void)hipFree(tmp);
    if (r) return r;
    if (e != hipSuccess) { set_error("th_state_download: %s", hipGetErrorString(e)); return E_HIP; }
    return OK;
}

// Diagnostics: reference-layout copy of the stage field (which = 1) or the RK4 accumulator
// (which = 2) instead of the state (which = 0).
extern "C" int th_state_download_buffer(th_state* s, int which, int comp, void* dst, int dst_on_host)
{
    if (!s || which < 0 || which > 2) { set_error("th_state_download_buffer: bad args"); return E_ARG; }
    if (which == 1 && !s->stage_kept) {
        set_error("th_state_download_buffer: the chained step of an unforced NS or MHD state keeps no stage field");
        return E_UNSUPPORTED;
    }
    void* keep = s->S;
    s->S = which == 0 ? s->S : which == 1 ? s->Vs : s->acc;
    int r = state_download(s, 0, comp, dst, dst_on_host, s->plan->ctx->stream);
    s->S = keep;
    return r;
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

// Pass `pass` of RK stage `stage` (1..4) for every box of s, launch-shape variant var.  Stage 1
// reads the state, later stages the stage field; pass 5 writes the stage field / accumulator
// (and the state in stage 4).
template <typename R> static int run_pass(th_state* s, int pass, int var, int stage, hipStream_t st, bool chain = false)
{
    th_plan* p = s->plan;
    const FG& g = s->fg;
    const cplx<R>* twx = p->twd<R>(g.N1);
    const cplx<R>* twy = p->twd<R>(g.N2);
    cplx<R>* S = (cplx<R>*)s->S;
    cplx<R>* Vs = (cplx<R>*)s->Vs;
    cplx<R>* acc = (cplx<R>*)s->acc;
    cplx<R>* X1 = (cplx<R>*)s->X1;
    cplx<R>* X2 = (cplx<R>*)s->X2;
    cplx<R>* X3 = (cplx<R>*)s->X3;
    const StageCoef* sc = s->dcoef + (size_t)(stage - 1) * s->nbox;
    const cplx<R>* src = stage == 1 ? S : Vs;
    const int ph = s->physics;
    const int nin = s->nin, nout = s->nout;
    ShellSel sel{0, 0};
    switch (pass) {
    case PASS_XINV:
        if (ph == TH_RB) return d_sx_inv<R>(g, src, X1, twx, p->twd<R>(4 * g.N1), st);
        return ph == TH_MHD ? d_xinv<R, 1>(g, src, X1, nin, twx, sel, var, st)
                            : d_xinv<R, 0>(g, src, X1, nin, twx, sel, var, st);
    case PASS_YINV: return d_yinv<R>(g, X1, X2, nin, twy, var, st);
    case PASS_Z: {
        const cplx<R>* twz = p->twd<R>(g.LZ);
        const cplx<R>* twn = p->twd<R>(g.N3);
        if (ph == TH_RB) return d_z<R, 3>(g, X2, X3, twz, twn, var, st);
        return ph == TH_MHD ? d_z<R, 1>(g, X2, X3, twz, twn, var, st) : d_z<R, 0>(g, X2, X3, twz, twn, var, st);
    }
    case PASS_YFWD: return d_yfwd<R>(g, X3, X1, nout, twy, var, st);
    default:
        if (ph == TH_RB) return d_sx_fwd_rb<R>(g, X1, S, Vs, acc, twx, p->twd<R>(4 * g.N1), sc, var, st);
        return ph == TH_MHD ? d_xfwd<R, 1>(g, X1, S, Vs, acc, twx, sc, var, st, chain)
                            : d_xfwd<R, 0>(g, X1, S, Vs, acc, twx, sc, var, st, chain);
    }
}

// Dissipation of the two field groups: (nu, eta) for NS / MHD; for RB (nu, kappa), or with Ra > 0
// and Pr > 0 the PRLARGE / ULARGE values sqrt(Pr/Ra), 1/sqrt(Pr Ra) (RB_slip_main.cc:238-239).
static void diss_of(const th_state* s, const th_phys* ph, double* d0, double* d1)
{
    *d0 = ph->nu;
    *d1 = ph->eta;
    if (s->physics == TH_RB) {
        *d1 = ph->kappa;
        if (ph->Ra > 0 && ph->Pr > 0) {
            *d0 = sqrt(ph->Pr / ph->Ra);
            *d1 = 1.0 / sqrt(ph->Pr * ph->Ra);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Forced step (force_field_proc 1 / 2, Compute_force_TO_rhs single_time_step.cc:351-364).  The
// reference subtracts the force F of the stage field from nlin before the projection, so the stage's
// rhs gains +P(F), P the projection.  Every stage output (the stage field, the accumulator, the state
// in stage 4) is linear in the rhs and F lives on a few thousand band modes, so instead of changing
// the x-forward epilogue, k_force_pre stores P(F) of the stage input for the band box before the
// stage's passes, and k_force_post adds the stage's share of it afterwards:
//   stage 1: Vs += dt/2 e_h P(F), acc += dt/6 e_1 P(F)      stage 2: Vs += dt/2 P(F), acc += dt/3 e_h P(F)
//   stage 3: Vs += dt e_h P(F),   acc += dt/3 e_h P(F)      stage 4: S += dt/6 P(F)
// with e_h = exp(-nu dt K^2 / 2), e_1 = exp(-nu dt K^2) (the integrating factors of rk_store).
// A chained forced step writes no stage field: in stages 2-4, k_force_recon recovers the stage input's band
// modes from the A lines (the x-inverse of that input, which the previous stage's x-forward wrote) by a direct
// DFT along x; the x-forward has transformed the stage output without the force share into the next stage's A
// lines, so k_force_post records the share it adds to the next stage's input (S in stage 4) per band mode, and
// k_force_chain_fix adds its x-inverse to the A lines of the band's columns.
// ------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ bool fused_stored(const FG& g, int lx, int ly, int lz)
{
    return lz < g.nkz && lkept(g.mx, lx) && lkept(g.my, ly);
}

// Chained forced step, stages 2-4: the stage input at every band mode's reality source (sx, sy, lz),
// v_c = (1/N1) sum_x A[c][x][sy][lz] e^{-2 pi i sx x / N1}, into vbox (6 doubles per band mode; 0 where the
// state stores no such mode).  One workgroup of 256 threads per band mode.
template <typename R>
__global__ __launch_bounds__(256) void k_force_recon(ForceBand b, FG g, const cplx<R>* __restrict__ A,
                                                     double* __restrict__ vbox)
{
    const long t = blockIdx.x;
    int lx, ly, lz, sx, sy;
    band_mode(b, t, lx, ly, lz);
    (void)reality_source(b, lx, ly, lz, sx, sy);
    __shared__ double red[6][256];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const bool st = fused_stored(g, sx, sy, lz);
    if (st) {
        const cplx<R>* col = A + (long)lq(g.my, sy) * g.KZP + lz;
        for (int x = threadIdx.x; x < g.N1; x += blockDim.x) {
            const long m = ((long)sx * x) % g.N1;   // e^{-2 pi i sx x / N1} from the exact phase index
            double sn, cs;
            sincospi(2.0 * (double)m / g.N1, &sn, &cs);
            for (int c = 0; c < 3; c++) {
                const cplx<R> a = col[c * g.cA + (long)x * g.sx];
                acc[2 * c] += (double)a.x * cs + (double)a.y * sn;
                acc[2 * c + 1] += (double)a.y * cs - (double)a.x * sn;
            }
        }
    }
    for (int j = 0; j < 6; j++) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 6; j++) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) vbox[6 * t + threadIdx.x] = st ? red[threadIdx.x][0] / g.N1 : 0.0;
}

// vbox (chained stages 2-4): the stage input's band modes from k_force_recon instead of src
template <typename R>
__global__ void k_force_pre(ForceBand b, FG g, const cplx<R>* __restrict__ src, const double* __restrict__ vbox,
                            double* __restrict__ fbox)
{
    const long n = (long)b.nx * b.ny * b.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int lx, ly, lz, sx, sy;
    band_mode(b, t, lx, ly, lz);
    const bool mirror = reality_source(b, lx, ly, lz, sx, sy);
    cplx<R> v[3] = {czero<R>(), czero<R>(), czero<R>()};
    if (vbox) {
        for (int c = 0; c < 3; c++) v[c] = mk<R>((R)vbox[6 * t + 2 * c], (R)vbox[6 * t + 2 * c + 1]);
    } else if (fused_stored(g, sx, sy, lz)) {
        const long i = (long)lq(g.mx, sx) * g.sx + (long)lq(g.my, sy) * g.KZP + lz;
        for (int c = 0; c < 3; c++) v[c] = src[c * g.cS + i];
    }
    double fr[3], fi[3];
    mode_force<R>(b, v, sx, sy, lz, fr, fi);
    if (mirror)
        for (int c = 0; c < 3; c++) fi[c] = -fi[c];
    const double K[3] = {kx_four(lx, g.N1) * g.f1, ky_of(ly, g.N2) * g.f2, lz * g.f3};
    const double k2 = ksq_f(K[0], K[1], K[2]);
    if (k2 > 0) {
        const double dr = (K[0] * fr[0] + K[1] * fr[1] + K[2] * fr[2]) / k2;
        const double di = (K[0] * fi[0] + K[1] * fi[1] + K[2] * fi[2]) / k2;
        for (int c = 0; c < 3; c++) {
            fr[c] -= K[c] * dr;
            fi[c] -= K[c] * di;
        }
    }
    for (int c = 0; c < 3; c++) {
        fbox[6 * t + 2 * c] = fr[c];
        fbox[6 * t + 2 * c + 1] = fi[c];
    }
}

// chained: no stage field to update (the next stage's A lines get the share through k_force_chain_fix) except
// the stage-3 output, which the chained x-forward keeps for a forced state
template <typename R>
__global__ void k_force_post(ForceBand b, FG g, const StageCoef* __restrict__ coef, cplx<R>* S, cplx<R>* Vs,
                             cplx<R>* acc, double* __restrict__ fbox, int chained)
{
    const long n = (long)b.nx * b.ny * b.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double* share = fbox + 6 * n + 6 * t;   // what this stage adds to the next stage's input (0: nothing)
    for (int j = 0; j < 6; j++) share[j] = 0.0;
    int lx, ly, lz;
    band_mode(b, t, lx, ly, lz);
    if (!fused_stored(g, lx, ly, lz)) return;
    double f[6];
    bool any = false;
    for (int j = 0; j < 6; j++) {
        f[j] = fbox[6 * t + j];
        any = any || f[j] != 0.0;
    }
    if (!any) return;
    const StageCoef sc = *coef;
    const double kx = kx_four(lx, g.N1) * g.f1, ky = ky_of(ly, g.N2) * g.f2, kz = lz * g.f3;
    const double kx2 = ksq_f(kx, 0.0, 0.0), kyz2 = ksq_f(0.0, ky, kz);
    const double eh = exp(sc.fh[0] * kx2) * exp(sc.fh[0] * kyz2), e1 = exp(sc.f1[0] * kx2) * exp(sc.f1[0] * kyz2);
    double cv = 0.0, ca = 0.0;
    if (sc.stage == 1) { cv = 0.5 * sc.dt * eh; ca = sc.dt / 6 * e1; }
    else if (sc.stage == 2) { cv = 0.5 * sc.dt; ca = sc.dt / 3 * eh; }
    else if (sc.stage == 3) { cv = sc.dt * eh; ca = sc.dt / 3 * eh; }
    const long i = (long)lq(g.mx, lx) * g.sx + (long)lq(g.my, ly) * g.KZP + lz;
    for (int c = 0; c < 3; c++) {
        const long j = c * g.cS + i;
        if (sc.stage < 4) {
            if (!chained || sc.stage == 3) {   // chained: only the kept stage-3 output (FG::keep3)
                const cplx<R> v = Vs[j];
                Vs[j] = mk<R>((R)((double)v.x + cv * f[2 * c]), (R)((double)v.y + cv * f[2 * c + 1]));
            }
            const cplx<R> a = acc[j];
            acc[j] = mk<R>((R)((double)a.x + ca * f[2 * c]), (R)((double)a.y + ca * f[2 * c + 1]));
            share[2 * c] = cv * f[2 * c];
            share[2 * c + 1] = cv * f[2 * c + 1];
        } else {
            const cplx<R> v = S[j];
            const double c4 = sc.dt / 6;
            S[j] = mk<R>((R)((double)v.x + c4 * f[2 * c]), (R)((double)v.y + c4 * f[2 * c + 1]));
            share[2 * c] = c4 * f[2 * c];
            share[2 * c + 1] = c4 * f[2 * c + 1];
        }
    }
}

// Chained forced step: A[c][x][column] += sum over the column's band modes lx of share * e^{+2 pi i lx x / N1}
// (the unnormalised inverse line transform of the shares k_force_post recorded).  One workgroup per band
// column (ty, tz) and component; columns the state does not store, or without a nonzero share, are left alone.
template <typename R>
__global__ void k_force_chain_fix(ForceBand b, FG g, cplx<R>* A, const double* __restrict__ share)
{
    const int ty = blockIdx.x / b.nz, tz = blockIdx.x % b.nz, c = blockIdx.y;
    int lx, ly, lz;
    band_mode(b, (long)ty * b.nz + tz, lx, ly, lz);
    if (lz >= g.nkz || !lkept(g.my, ly)) return;
    bool any = false;
    for (int tx = 0; tx < b.nx; tx++) {
        const double* d = share + 6 * (((long)tx * b.ny + ty) * b.nz + tz) + 2 * c;
        any = any || d[0] != 0.0 || d[1] != 0.0;
    }
    if (!any) return;
    cplx<R>* col = A + c * g.cA + (long)lq(g.my, ly) * g.KZP + lz;
    for (int x = threadIdx.x; x < g.N1; x += blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int tx = 0; tx < b.nx; tx++) {
            const long t = ((long)tx * b.ny + ty) * b.nz + tz;
            const double dr = share[6 * t + 2 * c], di = share[6 * t + 2 * c + 1];
            if (dr == 0.0 && di == 0.0) continue;
            band_mode(b, t, lx, ly, lz);
            const long m = ((long)lx * x) % g.N1;   // e^{+2 pi i lx x / N1} from the exact phase index
            double sn, cs;
            sincospi(2.0 * (double)m / g.N1, &sn, &cs);
            sr += dr * cs - di * sn;
            si += dr * sn + di * cs;
        }
        cplx<R>& a = col[(long)x * g.sx];
        a = mk<R>((R)((double)a.x + sr), (R)((double)a.y + si));
    }
}

// One RK4 step of every box in s, as 4 stages x 5 pass launches on stream st (forced: plus the
// band kernels around each stage); chained: 4 passes per stage, the x-inverse done by the previous
// stage's x-forward (the first stage's by the step before, or by the priming launch in th_step), and a forced
// chained stage adds k_force_recon before and k_force_chain_fix after.
template <typename R> static int enqueue_step(th_state* s, hipStream_t st, bool chained = false)
{
    const long nb = (long)s->fb.nx * s->fb.ny * s->fb.nz;
    const dim3 fgrid((unsigned)((nb + 255) / 256));
    for (int stage = 1; stage <= 4; stage++) {
        if (s->forced) {
            const cplx<R>* src = (const cplx<R>*)(stage == 1 ? s->S : s->Vs);
            const double* vbox = nullptr;
            if (chained && stage > 1) {   // no stage field: its band modes from the A lines
                vbox = (const double*)s->fbox + 12 * nb;
                hipLaunchKernelGGL((k_force_recon<R>), dim3((unsigned)nb), dim3(256), 0, st, s->fb, s->fg,
                                   (const cplx<R>*)s->X1, (double*)vbox);
                TH_CHECK_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL((k_force_pre<R>), fgrid, dim3(256), 0, st, s->fb, s->fg, src, vbox, s->fbox);
            TH_CHECK_HIP(hipGetLastError());
        }
        for (int k = 0; k < NPASS; k++) {
            if (chained && k == PASS_XINV) continue;
            TRY(run_pass<R>(s, k, s->var[k], stage, st, chained));
        }
        if (s->forced) {
            hipLaunchKernelGGL((k_force_post<R>), fgrid, dim3(256), 0, st, s->fb, s->fg,
                               (const StageCoef*)(s->dcoef + (size_t)(stage - 1) * s->nbox), (cplx<R>*)s->S,
                               (cplx<R>*)s->Vs, (cplx<R>*)s->acc, (double*)s->fbox, chained ? 1 : 0);
            TH_CHECK_HIP(hipGetLastError());
            if (chained) {
                hipLaunchKernelGGL((k_force_chain_fix<R>), dim3((unsigned)(s->fb.ny * s->fb.nz), 3), dim3(256), 0, st,
                                   s->fb, s->fg, (cplx<R>*)s->X1, (const double*)s->fbox + 6 * nb);
                TH_CHECK_HIP(hipGetLastError());
            }
        }
    }
    return OK;
}

extern "C" int th_state_set_force(th_state* s, const th_force* f)
{
    if (!s) { set_error("th_state_set_force: null state"); return E_ARG; }
    s->chain_ok = false;
    TH_CHECK_HIP(hipSetDevice(s->plan->ctx->device));
    // steps already enqueued read the force buffer and the captured graph: let them finish first
    TH_CHECK_HIP(hipStreamSynchronize(s->plan->ctx->stream));
    if (!f || f->proc == 0) {
        s->forced = false;
    } else {
        if (s->physics != TH_NS || s->nbox != 1) {
            set_error("th_state_set_force: forcing runs on single-box NS states");
            return E_UNSUPPORTED;
        }
        if (f->proc == 3 || f->proc == 4) {
            set_error("th_state_set_force: the Taylor-Green and ABC forces run on the plain stepper");
            return E_UNSUPPORTED;
        }
        ForceBand fb;
        TRY(force_setup(s->plan->g, f, fb));
        // every mode that can carry a force must be stored by the 2/3-dealiased state
        const long nb = (long)fb.nx * fb.ny * fb.nz;
        for (long t = 0; t < nb; t++) {
            int lx, ly, lz, sx, sy;
            band_mode(fb, t, lx, ly, lz);
            (void)reality_source(fb, lx, ly, lz, sx, sy);
            const double km = kmag3(kx_four(sx, fb.N1) * fb.f1, ky_of(sy, fb.N2) * fb.f2, lz * fb.f3);
            const bool can = km > fb.inner && km <= fb.outer &&
                             (!fb.dealias_force || kept_mode(0, 1, fb.N1, fb.N2, fb.N3, sx, sy, lz));
            if (can && !fused_stored(s->fg, lx, ly, lz)) {
                set_error("th_state_set_force: the band reaches mode (%d, %d, %d), which the 2/3-dealiased fused "
                          "state does not store; use th_step_plain_forced", kx_four(lx, fb.N1), ky_of(ly, fb.N2), lz);
                return E_UNSUPPORTED;
            }
        }
        // per band mode: P(F); the stage's share and the reconstructed stage input (chained steps)
        const size_t need = (size_t)nb * 18;
        if (need > s->fbox_cap) {
            // the old buffer goes away: the state is unforced (and its captured step void) until this succeeds
            s->forced = false;
            s->fg.keep3 = 0;
            if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
            if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
            if (s->fbox) TH_CHECK_HIP(hipFree(s->fbox));
            s->fbox = nullptr;
            s->fbox_cap = 0;
            if (hipMalloc((void**)&s->fbox, need * sizeof(double)) != hipSuccess) {
                s->fbox = nullptr;
                set_error("th_state_set_force: out of device memory");
                return E_NOMEM;
            }
            s->fbox_cap = need;
        }
        s->fb = fb;
        s->forced = true;
    }
    s->fg.keep3 = s->forced ? 1 : 0;
    // the captured step changes
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    return OK;
}

static int ensure_twiddles(th_plan* p, const FG& g, hipStream_t st)
{
    for (int L : {g.N1, g.N2, g.LZ, g.N3, g.scft ? 4 * g.N1 : g.N1})
        if (!p->tw.get(L, p->prec, st)) { set_error("twiddle table allocation failed"); return E_NOMEM; }
    return OK;
}

// whether a step runs this state chained (NS, forced or not, or MHD; N1 with a chained x-forward; any number of
// boxes)
static bool step_chained(const th_state* s)
{
    return (s->physics == TH_NS || s->physics == TH_MHD) &&
           (s->plan->prec ? d_xfwd_chainable<float>(s->fg) : d_xfwd_chainable<double>(s->fg));
}

extern "C" int th_step(th_plan* p, th_state* s, double dt, const th_phys* ph, int nsteps)
{
    if (!p || !s || !ph || s->plan != p || nsteps < 0) { set_error("th_step: bad args"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    hipStream_t st = p->ctx->stream;
    // twiddle tables must exist before graph capture (their creation allocates)
    TRY(ensure_twiddles(p, s->fg, st));
    double d0, d1;
    diss_of(s, ph, &d0, &d1);
    TRY(set_coefs(s, &dt, &d0, &d1, st));
    const bool chained = step_chained(s);
    if (s->gexec && s->graph_chained != chained) {
        (void)hipGraphExecDestroy(s->gexec);
        s->gexec = nullptr;
        if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    }
    if (!s->gexec) {
        TH_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int r = p->prec ? enqueue_step<float>(s, st, chained) : enqueue_step<double>(s, st, chained);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
        if (e != hipSuccess) { set_error("hipStreamEndCapture: %s", hipGetErrorString(e)); return E_HIP; }
        s->graph = graph;
        TH_CHECK_HIP(hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0));
        s->graph_chained = chained;
    }
    if (chained && !s->chain_ok && nsteps > 0)   // X1 <- x-inverse of S, the first stage's input
        TRY(p->prec ? run_pass<float>(s, PASS_XINV, s->var[PASS_XINV], 1, st)
                    : run_pass<double>(s, PASS_XINV, s->var[PASS_XINV], 1, st));
    for (int i = 0; i < nsteps; i++) TH_CHECK_HIP(hipGraphLaunch(s->gexec, st));
    if (nsteps > 0) {
        s->chain_ok = chained;
        s->stage_kept = !chained || s->forced;
    }
    return OK;
}


// Energy transfer from the state's velocity: one giver-masked fused pipeline per shell n gives
// T(n, m) (Shell_mult_all, |K| <= kmax_inside) and the same sums over |K| > kmax_inside ("corner",
// which Shell_mult_single counts); then Pi(R_i) = sum_{n<=i} sum_{m>i} [T(n,m) - corner(n,m)]
// (sphere i = union of shells 1..i and N[G] is linear in G).
// Per giver shell n (inner radii[n-1] < |K| <= radii[n], n = 1..ngiv) of the energy-transfer pipeline with
// helper H = state components hoff..hoff+2, giver G = components goff..goff+2 and each receiver R_r =
// components roff[r]..roff[r]+2 of the state, the sums over the modes of receiver shell m (Get_shell_index
// basis_basicfn.cc:288-299 on |K|) of 2 mult Re(N[H, G_n] . R_r*), N = EnergyTr_Compute_nlin
// (ET_compute_nlin.cc:85-120), split as Shell_mult_all splits them: Yb over |K| <= kmax_inside
// (four_ET.cc:142) and Yc over the corners above it.  Rows are [nr], nr = nsh + 1, indexed n * nr + m.
template <typename R>
static int et_shell_rows(th_state* s, const double* radii, int nsh, int ngiv, int hoff, int goff, const int* roff, int nrecv,
                         std::vector<double>* Yb, std::vector<double>* Yc)
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
    // helper H: inverse x, y once, kept in X2 comps 0..2 (B layout)
    TRY((d_xinv<R, 0>(g, S + (long)hoff * g.cS, X1, 3, twx, all, 0, st)));
    // the y passes take the step's tuned launch shapes (the same kernels, any component count)
    TRY(d_yinv<R>(g, X1, X2, 3, twy, s->var[PASS_YINV], st));
    std::vector<double> hostbuf(2 * nr);
    for (int r = 0; r < nrecv; r++) {
        Yb[r].assign((size_t)(ngiv + 1) * nr, 0.0);
        Yc[r].assign((size_t)(ngiv + 1) * nr, 0.0);
    }
    double* outd = s->partial + (size_t)nblk * (2 * nr);
    for (int n = 1; n <= ngiv; n++) {
        ShellSel sel{radii[n - 1], radii[n]};
        TRY((d_xinv<R, 2>(g, S + (long)goff * g.cS, X1, 3, twx, sel, 0, st)));
        TRY(d_yinv<R>(g, X1, X2 + 3 * g.cB, 3, twy, s->var[PASS_YINV], st));
        TRY((d_z<R, 2>(g, X2, X3, twz, twn, 0, st)));
        TRY(d_yfwd<R>(g, X3, X1, 9, twy, s->var[PASS_YFWD], st));
        for (int r = 0; r < nrecv; r++) {
            TRY(d_xfwd_et<R>(g, X1, S + (long)roff[r] * g.cS, twx, s->radii, nr, kin, s->partial, &nblk, st));
            launch_finalize((const double*)s->partial, nblk, 2 * nr, 2 * nr, outd, st);
            TH_CHECK_HIP(hipGetLastError());
            TH_CHECK_HIP(hipMemcpyAsync(hostbuf.data(), outd, 2 * nr * sizeof(double), hipMemcpyDeviceToHost, st));
            TH_CHECK_HIP(hipStreamSynchronize(st));
            for (int m = 0; m < nr; m++) {
                Yb[r][(size_t)n * nr + m] = hostbuf[m];
                Yc[r][(size_t)n * nr + m] = hostbuf[nr + m];
            }
        }
    }
    return OK;
}

// Compute_flux / Compute_shell_tr (compute_flux.cc:52-69, compute_shell_tr.cc:59-79) from shell givers:
// T(n,m) = -Yb(n,m) (Shell_mult_all, |K| <= kmax_inside) and Pi(R_i) = sum_{n<=i} sum_{m>i} [T(n,m) - Yc(n,m)]
// (Shell_mult_single over |K| > R_i counts the corners too), Pi(INF) = 0.
template <typename R> static int flux_impl(th_state* s, const double* radii, int nsh, double* Pi, double* T)
{
    const int nr = nsh + 1;
    std::vector<double> Yb, Yc;
    const int roff = 0;
    TRY(et_shell_rows<R>(s, radii, nsh, nsh - 1, 0, 0, &roff, 1, &Yb, &Yc));
    std::vector<double> Tm((size_t)nr * nr, 0.0);
    for (int n = 1; n < nsh; n++)
        for (int m = 0; m < nr; m++) Tm[(size_t)n * nr + m] = -Yb[(size_t)n * nr + m];
    if (T)
        for (size_t i = 0; i < (size_t)nr * nr; i++) T[i] = Tm[i];
    if (Pi) {
        Pi[0] = 0;
        for (int i = 1; i <= nsh; i++) {
            double acc_ = 0;
            for (int n = 1; n <= i && n < nsh; n++)
                for (int m = i + 1; m < nr; m++) acc_ += Tm[(size_t)n * nr + m] - Yc[(size_t)n * nr + m];
            Pi[i] = (i >= nsh) ? 0.0 : acc_;
        }
    }
    return OK;
}

// IncVF::Compute_flux(IncVF& W) compute_flux.cc:148-262 on the fused MHD state (U = comps 0..2, B = 3..5):
// for helper H and giver G in {U, B} and receiver R in {U, B}, Y_HGR(n, m) = Yb + Yc of et_shell_rows, and
// out_HGR(i) = sum_{n<=i} sum_{m>i} Y (giver inside R_i, receiver |K| > R_i: Prod_out_sphere_nlinV),
// in_HGR(i) = sum_{n<=i} sum_{0<m<=i} Y (Prod_in_sphere_nlinV), outout(i) = sum_{n>i} sum_{m>i} Y (giver U>).
// Rows as th_flux_mhd: u2u = -out_UUU, b2b = -out_UBB, u2b_in_out = out_BUB, u2b_in_in = in_BUB,
// b2u_in_out = out_UBU, b2u_in_in = in_UBU, u_out_b_out = outout_BUB; the Elsasser fluxes from N's
// bilinearity: Zp = -out over (Z- helper, Z+ giver, Z+ receiver) = -sum s_H out_HGR (s_U = 1, s_B = -1), and
// Zm = -sum s_G s_R out_HGR.
template <typename R> static int flux_mhd_impl(th_state* s, const double* radii, int nsh, double* out)
{
    const int nr = nsh + 1, W = nsh + 1;
    const int roff[2] = {0, 3};
    std::vector<double> Yb[4][2], Yc[4][2];
    // pipelines (H, G): 0 (U, U), 1 (U, B), 2 (B, U), 3 (B, B); givers up to the infinite shell where the
    // in-sphere sums of the last sphere or the outside givers need it
    const int hg[4][2] = {{0, 0}, {0, 3}, {3, 0}, {3, 3}};
    const int ngiv[4] = {nsh - 1, nsh, nsh, nsh - 1};
    for (int k = 0; k < 4; k++) TRY(et_shell_rows<R>(s, radii, nsh, ngiv[k], hg[k][0], hg[k][1], roff, 2, Yb[k], Yc[k]));
    auto Y = [&](int k, int r, int n, int m) {
        if (n > ngiv[k]) return 0.0;
        return Yb[k][r][(size_t)n * nr + m] + Yc[k][r][(size_t)n * nr + m];
    };
    auto sum = [&](int k, int r, int n0, int n1, int m0, int m1) {
        double a = 0;
        for (int n = n0; n <= n1; n++)
            for (int m = m0; m <= m1; m++) a += Y(k, r, n, m);
        return a;
    };
    for (int i = 0; i < 9 * W; i++) out[i] = 0.0;
    for (int i = 1; i <= nsh; i++) {
        double o[4][2];
        for (int k = 0; k < 4; k++)
            for (int r = 0; r < 2; r++) o[k][r] = sum(k, r, 1, i, i + 1, nsh);
        out[0 * W + i] = -o[0][0];                          // U<->U>  flux_self
        out[1 * W + i] = -o[1][1];                          // B<->B>  W.flux_self
        out[2 * W + i] = o[2][1];                           // U<->B>  flux_VF_in_out (helper B)
        out[3 * W + i] = sum(2, 1, 1, i, 1, i);             // U<->B<  flux_VF_in_in
        out[4 * W + i] = o[1][0];                           // B<->U>  W.flux_VF_in_out (helper U)
        out[5 * W + i] = sum(1, 0, 1, i, 1, i);             // B<->U<  W.flux_VF_in_in
        out[6 * W + i] = sum(2, 1, i + 1, nsh, i + 1, nsh); // U>->B>  flux_VF_out_out
        const double sH[4] = {1, 1, -1, -1}, sG[4] = {1, -1, 1, -1}, sR[2] = {1, -1};
        double zp = 0, zm = 0;
        for (int k = 0; k < 4; k++)
            for (int r = 0; r < 2; r++) {
                zp += sH[k] * o[k][r];
                zm += sG[k] * sR[r] * o[k][r];
            }
        out[7 * W + i] = -zp;
        out[8 * W + i] = -zm;
    }
    return OK;
}

extern "C" int th_flux_mhd_fused(th_plan* p, th_state* s, const double* radii, int nsph, double* out)
{
    if (s) s->chain_ok = false;   // the pipelines use X1..X3
    if (!p || !s || s->plan != p || !radii || nsph < 2 || !out) { set_error("th_flux_mhd_fused: bad args"); return E_ARG; }
    if (s->physics != TH_MHD) { set_error("th_flux_mhd_fused: MHD state required"); return E_UNSUPPORTED; }
    if (radii[nsph] < 1e3) { set_error("th_flux_mhd_fused: radii[nsph] must be the infinite radius"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? flux_mhd_impl<float>(s, radii, nsph, out) : flux_mhd_impl<double>(s, radii, nsph, out);
}

extern "C" int th_flux_shell_tr(th_plan* p, th_state* s, const double* radii, int nsh, double* Pi, double* T)
{
    if (s) s->chain_ok = false;   // the pipelines use X1..X3
    if (!p || !s || s->plan != p || !radii || nsh < 2) { set_error("th_flux_shell_tr: bad args"); return E_ARG; }
    if (s->physics != TH_NS) { set_error("th_flux_shell_tr: NS state required"); return E_UNSUPPORTED; }
    if (radii[nsh] < 1e3) { set_error("th_flux_shell_tr: radii[nsh] must be the infinite radius"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? flux_impl<float>(s, radii, nsh, Pi, T) : flux_impl<double>(s, radii, nsh, Pi, T);
}

// ------------------------------------------------------------------------------------------
// Per-pass timing of one fused RK stage with HIP events on the plan's stream, and each pass's
// algorithmic HBM bytes (compulsory reads + writes of kept data; padding excluded).  A chained state is timed as
// th_step runs it: the x-inverse once beforehand (not timed: 0 ms, 0 bytes) and the chained x-forward.
// ------------------------------------------------------------------------------------------
template <typename R> static int profile_impl(th_state* s, double dt, double nu, double eta, int stage, int nrep,
                                              double* ms, double* bytes)
{
    th_plan* p = s->plan;
    hipStream_t st = p->ctx->stream;
    const FG& g = s->fg;
    TRY(set_coefs(s, &dt, &nu, &eta, st));
    // stage 4 writes the state itself: keep a copy and put it back afterwards
    const size_t sbytes = (size_t)s->nbox * g.bS * s->esz;
    void* keep = nullptr;
    if (stage == 4) {
        if (hipMalloc(&keep, sbytes) != hipSuccess) { set_error("th_profile_passes: out of device memory"); return E_NOMEM; }
        TH_CHECK_HIP(hipMemcpyAsync(keep, s->S, sbytes, hipMemcpyDeviceToDevice, st));
    }
    const bool chained = step_chained(s);
    s->chain_ok = false;   // the passes below leave other data in X1 (and stage 4 restores S)
    if (chained) TRY(run_pass<R>(s, PASS_XINV, s->var[PASS_XINV], stage, st));
    hipEvent_t ev[6];
    for (int i = 0; i < 6; i++) TH_CHECK_HIP(hipEventCreate(&ev[i]));
    double tot[5] = {0, 0, 0, 0, 0};
    for (int rep = 0; rep < nrep; rep++) {
        TH_CHECK_HIP(hipEventRecord(ev[0], st));
        for (int k = 0; k < NPASS; k++) {
            if (!(chained && k == PASS_XINV)) TRY(run_pass<R>(s, k, s->var[k], stage, st, chained));
            TH_CHECK_HIP(hipEventRecord(ev[k + 1], st));
        }
        TH_CHECK_HIP(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; k++) {
            float t = 0;
            TH_CHECK_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            tot[k] += t;
        }
    }
    for (int i = 0; i < 6; i++) (void)hipEventDestroy(ev[i]);
    if (keep) {
        TH_CHECK_HIP(hipMemcpyAsync(s->S, keep, sbytes, hipMemcpyDeviceToDevice, st));
        TH_CHECK_HIP(hipStreamSynchronize(st));
        (void)hipFree(keep);
    }
    const double e = (double)sizeof(cplx<R>) * s->nbox;
    const int nin = s->nin, nout = s->nout, nst = s->ncomp;
    const double cS = (double)g.mx.nk * g.my.nk * g.nkz, cA = (double)g.N1 * g.my.nk * g.nkz,
                 cB = (double)g.N1 * g.N2 * g.nkz;
    int state_rw = (stage == 1 || stage == 4) ? 3 : 4;   // V0 (+acc) read, Vs / acc / S written
    double extra = 0;
    if (s->physics == TH_RB && stage > 1) extra = 2 * cS;  // V1 and T of the stage field (force)
    double b[5] = {e * (nst * cS + nin * cA), e * (nin * cA + nin * cB), e * (nin * cB + nout * cB),
                   e * (nout * cB + nout * cA), e * (nout * cA + state_rw * nst * cS + extra)};
    if (chained) {
        // no stage-field write (stages 1-3), the next stage's A lines written instead
        b[PASS_XFWD] = e * (nout * cA + (state_rw - (stage < 4 ? 1 : 0)) * nst * cS + nin * cA);
        b[PASS_XINV] = 0;
    }
    for (int k = 0; k < 5; k++) {
        ms[k] = (chained && k == PASS_XINV) ? 0.0 : tot[k] / nrep;
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
    double d0, d1;
    diss_of(s, ph, &d0, &d1);
    return p->prec ? profile_impl<float>(s, dt, d0, d1, stage, nrep, ms, bytes)
                   : profile_impl<double>(s, dt, d0, d1, stage, nrep, ms, bytes);
}

// ------------------------------------------------------------------------------------------
// Launch-shape autotuner: every variant of every pass is timed (HIP events, stage-2 pipeline on the
// state's own buffers, nrep launches after one warm-up) and the fastest is kept, in the state and
// in the context's tune cache.  Stage 2 writes only the stage field and the accumulator, which
// the next step's first stage overwrites, so the state itself is left unchanged.
// ------------------------------------------------------------------------------------------
template <typename R> static int tune_impl(th_state* s, int nrep, hipStream_t st, double* ms)
{
    hipEvent_t e0, e1;
    TH_CHECK_HIP(hipEventCreate(&e0));
    TH_CHECK_HIP(hipEventCreate(&e1));
    int r = OK;
    // a chained state's x-forward is timed as th_step runs it (with the next x-inverse folded in; it writes its
    // A lines over D, which the trials after it read as they would D: the work is the same)
    const bool chained = step_chained(s);
    for (int k = 0; k < NPASS && !r; k++) r = run_pass<R>(s, k, s->var[k], 2, st);   // live inputs for every pass
    // variants are timed in interleaved rounds and each keeps its best round, so a transient slowdown (clock
    // ramp, another queue's work) cannot decide the choice on its own
    constexpr int ROUNDS = 3;
    for (int k = 0; k < NPASS && !r; k++) {
        const bool ch = chained && k == PASS_XFWD;
        double tv[8];
        for (int v = 0; v < 8; v++) tv[v] = 1e300;
        for (int round = 0; round < ROUNDS && !r; round++)
            for (int v = 0; v < pass_nvar(k) && v < 8 && !r; v++) {
                if ((r = run_pass<R>(s, k, v, 2, st, ch))) break;
                if (hipEventRecord(e0, st) != hipSuccess) { r = E_HIP; break; }
                for (int i = 0; i < nrep && !r; i++) r = run_pass<R>(s, k, v, 2, st, ch);
                if (r || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { r = r ? r : E_HIP; break; }
                float t = 0;
                (void)hipEventElapsedTime(&t, e0, e1);
                if (t / nrep < tv[v]) tv[v] = t / nrep;
            }
        int best = s->var[k];
        double tbest = 1e300;
        for (int v = 0; v < pass_nvar(k) && v < 8; v++)
            if (tv[v] < tbest) { tbest = tv[v]; best = v; }
        s->var[k] = best;
        if (ms) ms[k] = tbest;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r == E_HIP) set_error("th_tune: HIP failure while timing");
    return r;
}

static int state_tune(th_state* s, const double* dt, const double* nu, const double* eta, int nrep, hipStream_t st)
{
    s->chain_ok = false;   // the trials run every pass on the state's buffers
    th_plan* p = s->plan;
    TRY(ensure_twiddles(p, s->fg, st));
    TRY(set_coefs(s, dt, nu, eta, st));
    TRY(p->prec ? tune_impl<float>(s, nrep, st, nullptr) : tune_impl<double>(s, nrep, st, nullptr));
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    tune_store(p->ctx, tune_key(s), std::vector<int>(s->var, s->var + NPASS));
    return OK;
}

extern "C" int th_tune(th_plan* p, th_state* s, double dt, const th_phys* ph, int nrep, int* var)
{
    if (!p || !s || s->plan != p || !ph || nrep < 1) { set_error("th_tune: bad args"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    double d0, d1;
    diss_of(s, ph, &d0, &d1);
    TRY(state_tune(s, &dt, &d0, &d1, nrep, p->ctx->stream));
    if (var)
        for (int k = 0; k < NPASS; k++) var[k] = s->var[k];
    return OK;
}

extern "C" int th_state_variants(const th_state* s, int* var)
{
    if (!s || !var) { set_error("th_state_variants: bad args"); return E_ARG; }
    for (int k = 0; k < NPASS; k++) var[k] = s->var[k];
    return OK;
}

extern "C" int th_state_set_variants(th_state* s, const int* var)
{
    if (!s || !var) { set_error("th_state_set_variants: bad args"); return E_ARG; }
    for (int k = 0; k < NPASS; k++)
        if (var[k] < 0 || var[k] >= pass_nvar(k)) { set_error("th_state_set_variants: pass %d has no variant %d", k, var[k]); return E_ARG; }
    for (int k = 0; k < NPASS; k++) s->var[k] = var[k];
    (void)hipSetDevice(s->plan->ctx->device);
    (void)hipStreamSynchronize(s->plan->ctx->stream);
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    return OK;
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
            const bool ch = step_chained(k.st);
            r = e->prec ? enqueue_step<float>(k.st, k.stream, ch) : enqueue_step<double>(k.st, k.stream, ch);
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
    // chained classes need the x-inverse of their state in X1 before the first stage (see th_step)
    if (nsteps > 0)
        for (th_ensemble::Cls& k : e->cls)
            if (step_chained(k.st) && !k.st->chain_ok)
                TRY(e->prec ? run_pass<float>(k.st, PASS_XINV, k.st->var[PASS_XINV], 1, st)
                            : run_pass<double>(k.st, PASS_XINV, k.st->var[PASS_XINV], 1, st));
    for (int i = 0; i < nsteps; i++) TH_CHECK_HIP(hipGraphLaunch(e->gexec, st));
    if (nsteps > 0)
        for (th_ensemble::Cls& k : e->cls) k.st->chain_ok = step_chained(k.st);
    return OK;
}

extern "C" int th_ensemble_variants(const th_ensemble* e, int box, int* var)
{
    if (!e || !var || box < 0 || box >= e->nbox) { set_error("th_ensemble_variants: bad args"); return E_ARG; }
    return th_state_variants(e->cls[e->cls_of[box]].st, var);
}

extern "C" int th_ensemble_tune(th_ensemble* e, const double* dt, const double* nu, int nrep)
{
    if (!e || !dt || !nu || nrep < 1) { set_error("th_ensemble_tune: bad args"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(e->ctx->device));
    hipStream_t st = e->ctx->stream;
    for (th_ensemble::Cls& k : e->cls) {
        e->dtc.clear();
        e->nuc.clear();
        for (int b : k.boxes) {
            e->dtc.push_back(dt[b]);
            e->nuc.push_back(nu[b]);
        }
        TRY(state_tune(k.st, e->dtc.data(), e->nuc.data(), nullptr, nrep, st));
    }
    if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
    if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
    return OK;
}
