// th_force.hip -- energy / helicity forcing of the velocity on a FOUR box (force_field_proc 1 and 2):
//   proc 1: IncFluid::Compute_force_const_energy_helicity_supply compute_force_ek_hk_supply.cc:51-157
//   proc 2: IncFluid::Compute_force_const_energy_helicity compute_force_const_ek_hk.cc:51-140
//   proc 3: IncFluid::Compute_force_Taylor_Green compute_force_TG.cc:50-104 (Setup_Taylor_Green_force_field
//           force_model.cc:57-88), a fixed four-mode field
//   proc 4: IncFluid::Compute_force_ABC compute_force_ABC.cc:52-68 (Setup_ABC_force_field force_model.cc:122-160),
//           a fixed six-mode field (NS only)
//   with IncVF::Force_alpha_beta force_alpha_beta.cc:60-89, Get_local_number_modes_in_shell_FOUR
//   four_basic.cc:48-90, Is_alias_array universal_misc.cc:169-181 + Dealias_force dealias.cc:78-84 and
//   Satisfy_reality_condition_force_field reality_cond.cc:90-95.
//
// The force is nonzero only on the band inner < |K| <= outer.  One thread per mode of the band's bounding
// box (kx in [-kx_max, kx_max], ky in [-ky_max, ky_max], kz in [0, kz_max], the reference's own loop
// bounds) computes alpha, beta and F = alpha V + beta (i K x V) in FP64 from the stored velocity, then
// either stores F (the force field; the rest of the array is cleared first) or subtracts it from the
// nonlinear term in place (Add_force add_force.cc:47-53).  A forced step therefore touches a few thousand
// modes instead of making three more whole-field passes.  Dealias_force and the reality condition are
// folded in: a mode that Satisfy_reality_array overwrites (kz = 0 and kz = N3/2 planes) takes the
// conjugate of the force computed at its source mode, after the dealias rule at that source.
#include <math.h>
#include "../../include/tarang_hip.h"
#include "th_internal.hpp"
#include "th_plan.hpp"
#include "th_force.hpp"

using namespace th;

namespace {

// MODE 0: Out = F on the box (the array was cleared); MODE 1: Out -= F
template <typename R, int MODE>
__global__ void k_force_band(ForceBand b, const cplx<R>* V0, const cplx<R>* V1, const cplx<R>* V2, cplx<R>* O0,
                             cplx<R>* O1, cplx<R>* O2)
{
    const long n = (long)b.nx * b.ny * b.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int lx, ly, lz, sx, sy;
    band_mode(b, t, lx, ly, lz);
    const bool mirror = reality_source(b, lx, ly, lz, sx, sy);
    const long si = ((long)sx * b.N2 + sy) * b.H + lz;
    const cplx<R> v[3] = {V0[si], V1[si], V2[si]};
    double fr[3], fi[3];
    mode_force<R>(b, v, sx, sy, lz, fr, fi);
    if (mirror)
        for (int c = 0; c < 3; c++) fi[c] = -fi[c];
    const long i = ((long)lx * b.N2 + ly) * b.H + lz;
    cplx<R>* O[3] = {O0, O1, O2};
    for (int c = 0; c < 3; c++) {
        if (MODE == 0) {
            O[c][i] = mk<R>((R)fr[c], (R)fi[c]);
        } else {
            const cplx<R> o = O[c][i];
            O[c][i] = mk<R>((R)((double)o.x - fr[c]), (R)((double)o.y - fi[c]));
        }
    }
}

template <typename R, int MODE>
int launch_band(const ForceBand& b, const cplx<R>* const V[3], cplx<R>* const O[3], hipStream_t st)
{
    const long n = (long)b.nx * b.ny * b.nz;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL((k_force_band<R, MODE>), dim3(blocks), dim3(256), 0, st, b, V[0], V[1], V[2], O[0], O[1],
                       O[2]);
    TH_CHECK_HIP(hipGetLastError());
    return OK;
}

}  // namespace

namespace th {

int force_setup(const Geo& g, const th_force* f, ForceBand& b)
{
    if (!f) { set_error("force: null parameters"); return E_ARG; }
    if (f->proc < 1 || f->proc > 4) {
        set_error("force: force_field_proc %d is not supported (1: energy/helicity supply, 2: constant energy/"
                  "helicity, 3: Taylor-Green, 4: ABC)", f->proc);
        return E_UNSUPPORTED;
    }
    if (g.basis != TH_FOUR) { set_error("force: only the FOUR basis is supported"); return E_UNSUPPORTED; }
    if (f->proc == 3 || f->proc == 4) {
        // Compute_force_Taylor_Green compute_force_TG.cc:50-63: k0 = (int) force_double_para(1) (inner_radius),
        // force_amp = force_double_para(2) (outer_radius).  Compute_force_ABC compute_force_ABC.cc:52-68: k0,
        // A, B, C, force_amp = force_double_para(1..5) (inner_radius, outer_radius, p3, p4, p5).  Set once, the
        // same at every stage.
        const char* name = f->proc == 3 ? "Taylor-Green" : "ABC";
        if (!isfinite(f->inner_radius) || !isfinite(f->outer_radius) ||
            (f->proc == 4 && (!isfinite(f->p3) || !isfinite(f->p4) || !isfinite(f->p5)))) {
            set_error("force: %s parameters must be finite", name);
            return E_ARG;
        }
        const int k0 = (int)f->inner_radius;
        bool kept = true;
        if (g.dealias) {
            if (f->proc == 3)
                kept = kept_mode(0, 1, g.N1, g.N2, g.N3, k0, k0, k0) &&
                       kept_mode(0, 1, g.N1, g.N2, g.N3, g.N1 - k0, g.N2 - k0, k0);
            else
                kept = kept_mode(0, 1, g.N1, g.N2, g.N3, k0, 0, 0) && kept_mode(0, 1, g.N1, g.N2, g.N3, g.N1 - k0, 0, 0) &&
                       kept_mode(0, 1, g.N1, g.N2, g.N3, 0, k0, 0) && kept_mode(0, 1, g.N1, g.N2, g.N3, 0, g.N2 - k0, 0) &&
                       kept_mode(0, 1, g.N1, g.N2, g.N3, 0, 0, k0);
        }
        if (k0 < 1 || 2 * k0 >= g.N1 || 2 * k0 >= g.N2 || 2 * k0 >= g.N3 || !kept) {
            set_error("force: %s k0 = %d must satisfy 1 <= k0 < N/2 on every axis and lie in the dealiased modes",
                      name, k0);
            return E_ARG;
        }
        b = ForceBand{};
        b.N1 = g.N1; b.N2 = g.N2; b.N3 = g.N3; b.H = g.H;
        b.f1 = g.kf[0]; b.f2 = g.kf[1]; b.f3 = g.kf[2];
        b.proc = f->proc;
        b.k0 = k0;
        b.per_mode = f->proc == 3 ? f->outer_radius : f->p5;
        b.abc[0] = f->outer_radius; b.abc[1] = f->p3; b.abc[2] = f->p4;
        b.nf = 0;
        b.dealias_force = 0;
        b.emin = MYEPS;
        b.kxm = k0; b.kym = k0;
        b.fullx = b.fully = 0;
        b.nx = 2 * k0 + 1; b.ny = 2 * k0 + 1; b.nz = k0 + 1;
        return OK;
    }
    if (!(f->outer_radius > f->inner_radius) || !(f->inner_radius >= 0) || !isfinite(f->outer_radius) ||
        !isfinite(f->p3) || !isfinite(f->p4)) {
        set_error("force: need 0 <= inner_radius < outer_radius and finite parameters");
        return E_ARG;
    }
    if (f->proc == 2 && !(fabs(f->p4) <= 1.0)) {
        set_error("force: h_by_k_E = %g outside [-1, 1] (the square roots of the constant-helicity force are "
                  "not real)", f->p4);
        return E_ARG;
    }
    const double inner = f->inner_radius, outer = f->outer_radius;
    const int kxm = (int)ceil(outer / g.kf[0]);
    const int kym = g.N2 > 1 ? (int)ceil(outer / g.kf[1]) : 0;
    const int kzm = g.N3 > 2 ? (int)ceil(outer / g.kf[2]) : 0;
    // nf: Get_local_number_modes_in_shell_FOUR four_basic.cc:48-90 (inner <= |K| < outer, kz > 0 twice)
    long nf = 0;
    const int kzc = g.N3 >= 2 ? (int)ceil(outer / g.kf[2]) : 0;
    for (int kx = -kxm; kx <= kxm; kx++)
        for (int ky = -kym; ky <= kym; ky++)
            for (int kz = 0; kz <= kzc; kz++) {
                const int l1 = kx >= 0 ? kx : kx + g.N1, l2 = ky >= 0 ? ky : ky + g.N2;
                const double km = kmag3(kx_four(l1, g.N1) * g.kf[0], ky_of(l2, g.N2) * g.kf[1], kz * g.kf[2]);
                if (km >= inner && km < outer) nf += (kz == 0) ? 1 : 2;
            }
    if (nf <= 0) {
        set_error("force: no wavevector with %g <= |K| < %g", inner, outer);
        return E_ARG;
    }
    b.N1 = g.N1; b.N2 = g.N2; b.N3 = g.N3; b.H = g.H;
    b.f1 = g.kf[0]; b.f2 = g.kf[1]; b.f3 = g.kf[2];
    b.proc = f->proc;
    b.inner = inner; b.outer = outer;
    b.per_mode = f->p3 / (double)nf;
    b.p4 = f->p4;
    b.nf = nf;
    // Is_alias_array universal_misc.cc:169-181: the band reaches past N/3 on some axis
    const int ax = (int)ceil(outer / g.kf[0]), ay = (int)ceil(outer / g.kf[1]), az = (int)ceil(outer / g.kf[2]);
    b.dealias_force = g.dealias && (ax > g.N1 / 3 || ay > g.N2 / 3 || az > g.N3 / 3);
    b.fullx = (2L * kxm + 1 >= g.N1);
    b.fully = (2L * kym + 1 >= g.N2);
    b.kxm = kxm; b.kym = kym;
    b.nx = b.fullx ? g.N1 : 2 * kxm + 1;
    b.ny = b.fully ? g.N2 : 2 * kym + 1;
    b.nz = (kzm < g.H - 1 ? kzm : g.H - 1) + 1;
    b.emin = MYEPS;
    return OK;
}

int force_setup_mhd(const Geo& g, const th_force* f, ForceBand& bu, ForceBand& bw)
{
    if (f && f->proc == 4) {
        set_error("force: the ABC force (force_field_proc 4) is not supported for MHD");
        return E_UNSUPPORTED;
    }
    TRY(force_setup(g, f, bu));
    if (!isfinite(f->p5) || !isfinite(f->p6)) {
        set_error("force: energy_supply_W / energy_level_W and the W helicity ratio must be finite");
        return E_ARG;
    }
    if (f->proc == 2 && !(fabs(f->p6) <= 1.0)) {
        set_error("force: h_by_k_E_W = %g outside [-1, 1] (the square roots of the constant-helicity force are "
                  "not real)", f->p6);
        return E_ARG;
    }
    bw = bu;                                  // same band and nf; W's own level and helicity ratio
    if (f->proc == 3) {                       // compute_force_TG.cc:90-104: forceW_amp = force_double_para(3)
        if (!isfinite(f->p3)) { set_error("force: Taylor-Green forceW_amp must be finite"); return E_ARG; }
        bw.per_mode = f->p3;
        return OK;
    }
    bw.per_mode = f->p5 / (double)bu.nf;
    bw.p4 = f->p6;
    // compute_force_const_ek_hk.cc:376-402: W's modes are forced when W.CV_Modal_energy > 0, U's (and both
    // in the supply scheme, compute_force_ek_hk_supply.cc:377-408) when it exceeds MYEPS
    bw.emin = f->proc == 2 ? 0.0 : MYEPS;
    return OK;
}

template <typename R> int force_subtract(const ForceBand& b, const cplx<R>* const V[3], cplx<R>* const nl[3],
                                         hipStream_t st)
{
    return launch_band<R, 1>(b, V, nl, st);
}
template <typename R> int force_write(const ForceBand& b, const cplx<R>* const V[3], cplx<R>* const F[3], size_t n,
                                      hipStream_t st)
{
    for (int c = 0; c < 3; c++) TH_CHECK_HIP(hipMemsetAsync(F[c], 0, n * sizeof(cplx<R>), st));
    return launch_band<R, 0>(b, V, F, st);
}
template int force_write<float>(const ForceBand&, const cplx<float>* const[3], cplx<float>* const[3], size_t, hipStream_t);
template int force_write<double>(const ForceBand&, const cplx<double>* const[3], cplx<double>* const[3], size_t,
                                 hipStream_t);
template int force_subtract<float>(const ForceBand&, const cplx<float>* const[3], cplx<float>* const[3], hipStream_t);
template int force_subtract<double>(const ForceBand&, const cplx<double>* const[3], cplx<double>* const[3],
                                    hipStream_t);

}  // namespace th

template <typename R> static int compute_force(th_plan* p, void* const V[3], void* F[3], const th_force* f, long* nf)
{
    ForceBand b;
    TRY(force_setup(p->g, f, b));
    const cplx<R>* v[3] = {(const cplx<R>*)V[0], (const cplx<R>*)V[1], (const cplx<R>*)V[2]};
    cplx<R>* o[3] = {(cplx<R>*)F[0], (cplx<R>*)F[1], (cplx<R>*)F[2]};
    TRY(force_write<R>(b, v, o, p->g.ncplx(), p->ctx->stream));
    if (nf) *nf = b.nf;
    return OK;
}

// Compute_force(IncVF& W): both forces, each zero when its level per mode is not above MYEPS
template <typename R> static int compute_force_mhd(th_plan* p, void* const U[3], void* const B[3], void* FU[3],
                                                  void* FW[3], const th_force* f, long* nf)
{
    ForceBand bu, bw;
    TRY(force_setup_mhd(p->g, f, bu, bw));
    const ForceBand* b[2] = {&bu, &bw};
    void* const* src[2] = {U, B};
    void** dst[2] = {FU, FW};
    for (int part = 0; part < 2; part++) {
        const cplx<R>* v[3] = {(const cplx<R>*)src[part][0], (const cplx<R>*)src[part][1], (const cplx<R>*)src[part][2]};
        cplx<R>* o[3] = {(cplx<R>*)dst[part][0], (cplx<R>*)dst[part][1], (cplx<R>*)dst[part][2]};
        if (b[part]->proc == 3 || b[part]->per_mode > MYEPS) {
            TRY(force_write<R>(*b[part], v, o, p->g.ncplx(), p->ctx->stream));
        } else {
            for (int c = 0; c < 3; c++)
                TH_CHECK_HIP(hipMemsetAsync(o[c], 0, p->g.ncplx() * sizeof(cplx<R>), p->ctx->stream));
        }
    }
    if (nf) *nf = bu.nf;
    return OK;
}

extern "C" int th_compute_force_mhd(th_plan* p, void* const U[3], void* const B[3], void* FU[3], void* FW[3],
                                    const th_force* f, long* nf)
{
    if (!p || !U || !B || !FU || !FW) { set_error("th_compute_force_mhd: null"); return E_ARG; }
    for (int c = 0; c < 3; c++)
        if (!U[c] || !B[c] || !FU[c] || !FW[c]) { set_error("th_compute_force_mhd: null"); return E_ARG; }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? compute_force_mhd<float>(p, U, B, FU, FW, f, nf) : compute_force_mhd<double>(p, U, B, FU, FW, f, nf);
}

extern "C" int th_compute_force(th_plan* p, void* const V[3], void* F[3], const th_force* f, long* nf)
{
    if (!p || !V || !F || !V[0] || !V[1] || !V[2] || !F[0] || !F[1] || !F[2]) {
        set_error("th_compute_force: null");
        return E_ARG;
    }
    TH_CHECK_HIP(hipSetDevice(p->ctx->device));
    return p->prec ? compute_force<float>(p, V, F, f, nf) : compute_force<double>(p, V, F, f, nf);
}
