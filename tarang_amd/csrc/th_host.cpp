// th_host.cpp -- the C++ host classes of tarang_host.hpp over the C-ABI.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdexcept>
#include <math.h>
#include <string.h>
#include <string>
#include <iomanip>
#include "tarang_host.hpp"

namespace tarang {

void check(int rc, const char* what)
{
    if (rc != TH_OK) throw std::runtime_error(std::string(what) + ": " + th_last_error());
}

static void hipc(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
// real part of the k = 0 mode of each component (the mean field)
void mean_modes(void* const V[3], size_t esize, double out[3])
{
    for (int i = 0; i < 3; i++) {
        double z[2] = {0, 0};
        float zf[2] = {0, 0};
        if (esize == 16) {
            hipc(hipMemcpy(z, V[i], 16, hipMemcpyDeviceToHost), "mean mode");
        } else {
            hipc(hipMemcpy(zf, V[i], 8, hipMemcpyDeviceToHost), "mean mode");
            z[0] = zf[0];
        }
        out[i] = z[0];
    }
}
}  // namespace

Context::Context(int device) { check(th_ctx_create(&h, device, nullptr), "th_ctx_create"); }
Context::~Context() { th_ctx_destroy(h); }

DField::DField(size_t elems, size_t esize) : n_(elems), es_(esize)
{
    hipc(hipMalloc(&p_, n_ * es_), "hipMalloc");
    hipc(hipMemset(p_, 0, n_ * es_), "hipMemset");
}
DField::~DField()
{
    if (p_) (void)hipFree(p_);
}
DField::DField(DField&& o) noexcept : p_(o.p_), n_(o.n_), es_(o.es_) { o.p_ = nullptr; }
DField& DField::operator=(DField&& o) noexcept
{
    if (this != &o) {
        if (p_) (void)hipFree(p_);
        p_ = o.p_;
        n_ = o.n_;
        es_ = o.es_;
        o.p_ = nullptr;
    }
    return *this;
}
void DField::upload(const void* host) { hipc(hipMemcpy(p_, host, bytes(), hipMemcpyHostToDevice), "upload"); }
void DField::download(void* host) const { hipc(hipMemcpy(host, p_, bytes(), hipMemcpyDeviceToHost), "download"); }
void DField::copy_from(const DField& o) { hipc(hipMemcpy(p_, o.p_, bytes(), hipMemcpyDeviceToDevice), "copy"); }
void DField::zero() { hipc(hipMemset(p_, 0, bytes()), "zero"); }

// ---------------------------------------------------------------- CVF
CVF::CVF(Context& ctx, const Grid& g) : grid(g), ctx_(ctx)
{
    const int basis = g.basis_type == "SCFT" ? TH_SCFT : TH_FOUR;
    const int prec = g.precision == "FP32_MIXED" ? TH_FP32_MIXED : TH_FP64;
    const double kf[3] = {g.kfactor[1], g.kfactor[2], g.kfactor[3]};
    check(th_plan_create(ctx.h, &plan_, basis, g.N[1], g.N[2], g.N[3], kf, prec, g.alias_switch != "ALIAS"),
          "th_plan_create");
    elems_ = th_plan_field_elems(plan_);
    esize_ = th_plan_elem_bytes(plan_);
    V1 = DField(elems_, esize_);
    V2 = DField(elems_, esize_);
    V3 = DField(elems_, esize_);
}
CVF::~CVF() { th_plan_destroy(plan_); }

void CVF::CV_Forward_transform()
{
    check(th_forward(plan_, V1.data(), 1), "CV_Forward_transform");
    check(th_forward(plan_, V2.data(), 0), "CV_Forward_transform");
    check(th_forward(plan_, V3.data(), 0), "CV_Forward_transform");
}
void CVF::CV_Inverse_transform()
{
    check(th_inverse(plan_, V1.data(), 1), "CV_Inverse_transform");
    check(th_inverse(plan_, V2.data(), 0), "CV_Inverse_transform");
    check(th_inverse(plan_, V3.data(), 0), "CV_Inverse_transform");
}
void CVF::CV_Compute_totalenergy_diss()
{
    double e[3], s[3];
    for (int i = 0; i < 3; i++) {
        check(th_energy(plan_, v(i), nullptr, &e[i]), "th_energy");
        check(th_Sn(plan_, v(i), 2, &s[i]), "th_Sn");
    }
    CV_total_energy = e[0] + e[1] + e[2];
    CV_total_dissipation = 2 * (s[0] + s[1] + s[2]);
}
int CVF::shell_ek_size() const { return th_shell_ek_size(plan_); }

void CVF::CV_output(const std::string& path, double time, bool append)
{
    std::vector<char> host(3 * elems_ * esize_);
    const void* h[3];
    for (int i = 0; i < 3; i++) {
        DField const& f = i == 0 ? V1 : i == 1 ? V2 : V3;
        f.download(host.data() + i * elems_ * esize_);
        h[i] = host.data() + i * elems_ * esize_;
    }
    const int N[3] = {grid.N[1], grid.N[2], grid.N[3]};
    const int prec = esize_ == 16 ? TH_FP64 : TH_FP32_MIXED;
    check(th_field_write(path.c_str(), N, prec, 3, h, time, append ? 1 : 0), "CV_output");
}

double CVF::CV_input(const std::string& path, int snapshot)
{
    std::vector<char> host(3 * elems_ * esize_);
    void* h[3];
    for (int i = 0; i < 3; i++) h[i] = host.data() + i * elems_ * esize_;
    const int N[3] = {grid.N[1], grid.N[2], grid.N[3]};
    const int prec = esize_ == 16 ? TH_FP64 : TH_FP32_MIXED;
    double t = 0;
    check(th_field_read(path.c_str(), N, prec, 3, h, snapshot, &t), "CV_input");
    V1.upload(h[0]);
    V2.upload(h[1]);
    V3.upload(h[2]);
    return t;
}
void CVF::CV_output(const std::string& path, const int Nreduced[3], double time, bool append)
{
    std::vector<char> host(3 * elems_ * esize_);
    const void* h[3];
    for (int i = 0; i < 3; i++) {
        DField const& f = i == 0 ? V1 : i == 1 ? V2 : V3;
        f.download(host.data() + i * elems_ * esize_);
        h[i] = host.data() + i * elems_ * esize_;
    }
    const int N[3] = {grid.N[1], grid.N[2], grid.N[3]};
    const int prec = esize_ == 16 ? TH_FP64 : TH_FP32_MIXED;
    check(th_field_write_reduced(path.c_str(), N, Nreduced, grid.basis_type == "SCFT" ? TH_SCFT : TH_FOUR, prec, 3, h, time,
                                 append ? 1 : 0),
          "CV_output(Nreduced)");
}

double CVF::CV_input(const std::string& path, const int Nreduced[3], int snapshot)
{
    std::vector<char> host(3 * elems_ * esize_);
    void* h[3];
    for (int i = 0; i < 3; i++) h[i] = host.data() + i * elems_ * esize_;
    const int N[3] = {grid.N[1], grid.N[2], grid.N[3]};
    const int prec = esize_ == 16 ? TH_FP64 : TH_FP32_MIXED;
    double t = 0;
    check(th_field_read_reduced(path.c_str(), N, Nreduced, grid.basis_type == "SCFT" ? TH_SCFT : TH_FOUR, prec, 3, h, snapshot,
                                &t),
          "CV_input(Nreduced)");
    V1.upload(h[0]);
    V2.upload(h[1]);
    V3.upload(h[2]);
    return t;
}

void CVF::CV_Compute_shell_spectrum(std::vector<double>& result)
{
    int n = shell_ek_size();
    result.assign(n, 0.0);
    std::vector<double> tmp(n);
    for (int i = 0; i < 3; i++) {
        check(th_shell_spectrum(plan_, v(i), nullptr, 0, tmp.data(), nullptr), "th_shell_spectrum");
        for (int k = 0; k < n; k++) result[k] += tmp[k];
    }
}
void CVF::CV_Dealias()
{
    for (int i = 0; i < 3; i++) check(th_dealias(plan_, v(i)), "th_dealias");
}

// ---------------------------------------------------------------- RVF / NLIN
RVF::RVF(CVF& o) : o_(o)
{
    size_t n = th_plan_field_elems(o.plan()), e = th_plan_elem_bytes(o.plan());
    V1r = DField(n, e);
    V2r = DField(n, e);
    V3r = DField(n, e);
}
void RVF::RV_Inverse_transform()
{
    check(th_inverse(o_.plan(), V1r.data(), 1), "RV_Inverse_transform");
    check(th_inverse(o_.plan(), V2r.data(), 0), "RV_Inverse_transform");
    check(th_inverse(o_.plan(), V3r.data(), 0), "RV_Inverse_transform");
}
void RVF::RV_Forward_transform_RSprod()
{
    check(th_forward(o_.plan(), V1r.data(), 1), "RV_Forward_transform_RSprod");
    check(th_forward(o_.plan(), V2r.data(), 0), "RV_Forward_transform_RSprod");
    check(th_forward(o_.plan(), V3r.data(), 1), "RV_Forward_transform_RSprod");
}

NLIN::NLIN(CVF& o) : o_(o)
{
    size_t n = th_plan_field_elems(o.plan()), e = th_plan_elem_bytes(o.plan());
    nlin1 = DField(n, e);
    nlin2 = DField(n, e);
    nlin3 = DField(n, e);
    F = DField(n, e);
}
void NLIN::NLIN_diag_Forward_transform_derivative()
{
    th_plan* p = o_.plan();
    check(th_forward(p, nlin1.data(), 0), "NLIN forward");
    check(th_forward(p, nlin2.data(), 0), "NLIN forward");
    check(th_forward(p, nlin3.data(), 0), "NLIN forward");
    check(th_deriv(p, 1, nlin1.data(), nlin1.data(), 0), "Xderiv");
    check(th_deriv(p, 2, nlin2.data(), nlin2.data(), 0), "Yderiv");
    check(th_deriv(p, 3, nlin3.data(), nlin3.data(), 0), "Zderiv");
}

// ---------------------------------------------------------------- EnergyTr
EnergyTr::EnergyTr(const CVF& f, int nsh, int scheme) : no_shells(nsh), no_spheres(nsh), et_f_(&f)
{
    shell_radius.assign(nsh + 1, 0.0);
    if (scheme == 0) {
        double mri = 1.0 * th_max_radius_inside(f.plan());
        shell_radius[0] = 0;
        shell_radius[1] = 2;
        shell_radius[2] = 4;
        shell_radius[3] = 8;
        shell_radius[nsh - 2] = mri / 2;
        shell_radius[nsh - 1] = mri;
        if (nsh > 6) {
            double s = log2(mri / 16.0) / (nsh - 5);
            for (int i = 4; i <= nsh - 3; i++) shell_radius[i] = 8 * pow(2, s * (i - 3));
        }
    } else {
        for (int i = 0; i < nsh; i++) shell_radius[i] = i;
    }
    shell_radius[nsh] = 10000;      // INF_RADIUS extern_def_vars.h:88
    sphere_radius = shell_radius;
    flux_self.assign(nsh + 1, 0.0);
    shelltoshell_self.assign((size_t)(nsh + 1) * (nsh + 1), 0.0);
}
// EnergyTr.cc:176-360 (ring_shell_input_scheme / cylinder_shell_input_scheme 0 and the linear sector and slab
// schemes 0); the cylinder radii of scheme 0 take their log-spaced middle when no_ring_shells > 5, as the
// reference tests that count there (EnergyTr.cc:318).
void EnergyTr::Init_ring_tr(int nring, int nsect, int ring_scheme, int ncyl, int nslab, int cyl_scheme, int axis)
{
    if (axis < 1 || axis > 3) throw std::runtime_error("Init_ring_tr: anisotropy axis must be 1, 2 or 3");
    anisotropy_axis = axis;
    auto radii = [](int n, int scheme, double kmax) {
        std::vector<double> r(n + 1, 0.0);
        if (scheme == 0) {
            r[0] = 0;
            r[1] = 4;
            r[2] = 8;
            r[n - 2] = kmax / 2;
            r[n - 1] = kmax;
        } else {
            for (int i = 0; i < n; i++) r[i] = i;
        }
        r[n] = 10000;
        return r;
    };
    if (nring > 0) {
        if (nring < 3 || nsect < 1) throw std::runtime_error("Init_ring_tr: need >= 3 ring shells and >= 1 sector");
        no_ring_shells = nring;
        no_sectors_ring_tr = nsect;
        const double kmax = 1.0 * th_max_radius_inside(et_f_->plan());
        ring_shell_radius = radii(nring, ring_scheme, kmax);
        if (ring_scheme == 0 && nring > 5) {
            const double sc = log2(kmax / 16.0) / (nring - 4);
            for (int i = 3; i <= nring - 3; i++) ring_shell_radius[i] = 8 * pow(2, sc * (i - 2));
        }
        sector_angle_ring_tr.assign(nsect + 1, 0.0);
        const double dth = (M_PI / 2) / nsect;              // Get_max_polar_angle = pi/2
        for (int i = 0; i < nsect; i++) sector_angle_ring_tr[i] = i * dth;
        sector_angle_ring_tr[nsect] = M_PI / 2;
    }
    if (ncyl > 0) {
        if (ncyl < 3 || nslab < 1) throw std::runtime_error("Init_ring_tr: need >= 3 cylinder shells and >= 1 slab");
        no_cylinder_shells = ncyl;
        no_cylinder_kpll_slabs = nslab;
        const double kmax = 1.0 * (th_cylinder_shell_ek_size(et_f_->plan(), axis) - 1);   // Anis_max_Krho_radius_inside
        cylinder_shell_radius = radii(ncyl, cyl_scheme, kmax);
        if (cyl_scheme == 0 && no_ring_shells > 5) {
            const double sc = log2(kmax / 16.0) / (ncyl - 4);
            for (int i = 3; i <= ncyl - 3; i++) cylinder_shell_radius[i] = 8 * pow(2, sc * (i - 2));
        }
        // Anis_max_Kpll four_inline.h:484-514: (N_axis / 3) kfactor_axis on a dealiased grid, N_axis / 2 otherwise
        const Grid& g = et_f_->grid;
        const double kpmax = (g.N[axis] / (g.alias_switch == "DEALIAS" ? 3 : 2)) * g.kfactor[axis];
        cylinder_kpll_array_tr.assign(nslab + 1, 0.0);
        const double dk = (kpmax - 0.0) / nslab;
        for (int i = 1; i < nslab; i++) cylinder_kpll_array_tr[i] = 0.0 + i * dk;
        cylinder_kpll_array_tr[nslab] = kpmax;
    }
}

void EnergyTr::Set_shell_radii(const std::vector<double>& r)
{
    if ((int)r.size() < no_shells) throw std::runtime_error("Set_shell_radii: need no_shells radii");
    for (int i = 0; i < no_shells; i++) shell_radius[i] = r[i];
    shell_radius[no_shells] = 10000;
    sphere_radius = shell_radius;
}

// ---------------------------------------------------------------- IncVF
IncVF::IncVF(Context& ctx, const Grid& g, double nu, int nsh, int scheme)
    : CVF(ctx, g), NLIN(static_cast<CVF&>(*this)), EnergyTr(static_cast<CVF&>(*this), nsh, scheme),
      dissipation_coefficient(nu)
{
}

void IncVF::Compute_nlin()
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    void* N[3] = {nlin1.data(), nlin2.data(), nlin3.data()};
    check(th_nlin_ns(plan_, V, N), "Compute_nlin");
}
void IncVF::Compute_pressure()
{
    void* N[3] = {nlin1.data(), nlin2.data(), nlin3.data()};
    check(th_pressure(plan_, N, F.data()), "Compute_pressure");
}
void IncVF::Compute_flux()
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    check(th_flux_ns(plan_, V, sphere_radius.data(), no_spheres, flux_self.data()), "Compute_flux");
}
void IncVF::Compute_shell_tr()
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    check(th_shell_tr_ns(plan_, V, shell_radius.data(), no_shells, shelltoshell_self.data()), "Compute_shell_tr");
}
void IncVF::ring_table(bool cyl, void* const H[3], void* const F[3], void* const R[3], int ncomp, double sign, int ngiv,
                       std::vector<double>& out)
{
    const std::vector<double>& r = cyl ? cylinder_shell_radius : ring_shell_radius;
    const std::vector<double>& e = cyl ? cylinder_kpll_array_tr : sector_angle_ring_tr;
    const int nr = cyl ? no_cylinder_shells : no_ring_shells, ns = cyl ? no_cylinder_kpll_slabs : no_sectors_ring_tr;
    if (nr < 3) throw std::runtime_error("ring transfers: Init_ring_tr has not set these rings");
    const size_t M = (size_t)(nr + 1) * (ns + 1);
    out.assign(M * M, 0.0);
    check(th_ring_tr(plan_, cyl ? 1 : 0, anisotropy_axis, H, F, R, ncomp, r.data(), nr, e.data(), ns, ngiv, sign,
                     out.data()),
          cyl ? "Compute_cylinder_ring_tr" : "Compute_ring_tr");
}
void IncVF::Compute_ring_tr()
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    ring_table(false, V, V, V, 3, -1.0, no_ring_shells - 1, ring_to_ring_self);
}
void IncVF::Compute_cylinder_ring_tr()
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    ring_table(true, V, V, V, 3, -1.0, no_cylinder_shells - 1, cylinder_ring_to_ring_self);
}
// compute_ring_tr.cc:174-291 / the cylinder counterpart: W to W (helper U), U to W (helper W, + sign), and the
// Elsasser pair after UB_to_Elsasser_field (converted back at the end, as the reference does).  The reference
// loops its MHD giver shells up to no_shells - 1 (the shell-to-shell count); here up to no_ring_shells - 1,
// the extent of the ring radii.
void IncVF::Compute_ring_tr(IncVF& W)
{
    W.no_ring_shells = no_ring_shells;
    W.no_sectors_ring_tr = no_sectors_ring_tr;
    Compute_ring_tr();
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    const int ng = no_ring_shells - 1;
    ring_table(false, U, B, B, 3, -1.0, ng, W.ring_to_ring_self);
    ring_table(false, B, U, B, 3, 1.0, ng, ring_to_ring_VF);
    check(th_ub_to_elsasser(plan_, U, B, 0), "UB_to_Elsasser_field");
    ring_table(false, B, U, U, 3, -1.0, ng, ring_to_ring_Elsasser);
    ring_table(false, U, B, B, 3, -1.0, ng, W.ring_to_ring_Elsasser);
    check(th_ub_to_elsasser(plan_, U, B, 1), "Elsasser_to_UB_field");
    check(th_sync(plan_), "th_sync");
}
void IncVF::Compute_cylinder_ring_tr(IncVF& W)
{
    W.no_cylinder_shells = no_cylinder_shells;
    W.no_cylinder_kpll_slabs = no_cylinder_kpll_slabs;
    Compute_cylinder_ring_tr();
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    const int ng = no_cylinder_shells - 1;
    ring_table(true, U, B, B, 3, -1.0, ng, W.cylinder_ring_to_ring_self);
    ring_table(true, B, U, B, 3, 1.0, ng, cylinder_ring_to_ring_VF);
    check(th_ub_to_elsasser(plan_, U, B, 0), "UB_to_Elsasser_field");
    ring_table(true, B, U, U, 3, -1.0, ng, cylinder_ring_to_ring_Elsasser);
    ring_table(true, U, B, B, 3, -1.0, ng, W.cylinder_ring_to_ring_Elsasser);
    check(th_ub_to_elsasser(plan_, U, B, 1), "Elsasser_to_UB_field");
    check(th_sync(plan_), "th_sync");
}

void IncVF::Setup_Taylor_Green_field(int k0, double amp)
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    check(th_ic_taylor_green(plan_, V, k0, amp), "Setup_Taylor_Green_field");
}

// ---------------------------------------------------------------- IncSF
IncSF::IncSF(Context& ctx, const Grid& g, double kappa) : diffusion_coefficient(kappa)
{
    const int basis = g.basis_type == "SCFT" ? TH_SCFT : TH_FOUR;
    const int prec = g.precision == "FP32_MIXED" ? TH_FP32_MIXED : TH_FP64;
    const double kf[3] = {g.kfactor[1], g.kfactor[2], g.kfactor[3]};
    check(th_plan_create(ctx.h, &plan_, basis, g.N[1], g.N[2], g.N[3], kf, prec, g.alias_switch != "ALIAS"),
          "th_plan_create");
    F = DField(th_plan_field_elems(plan_), th_plan_elem_bytes(plan_));
}
IncSF::~IncSF() { th_plan_destroy(plan_); }
void IncSF::CS_Compute_totalenergy_diss()
{
    double s;
    check(th_energy(plan_, F.data(), nullptr, &CS_total_energy), "CS_Compute_totalenergy_diss");
    check(th_Sn(plan_, F.data(), 2, &s), "CS_Compute_totalenergy_diss");
    CS_total_dissipation = 2 * s;
}
void IncSF::CS_Compute_entropy() { check(th_entropy_scalar(plan_, F.data(), &CS_entropy), "CS_Compute_entropy"); }

// ---------------------------------------------------------------- IncFluid
IncFluid::IncFluid(Context& ctx, const Grid& g, double nu, double Tinit, double tdt_fixed, const std::string& scheme,
                   bool fused, int nsh, int shell_scheme)
    : IncVF(ctx, g, nu, nsh, shell_scheme), Tfinal(Tinit), Tnow(Tinit), Tdt(tdt_fixed), Tdt_fixed(tdt_fixed),
      integ_scheme(scheme), Tinit_(Tinit),
      fused_(fused && g.basis_type == "FOUR" && scheme == "RK4")
{
    if (fused_) check(th_state_create(plan_, &state_, TH_NS), "th_state_create");
}
IncFluid::~IncFluid()
{
    if (state_) th_state_destroy(state_);
    if (mstate_) th_state_destroy(mstate_);
    if (rstate_) th_state_destroy(rstate_);
}
void IncFluid::Compute_rhs()
{
    void* N[3] = {nlin1.data(), nlin2.data(), nlin3.data()};
    check(th_rhs(plan_, N, F.data()), "Compute_rhs");
}
double IncFluid::Get_dt()
{
    CV_Compute_totalenergy_diss();
    double kmax = th_max_radius_inside(plan_);
    double dx = (2 * M_PI) / kmax;
    double dt = dx / (sqrt(CV_total_energy) * 20);
    return dt < Tdt_fixed ? dt : Tdt_fixed;
}
void IncFluid::Output_field(const std::string& path) { CV_output(path, Tnow, true); }
void IncFluid::Output_field_reduced(const std::string& path, const int Nreduced[3]) { CV_output(path, Nreduced, Tnow, true); }

double IncFluid::Init_cond_reduced(const std::string& path, const int Nreduced[3], bool reality, int snapshot)
{
    const double t = CV_input(path, Nreduced, snapshot);
    if (reality)
        for (DField* f : {&V1, &V2, &V3}) check(th_satisfy_reality(plan_, f->data()), "Satisfy_reality_condition_field");
    if (grid.alias_switch == "DEALIAS") CV_Dealias();
    return t;
}

double IncFluid::Init_cond(const std::string& path, bool reality, int snapshot)
{
    const double t = CV_input(path, snapshot);
    if (reality)
        for (DField* f : {&V1, &V2, &V3}) check(th_satisfy_reality(plan_, f->data()), "Satisfy_reality_condition_field");
    if (grid.alias_switch == "DEALIAS") CV_Dealias();
    return t;
}

// Blitz++'s stream output of arrays (blitz/array/io.cc), as the reference's shell.d prints them: a 1-D array as
// its extent, a newline, " [ ", every value in setw(9) followed by a space, a newline and two spaces after
// every 7th value, then " ]"; a 2-D array as "rows x columns", a newline, "[ ", the rows (each value in setw(9)
// and a space, a newline and two spaces after every 7th value of a row, a newline and two spaces between rows)
// and "]" with a newline.
static void blitz_out_1d(std::ostream& os, const double* a, int n)
{
    os << n << std::endl << " [ ";
    for (int i = 0; i < n; i++) {
        os << std::setw(9) << a[i] << " ";
        if (!((i + 1) % 7)) os << std::endl << "  ";
    }
    os << " ]";
}
static void blitz_out_2d(std::ostream& os, const double* a, int ld, int r0, int r1, int c0, int c1)
{
    os << (r1 - r0 + 1) << " x " << (c1 - c0 + 1) << std::endl << "[ ";
    for (int i = r0; i <= r1; i++) {
        for (int j = c0; j <= c1; j++) {
            os << std::setw(9) << a[(size_t)i * ld + j] << " ";
            if (!((j + 1 - c0) % 7)) os << std::endl << "  ";
        }
        if (i != r1) os << std::endl << "  ";
    }
    os << "]" << std::endl;
}

void IncFluid::Open_output_files(const std::string& data_dir, const std::string& prefix_str)
{
    const std::string out = data_dir + "/out/";
    global_file.open(out + "glob.d");
    spectrum_file.open(out + "spectrum.d");
    flux_file.open(out + "flux.d");
    shell_to_shell_file.open(out + "shell.d");
    if (!global_file || !spectrum_file || !flux_file || !shell_to_shell_file)
        throw std::runtime_error("Open_output_files: cannot open files in " + out);
    global_file.setf(std::ios::fixed, std::ios::floatfield);
    global_file.precision(12);                        // MY_PRECISION extern_def_vars.h:78
    std::ofstream& X = shell_to_shell_file;            // Output_prefix file_operations.cc:108-121
    X << prefix_str << std::endl;
    X << "%% Grid size " << grid.N[1] << " " << grid.N[2] << " " << grid.N[3] << std::endl;
    X << "%% Basis, Integration scheme:  " << grid.basis_type << " " << integ_scheme << std::endl;
    X << "%% Tinit, Tfinal, Tdt_fixed:  " << Tinit_ << "    " << Tfinal << "    " << Tdt_fixed << std::endl << std::endl;
    X << "%% globalvar_anisotropy_switch: " << anisotropy_switch << std::endl;
    X << "%% globalvar_waveno_switch: " << waveno_switch << std::endl;
    X << "%% Shellradius = ";
    blitz_out_1d(X, shell_radius.data() + 1, no_shells);
    X << std::endl;
    if (no_ring_shells >= 3) {
        ring_to_ring_file.open(out + "ring_to_ring.d");
        if (!ring_to_ring_file) throw std::runtime_error("Open_output_files: cannot open ring_to_ring.d in " + out);
    }
    if (no_cylinder_shells >= 3) {
        cylinder_ring_to_ring_file.open(out + "cyl_ring_to_ring.d");
        if (!cylinder_ring_to_ring_file) throw std::runtime_error("Open_output_files: cannot open cyl_ring_to_ring.d in " + out);
    }
}

// Blitz++'s stream output of an N-d array: the extents joined by " x ", a newline, "[ ", the elements in
// row-major order, each in setw(9) followed by a space, a newline and two spaces after every 7th, then "]" and
// a newline.  Here the 4-D sub-table [1..n-1][1..m][1..n-1][1..m] of a (n+1, m+1, n+1, m+1) table.
static void blitz_out_4d(std::ostream& os, const std::vector<double>& t, int n, int m)
{
    const size_t M = (size_t)(n + 1) * (m + 1);
    os << (n - 1) << " x " << m << " x " << (n - 1) << " x " << m << std::endl << "[ ";
    long p = 0;
    for (int a = 1; a < n; a++)
        for (int b = 1; b <= m; b++)
            for (int c = 1; c < n; c++)
                for (int d = 1; d <= m; d++) {
                    os << std::setw(9) << t[(size_t)(a * (m + 1) + b) * M + (size_t)(c * (m + 1) + d)] << " ";
                    if (!(++p % 7)) os << std::endl << "  ";
                }
    os << "]" << std::endl;
}

// the receiver-ring force feed Ring_mult_all(A, B) (compute_force_feed.cc:162-250), rows 1..n-1 x cols 1..m;
// A = nullptr prints zeros (an unforced run)
void IncFluid::ring_feed(bool cyl, std::ostream& os, void* const* A, void* const* B, int ncomp)
{
    const std::vector<double>& r = cyl ? cylinder_shell_radius : ring_shell_radius;
    const std::vector<double>& e = cyl ? cylinder_kpll_array_tr : sector_angle_ring_tr;
    const int n = cyl ? no_cylinder_shells : no_ring_shells, m = cyl ? no_cylinder_kpll_slabs : no_sectors_ring_tr;
    std::vector<double> f((size_t)(n + 1) * (m + 1), 0.0);
    if (A) check(th_ring_mult_all(plan_, cyl ? 1 : 0, anisotropy_axis, A, B, ncomp, r.data(), n, e.data(), m, f.data()),
                 "Compute_force_feed_ring");
    blitz_out_2d(os, f.data(), m + 1, 1, n - 1, 1, m);
}

// One output time of ring_to_ring.d / cyl_ring_to_ring.d: "%% Time = t", each named 4-D table, then each named
// force feed (output_ET.cc:427-760).
void IncFluid::ring_out(bool cyl, const std::vector<std::pair<const char*, const std::vector<double>*>>& tables,
                        const std::vector<RingFeed>& feeds)
{
    std::ofstream& os = cyl ? cylinder_ring_to_ring_file : ring_to_ring_file;
    const int n = cyl ? no_cylinder_shells : no_ring_shells, m = cyl ? no_cylinder_kpll_slabs : no_sectors_ring_tr;
    os << "%% Time = " << Tnow << std::endl;
    for (size_t k = 0; k < tables.size(); k++) {
        os << tables[k].first << std::endl;
        blitz_out_4d(os, *tables[k].second, n, m);
        os << std::endl;
        if (k + 1 == tables.size()) os << std::endl;
    }
    for (const RingFeed& f : feeds) {
        os << f.name << std::endl;
        ring_feed(cyl, os, f.A, f.B, f.ncomp);
        os << std::endl << std::endl;
    }
}

void IncFluid::Output_ring_to_ring() { ring_out_ns(false); }
void IncFluid::Output_cylinder_ring_to_ring() { ring_out_ns(true); }
void IncFluid::ring_out_ns(bool cyl)
{
    if ((cyl ? no_cylinder_shells : no_ring_shells) < 3) return;   // ET_anisotropic_*_switch = 0
    if (cyl) Compute_cylinder_ring_tr();
    else Compute_ring_tr();
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    Force_field F(*this);                               // zero before a forced step and for unforced runs
    void* Fp[3] = {F.f[0].data(), F.f[1].data(), F.f[2].data()};
    ring_out(cyl, {{"%% U to U: ", cyl ? &cylinder_ring_to_ring_self : &ring_to_ring_self}},
             {RingFeed{"%% Force feed: ", has_force_ ? Fp : nullptr, V, 3}});
}

// Compute_ring_tr_RB / _scalar compute_ring_tr.cc:90-170 (PRLARGE / ULARGE): U to U, then T to T with helper U
void IncFluid::Compute_ring_tr(IncSF& T)
{
    Compute_ring_tr();
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    void* Tf[3] = {T.F.data(), nullptr, nullptr};
    ring_table(false, V, Tf, Tf, 1, -1.0, no_ring_shells - 1, ring_to_ring_SF);
}
void IncFluid::Compute_cylinder_ring_tr(IncSF& T)
{
    Compute_cylinder_ring_tr();
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    void* Tf[3] = {T.F.data(), nullptr, nullptr};
    ring_table(true, V, Tf, Tf, 1, -1.0, no_cylinder_shells - 1, cylinder_ring_to_ring_SF);
}

// Output_ring_to_ring_scalar output_ET.cc:468-499 and its cylinder form: U to U, T to T, then the force feeds
// (Force, V) and (T.Force, T) with the last RK stage's buoyancy forces of the RB step
void IncFluid::Output_ring_to_ring(IncSF& T) { ring_out_scalar(false, T); }
void IncFluid::Output_cylinder_ring_to_ring(IncSF& T) { ring_out_scalar(true, T); }
void IncFluid::ring_out_scalar(bool cyl, IncSF& T)
{
    if ((cyl ? no_cylinder_shells : no_ring_shells) < 3) return;
    if (cyl) Compute_cylinder_ring_tr(T);
    else Compute_ring_tr(T);
    DField fv[3], ft;
    rb_forces(fv, ft);
    DField z(elems_, esize_);
    void* F[3] = {fv[0].data(), fv[1].data(), fv[2].data()};
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    void* A[3] = {ft.data(), z.data(), z.data()};
    void* B[3] = {T.F.data(), z.data(), z.data()};
    ring_out(cyl, {{"%% U to U: ", cyl ? &cylinder_ring_to_ring_self : &ring_to_ring_self},
                   {"%% T to T: ", cyl ? &cylinder_ring_to_ring_SF : &ring_to_ring_SF}},
             {RingFeed{"%% Force feed (Fv * V): ", F, V, 3}, RingFeed{"%% Force feed (Ftheta * theta): ", A, B, 1}});
}

// Output_ring_to_ring(IncVF& W) output_ET.cc:512-555 and its cylinder form: U to U, W to W, U to W, Zp to Zp,
// Zm to Zm, then the force feeds (Force, U) and (W.Force, W) (no mean field B0, so no B0 transfer block)
void IncFluid::Output_ring_to_ring(IncVF& W) { ring_out_mhd(false, W); }
void IncFluid::Output_cylinder_ring_to_ring(IncVF& W) { ring_out_mhd(true, W); }
void IncFluid::ring_out_mhd(bool cyl, IncVF& W)
{
    if ((cyl ? no_cylinder_shells : no_ring_shells) < 3) return;
    if (cyl) Compute_cylinder_ring_tr(W);
    else Compute_ring_tr(W);
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    DField fu[3], fw[3];
    if (mhd_forced_) mhd_forces(fu, fw);
    void* Fu[3] = {fu[0].data(), fu[1].data(), fu[2].data()};
    void* Fw[3] = {fw[0].data(), fw[1].data(), fw[2].data()};
    ring_out(cyl,
             {{"%% U to U: ", cyl ? &cylinder_ring_to_ring_self : &ring_to_ring_self},
              {"%% W to W: ", cyl ? &W.cylinder_ring_to_ring_self : &W.ring_to_ring_self},
              {"%% U to W: ", cyl ? &cylinder_ring_to_ring_VF : &ring_to_ring_VF},
              {"%% Zp to Zp: ", cyl ? &cylinder_ring_to_ring_Elsasser : &ring_to_ring_Elsasser},
              {"%% Zm to Zm: ", cyl ? &W.cylinder_ring_to_ring_Elsasser : &W.ring_to_ring_Elsasser}},
             {RingFeed{"%% Force feed (Fv * V): ", mhd_forced_ ? Fu : nullptr, U, 3},
              RingFeed{"%% Force feed (Fw * W): ", mhd_forced_ ? Fw : nullptr, B, 3}});
}

void IncFluid::Output_shell_to_shell()
{
    Compute_shell_tr();
    shell_to_shell_file << "%% Time = " << Tnow << std::endl;
    shell_to_shell_file << "%% U to U: " << std::endl;
    blitz_out_2d(shell_to_shell_file, shelltoshell_self.data(), no_shells + 1, 1, no_shells - 1, 1, no_shells);
    shell_to_shell_file << std::endl << std::endl << std::endl;
}

void IncFluid::Output_shell_to_shell(IncSF& T)
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    shelltoshell_SF.assign(shelltoshell_self.size(), 0.0);
    check(th_shell_tr_scalar(plan_, V, T.F.data(), shell_radius.data(), no_shells, shelltoshell_self.data(),
                             shelltoshell_SF.data()), "Compute_shell_tr(T)");
    const int ld = no_shells + 1;
    shell_to_shell_file << "%% Time = " << Tnow << std::endl;
    shell_to_shell_file << "%% U to U: " << std::endl;
    blitz_out_2d(shell_to_shell_file, shelltoshell_self.data(), ld, 1, no_shells - 1, 1, no_shells);
    shell_to_shell_file << std::endl << "%% T to T: " << std::endl;
    blitz_out_2d(shell_to_shell_file, shelltoshell_SF.data(), ld, 1, no_shells - 1, 1, no_shells);
    shell_to_shell_file << std::endl << std::endl << std::endl;
}

void IncFluid::Output_shell_to_shell(IncVF& W)
{
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    const int ld = no_shells + 1;
    const size_t M = (size_t)ld * ld;
    shelltoshell_mhd.assign(5 * M, 0.0);
    check(th_shell_tr_mhd(plan_, U, B, shell_radius.data(), no_shells, shelltoshell_mhd.data()), "Compute_shell_tr(W)");
    std::copy(shelltoshell_mhd.begin(), shelltoshell_mhd.begin() + M, shelltoshell_self.begin());
    if (W.shelltoshell_self.size() == M)
        std::copy(shelltoshell_mhd.begin() + M, shelltoshell_mhd.begin() + 2 * M, W.shelltoshell_self.begin());
    const char* names[5] = {"%% U to U: ", "%% W to W: ", "%% U to W: ", "%% Zp to Zp: ", "%% Zm to Zm: "};
    shell_to_shell_file << "%% Time = " << Tnow << std::endl;
    for (int k = 0; k < 5; k++) {
        shell_to_shell_file << names[k] << std::endl;
        blitz_out_2d(shell_to_shell_file, shelltoshell_mhd.data() + k * M, ld, 1, no_shells - 1, 1, no_shells);
        shell_to_shell_file << std::endl;
    }
    shell_to_shell_file << std::endl << std::endl << std::endl << std::endl;
}

void IncFluid::Output_global()
{
    CV_Compute_totalenergy_diss();
    double h[4], entropy;
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    check(th_helicity(plan_, V, h), "CV_Compute_total_helicity");
    check(th_entropy(plan_, V, &entropy), "CV_Compute_entropy");
    double mean[3];
    mean_modes(V, esize_, mean);
    const double kmax = th_max_radius_inside(plan_);
    const double total_diss = dissipation_coefficient * CV_total_dissipation;
    const double kolm_scale_u = sqrt(sqrt((pow(dissipation_coefficient, 3) / total_diss)));
    const double kmax_eta = kmax * kolm_scale_u;
    const double Rlambda = 2 * CV_total_energy * sqrt(15 / (dissipation_coefficient * total_diss));
    global_file << Tnow << "\t" << CV_total_energy << " " << CV_total_dissipation << " " << total_diss << "    " << h[0]
                << " " << h[1] << "    " << h[2] << " " << h[3] << "    " << entropy << "    " << kmax_eta << "    "
                << dissipation_coefficient << "    " << Rlambda << "    " << Tdt << "    " << mean[0] << " " << mean[1]
                << " " << mean[2] << std::endl;
}

void IncFluid::Output_shell_spectrum()
{
    spectrum_file << "%% Time = " << Tnow << std::endl;
    const int n = shell_ek_size();
    std::vector<double> ek[3], dk(n, 0.0), t(n), hk[3];
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    for (int c = 0; c < 3; c++) {
        ek[c].resize(n);
        hk[c].resize(n);
        check(th_shell_spectrum(plan_, V[c], nullptr, 0, ek[c].data(), nullptr), "CV_Compute_shell_spectrum");
        check(th_shell_spectrum(plan_, V[c], nullptr, 2, t.data(), nullptr), "CV_Compute_shell_spectrum");
        for (int i = 0; i < n; i++) dk[i] += t[i];
    }
    for (int i = 0; i < n; i++) dk[i] *= 2.0;
    check(th_shell_spectrum_helicity(plan_, V, hk[0].data(), hk[1].data(), hk[2].data()), "CV_Compute_shell_spectrum_helicity");
    // Compute_force_shell_spectrum force_spectrum.cc:48-53: 2 sum_c Re(V_c conj F_c) per shell (0 unforced)
    std::vector<double> fk(n, 0.0);
    if (has_force_) {
        Force_field F(*this);
        for (int c = 0; c < 3; c++) {
            check(th_shell_spectrum(plan_, V[c], F.f[c].data(), 0, t.data(), nullptr), "Compute_force_shell_spectrum");
            for (int i = 0; i < n; i++) fk[i] += t[i];
        }
        for (int i = 0; i < n; i++) fk[i] *= 2.0;
    }
    for (int i = 0; i < n; i++)
        spectrum_file << i << "\t" << ek[0][i] << " " << ek[1][i] << " " << ek[2][i] << "\t"
                      << dissipation_coefficient * dk[i] << "\t" << hk[0][i] << "  " << hk[1][i] << "  " << hk[2][i] << "\t"
                      << fk[i] << "\t" << std::endl;
    spectrum_file << std::endl;
}

void IncFluid::Output_flux()
{
    flux_file << "%% Time = " << Tnow << std::endl;
    Compute_flux();
    // Compute_force_feed_shell compute_force_feed.cc:70-76: Shell_mult_all(F, V) over shell_radius
    std::vector<double> feed(shell_radius.size(), 0.0);
    if (has_force_) {
        Force_field F(*this);
        void* Fp[3] = {F.f[0].data(), F.f[1].data(), F.f[2].data()};
        void* V[3] = {V1.data(), V2.data(), V3.data()};
        check(th_shell_mult_all(plan_, Fp, V, shell_radius.data(), (int)shell_radius.size(), feed.data()),
              "Compute_force_feed_shell");
    }
    for (int i = 1; i <= no_spheres; i++)
        flux_file << i << " " << sphere_radius[i] << " " << flux_self[i] << "  " << (i < (int)feed.size() ? feed[i] : 0.0)
                  << "     " << std::endl;
    flux_file << std::endl;
}

// ---------------------------------------------------------------- IncFluid with W (MHD)

double IncFluid::Get_dt(IncVF& W)
{
    CV_Compute_totalenergy_diss();
    W.CV_Compute_totalenergy_diss();
    const double kmax = th_max_radius_inside(plan_);
    const double dx = (2 * M_PI) / kmax;
    double dt = dx / (sqrt(CV_total_energy + W.CV_total_energy) * 20);
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    double b0[3];
    mean_modes(B, esize_, b0);
    const double b0sqr = b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2];
    if (b0sqr > 1e-15) dt = std::min(dt, dx / (sqrt(b0sqr) * 20));
    return std::min(dt, Tdt_fixed);
}

void IncFluid::Time_advance(IncVF& W)
{
    if (has_force_) throw std::runtime_error("Time_advance(W): Set_force(proc, ...) forces NS runs; use Set_force(W, ...)");
    th_phys ph{dissipation_coefficient, W.dissipation_coefficient, 0, 0, 0};
    void* f[6] = {V1.data(), V2.data(), V3.data(), W.V1.data(), W.V2.data(), W.V3.data()};
    const int sch = integ_scheme == "EULER" ? 0 : integ_scheme == "RK2" ? 1 : 2;
    if (mhd_forced_) {
        void* const U[3] = {f[0], f[1], f[2]};
        void* const B[3] = {f[3], f[4], f[5]};
        if (!mlast_[0].data())
            for (int c = 0; c < 6; c++) mlast_[c] = DField(elems_, esize_);
        void* const last[6] = {mlast_[0].data(), mlast_[1].data(), mlast_[2].data(),
                               mlast_[3].data(), mlast_[4].data(), mlast_[5].data()};
        check(th_step_plain_forced_mhd(plan_, sch, U, B, Tdt, &ph, &mhd_force_, last), "th_step_plain_forced_mhd");
        mhd_stepped_ = true;
        check(th_sync(plan_), "th_sync");
        return;
    }
    if (fused_) {
        if (!mstate_) check(th_state_create(plan_, &mstate_, TH_MHD), "th_state_create");
        for (int i = 0; i < 6; i++) check(th_state_upload(mstate_, i, f[i], 0), "th_state_upload");
        check(th_step(plan_, mstate_, Tdt, &ph, 1), "th_step");
        for (int i = 0; i < 6; i++) check(th_state_download(mstate_, i, f[i], 0), "th_state_download");
        check(th_sync(plan_), "th_sync");
        return;
    }
    check(th_step_plain(plan_, TH_MHD, sch, f, Tdt, &ph), "th_step_plain");
    check(th_sync(plan_), "th_sync");
}

// Runs that step on the fused MHD state (its fields are 2/3-dealiased after every step) take the fused
// giver-shell pipeline (th_flux_mhd_fused); otherwise the plain sphere loop of the reference (th_flux_mhd).
void IncFluid::Compute_flux(IncVF& W)
{
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    flux_mhd.assign((size_t)9 * (no_spheres + 1), 0.0);
    if (mstate_ && no_spheres >= 2) {
        void* f[6] = {U[0], U[1], U[2], B[0], B[1], B[2]};
        for (int i = 0; i < 6; i++) check(th_state_upload(mstate_, i, f[i], 0), "th_state_upload");
        check(th_flux_mhd_fused(plan_, mstate_, sphere_radius.data(), no_spheres, flux_mhd.data()), "Compute_flux(W)");
        return;
    }
    check(th_flux_mhd(plan_, U, B, sphere_radius.data(), no_spheres, flux_mhd.data()), "Compute_flux(W)");
}

void IncFluid::Output_global(IncVF& W)
{
    CV_Compute_totalenergy_diss();
    W.CV_Compute_totalenergy_diss();
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    double Hc = 0.0;
    for (int c = 0; c < 3; c++) {        // Get_cross_helicity hc.cc:48-53
        double e;
        check(th_energy(plan_, U[c], B[c], &e), "Get_cross_helicity");
        Hc += e;
    }
    double hu[4], hw[4], su, sw;
    check(th_helicity(plan_, U, hu), "CV_Compute_total_helicity");
    check(th_helicity(plan_, B, hw), "W.CV_Compute_total_helicity");
    check(th_entropy(plan_, U, &su), "CV_Compute_entropy");
    check(th_entropy(plan_, B, &sw), "W.CV_Compute_entropy");
    double mu[3], mb[3];
    mean_modes(U, esize_, mu);
    mean_modes(B, esize_, mb);
    const double nu = dissipation_coefficient, eta = W.dissipation_coefficient;
    const double kmax = th_max_radius_inside(plan_);
    const double total_diss = nu * CV_total_dissipation;
    const double kmax_eta1 = kmax * sqrt(sqrt(pow(nu, 3) / total_diss));
    const double tempvar = nu / eta;
    const double kmax_eta2 = kmax_eta1 * sqrt(sqrt(tempvar)) / tempvar;
    const double Rlambda = 2 * CV_total_energy * sqrt(15 / (nu * total_diss));
    global_file << Tnow << "\t" << CV_total_energy << " " << W.CV_total_energy << "   " << CV_total_dissipation << " "
                << W.CV_total_dissipation << " " << total_diss << " " << eta * W.CV_total_dissipation << "    " << Hc
                << "    " << hu[0] << " " << hu[1] << " " << hw[0] << " " << hw[1] << "    " << hu[2] << " " << hu[3]
                << "    " << hw[2] << " " << hw[3] << "    " << su << " " << sw << "   " << kmax_eta1 << " " << kmax_eta2
                << "    " << nu << " " << eta << "    " << Rlambda << "    " << Tdt << "    " << mu[0] << " " << mu[1]
                << " " << mu[2] << "\t" << mb[0] << " " << mb[1] << " " << mb[2] << std::endl;
}

// Columns as output_energy.cc:478-497.  shell_ek_cross is printed but never filled on this path in the
// reference (Compute_cross_helicity_shell_spectrum hc.cc:67-71 has no caller), so it is 0; the force spectra
// of an unforced MHD run are 0.
void IncFluid::Output_shell_spectrum(IncVF& W)
{
    spectrum_file << "%% Time = " << Tnow << std::endl;
    const int n = shell_ek_size();
    void* U[3] = {V1.data(), V2.data(), V3.data()};
    void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
    std::vector<double> eu[3], ew[3], hu[3], hw[3], du(n, 0.0), dw(n, 0.0), t(n);
    for (int c = 0; c < 3; c++) {
        eu[c].resize(n);
        ew[c].resize(n);
        hu[c].resize(n);
        hw[c].resize(n);
        check(th_shell_spectrum(plan_, U[c], nullptr, 0, eu[c].data(), nullptr), "CV_Compute_shell_spectrum");
        check(th_shell_spectrum(plan_, B[c], nullptr, 0, ew[c].data(), nullptr), "W.CV_Compute_shell_spectrum");
        check(th_shell_spectrum(plan_, U[c], nullptr, 2, t.data(), nullptr), "CV_Compute_shell_spectrum");
        for (int i = 0; i < n; i++) du[i] += 2.0 * t[i];
        check(th_shell_spectrum(plan_, B[c], nullptr, 2, t.data(), nullptr), "W.CV_Compute_shell_spectrum");
        for (int i = 0; i < n; i++) dw[i] += 2.0 * t[i];
    }
    check(th_shell_spectrum_helicity(plan_, U, hu[0].data(), hu[1].data(), hu[2].data()), "CV_Compute_shell_spectrum_helicity");
    check(th_shell_spectrum_helicity(plan_, B, hw[0].data(), hw[1].data(), hw[2].data()), "W.CV_Compute_shell_spectrum_helicity");
    // Compute_force_shell_spectrum force_spectrum.cc:48-53 of U and of W (0 for an unforced run)
    std::vector<double> fku(n, 0.0), fkw(n, 0.0);
    if (mhd_forced_) {
        DField fu[3], fw[3];
        mhd_forces(fu, fw);
        for (int c = 0; c < 3; c++) {
            check(th_shell_spectrum(plan_, U[c], fu[c].data(), 0, t.data(), nullptr), "Compute_force_shell_spectrum");
            for (int i = 0; i < n; i++) fku[i] += 2.0 * t[i];
            check(th_shell_spectrum(plan_, B[c], fw[c].data(), 0, t.data(), nullptr), "W.Compute_force_shell_spectrum");
            for (int i = 0; i < n; i++) fkw[i] += 2.0 * t[i];
        }
    }
    const double nu = dissipation_coefficient, eta = W.dissipation_coefficient;
    for (int i = 0; i < n; i++)
        spectrum_file << i << "\t" << eu[0][i] << " " << eu[1][i] << " " << eu[2][i] << "\t" << ew[0][i] << " " << ew[1][i]
                      << " " << ew[2][i] << "\t" << nu * du[i] << " " << eta * dw[i] << "    " << 0.0 << "    " << hu[0][i]
                      << " " << hu[1][i] << " " << hu[2][i] << "\t" << hw[0][i] << " " << hw[1][i] << " " << hw[2][i]
                      << "\t" << fku[i] << " " << fkw[i] << std::endl;
    spectrum_file << std::endl;
}

// Columns as output_ET.cc:139-156: flux_self, flux_VF_in_out, flux_VF_in_in of U, then of W and
// W.flux_VF_out_out, flux_Elsasser of U and W, the force feeds.  Compute_flux(W) stores VF_out_out in U
// (compute_flux.cc:232), so W's printed value stays 0; the force feeds of an unforced run are 0.
void IncFluid::Output_flux(IncVF& W)
{
    flux_file << "%% Time = " << Tnow << std::endl;
    Compute_flux(W);
    const size_t s = (size_t)no_spheres + 1;
    auto F = [&](int row, int i) { return flux_mhd[row * s + i]; };
    // Compute_force_feed_shell(W) compute_force_feed.cc:130-139: Shell_mult_all(F, U) and (W.Force, W)
    std::vector<double> feu(shell_radius.size(), 0.0), few(shell_radius.size(), 0.0);
    if (mhd_forced_) {
        DField fu[3], fw[3];
        mhd_forces(fu, fw);
        void* Fu[3] = {fu[0].data(), fu[1].data(), fu[2].data()};
        void* Fw[3] = {fw[0].data(), fw[1].data(), fw[2].data()};
        void* U[3] = {V1.data(), V2.data(), V3.data()};
        void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
        check(th_shell_mult_all(plan_, Fu, U, shell_radius.data(), (int)shell_radius.size(), feu.data()),
              "Compute_force_feed_shell");
        check(th_shell_mult_all(plan_, Fw, B, shell_radius.data(), (int)shell_radius.size(), few.data()),
              "Compute_force_feed_shell(W)");
    }
    auto at = [](const std::vector<double>& v, int i) { return i < (int)v.size() ? v[i] : 0.0; };
    for (int i = 1; i <= no_spheres; i++)
        flux_file << i << " " << sphere_radius[i] << " " << F(0, i) << " " << F(2, i) << " " << F(3, i) << "   "
                  << F(1, i) << " " << F(4, i) << " " << F(5, i) << " " << 0.0 << "    " << F(7, i) << " " << F(8, i)
                  << "    " << at(feu, i) << " " << at(few, i) << "    " << std::endl;
    flux_file << std::endl;
}

// ---------------------------------------------------------------- IncFluid with T (RB)
double IncFluid::Get_dt(IncSF&) { return Get_dt(); }

void IncFluid::Time_advance(IncSF& T)
{
    if (has_force_) throw std::runtime_error("Time_advance(T): Set_force applies to NS runs only");
    th_phys ph{dissipation_coefficient, 0, T.diffusion_coefficient, Rayleigh, Prandtl};
    void* f[4] = {V1.data(), V2.data(), V3.data(), T.F.data()};
    if (grid.basis_type == "SCFT" && integ_scheme == "RK4") {
        if (!rstate_) check(th_state_create(plan_, &rstate_, TH_RB), "th_state_create");
        for (int i = 0; i < 4; i++) check(th_state_upload(rstate_, i, f[i], 0), "th_state_upload");
        check(th_step(plan_, rstate_, Tdt, &ph, 1), "th_step");
        for (int i = 0; i < 4; i++) check(th_state_download(rstate_, i, f[i], 0), "th_state_download");
        check(th_sync(plan_), "th_sync");
        rb_stepped_ = true;
        return;
    }
    const int sch = integ_scheme == "EULER" ? 0 : integ_scheme == "RK2" ? 1 : 2;
    check(th_step_plain(plan_, TH_RB, sch, f, Tdt, &ph), "th_step_plain");
    check(th_sync(plan_), "th_sync");
}

double IncFluid::Get_Nusselt_no(IncSF& T)
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    double nu;
    check(th_nusselt(plan_, V, T.F.data(), Rayleigh, Prandtl, &nu), "Get_Nusselt_no");
    return nu;
}

void IncFluid::Output_global(IncSF& T)
{
    CV_Compute_totalenergy_diss();
    T.CS_Compute_totalenergy_diss();
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    double h[4], sv;
    check(th_entropy(plan_, V, &sv), "CV_Compute_entropy");
    T.CS_Compute_entropy();
    check(th_helicity(plan_, V, h), "CV_Compute_total_helicity");
    const double nusselt_no = Get_Nusselt_no(T);
    double mv[3], mt[3];
    void* Tp[3] = {T.F.data(), T.F.data(), T.F.data()};
    mean_modes(V, esize_, mv);
    mean_modes(Tp, esize_, mt);
    const double nu = dissipation_coefficient, kappa = T.diffusion_coefficient;
    const double kmax = th_max_radius_inside(plan_);
    const double total_diss = nu * CV_total_dissipation;
    const double kmax_eta1 = kmax * sqrt(sqrt(pow(nu, 3) / total_diss));
    const double tempvar = nu / kappa;
    const double kmax_eta2 = kmax_eta1 * sqrt(sqrt(tempvar)) / tempvar;
    const double Rlambda = 2 * CV_total_energy * sqrt(15 / (nu * total_diss));
    global_file << Tnow << "\t" << CV_total_energy << " " << T.CS_total_energy << "\t" << CV_total_dissipation << " "
                << T.CS_total_dissipation << "\t" << total_diss << " " << kappa * T.CS_total_dissipation << "    "
                << nusselt_no << "    " << h[0] << " " << h[1] << "    " << h[2] << " " << h[3] << "    " << sv << " "
                << T.CS_entropy << "   " << kmax_eta1 << " " << kmax_eta2 << "    " << nu << " " << kappa << "    "
                << Rlambda << "    " << Tdt << "    " << mv[0] << " " << mv[1] << " " << mv[2] << "\t" << mt[0]
                << std::endl;
}

// Compute_force_RB compute_force_RB.cc:51-105 (PRLARGE, ULARGE) of the last RK stage of Time_advance(T), which is
// what Force1..3 and T.Force hold when RB_slip_main.cc:363-402 writes its outputs: Force1 = T, Force2 = Force3 = 0,
// T.Force = V1, all of the stage field and dealiased.  The fused step keeps that field in its stage buffer; before
// the first step (and on the plain stepper, which does not report it) the forces are left zero.
void IncFluid::rb_forces(DField fv[3], DField& ft)
{
    for (int c = 0; c < 3; c++) fv[c] = DField(elems_, esize_);
    ft = DField(elems_, esize_);
    if (!rb_stepped_ || !rstate_) return;
    check(th_state_download_buffer(rstate_, 1, 3, fv[0].data(), 0), "th_state_download_buffer");
    check(th_state_download_buffer(rstate_, 1, 0, ft.data(), 0), "th_state_download_buffer");
}

// Columns as output_energy.cc:429-446.  The scalar dissipation column prints IncFluid's own CS_shell_dissk
// (inherited through NLIN : CSF, Nlin.h:51), which nothing computes, so it is 0 (Csf.cc:120-121).
void IncFluid::Output_shell_spectrum(IncSF& T)
{
    spectrum_file << "%% Time = " << Tnow << std::endl;
    const int n = shell_ek_size();
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    std::vector<double> ek[3], hk[3], cr[3], et(n), dk(n, 0.0), t(n), fvk(n), ftk(n);
    for (int c = 0; c < 3; c++) {
        ek[c].resize(n);
        hk[c].resize(n);
        cr[c].resize(n);
        check(th_shell_spectrum(plan_, V[c], nullptr, 0, ek[c].data(), nullptr), "CV_Compute_shell_spectrum");
        check(th_shell_spectrum(plan_, V[c], nullptr, 2, t.data(), nullptr), "CV_Compute_shell_spectrum");
        for (int i = 0; i < n; i++) dk[i] += 2.0 * t[i];
        // Compute_cross_vT_shell_spectrum hc.cc:112-124
        check(th_shell_spectrum(plan_, V[c], T.F.data(), 0, cr[c].data(), nullptr), "Compute_cross_vT_shell_spectrum");
    }
    check(th_shell_spectrum(plan_, T.F.data(), nullptr, 0, et.data(), nullptr), "T.CS_Compute_shell_spectrum");
    check(th_shell_spectrum_helicity(plan_, V, hk[0].data(), hk[1].data(), hk[2].data()), "CV_Compute_shell_spectrum_helicity");
    DField fv[3], ft;
    rb_forces(fv, ft);
    // Compute_force_shell_spectrum force_spectrum.cc:48-53 (Force2 = Force3 = 0) and IncSF's :64-68
    check(th_shell_spectrum(plan_, V[0], fv[0].data(), 0, fvk.data(), nullptr), "Compute_force_shell_spectrum");
    check(th_shell_spectrum(plan_, T.F.data(), ft.data(), 0, ftk.data(), nullptr), "T.Compute_force_shell_spectrum");
    const double nu = dissipation_coefficient;
    for (int i = 0; i < n; i++)
        spectrum_file << i << "\t" << ek[0][i] << " " << ek[1][i] << " " << ek[2][i] << "\t" << et[i] << "\t" << nu * dk[i]
                      << " " << T.diffusion_coefficient * 0.0 << "\t" << cr[0][i] << " " << cr[1][i] << " " << cr[2][i]
                      << "\t" << hk[0][i] << "  " << hk[1][i] << "  " << hk[2][i] << "\t" << 2 * fvk[i] << " " << 2 * ftk[i]
                      << std::endl;
    spectrum_file << std::endl;
}

void IncFluid::Compute_flux(IncSF& T)
{
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    std::vector<double> out((size_t)2 * (no_spheres + 1), 0.0);
    check(th_flux_scalar(plan_, V, T.F.data(), sphere_radius.data(), no_spheres, out.data()), "Compute_flux_RB");
    flux_self.assign(out.begin(), out.begin() + no_spheres + 1);
    flux_SF.assign(out.begin() + no_spheres + 1, out.end());
}

// Columns as output_ET.cc:96-104: flux_self, flux_SF, then the force feeds of Compute_force_feed_shell_scalar
// compute_force_feed.cc:94-100 (Shell_mult_all(Force, V) and Shell_mult_all(T.Force, T.F) over shell_radius).
void IncFluid::Output_flux(IncSF& T)
{
    flux_file << "%% Time = " << Tnow << std::endl;
    Compute_flux(T);
    DField fv[3], ft;
    rb_forces(fv, ft);
    DField z(elems_, esize_);
    std::vector<double> feedV(shell_radius.size(), 0.0), feedT(shell_radius.size(), 0.0);
    void* F[3] = {fv[0].data(), fv[1].data(), fv[2].data()};
    void* V[3] = {V1.data(), V2.data(), V3.data()};
    void* A[3] = {ft.data(), z.data(), z.data()};
    void* B[3] = {T.F.data(), z.data(), z.data()};
    check(th_shell_mult_all(plan_, F, V, shell_radius.data(), (int)shell_radius.size(), feedV.data()), "Compute_force_feed_shell");
    check(th_shell_mult_all(plan_, A, B, shell_radius.data(), (int)shell_radius.size(), feedT.data()), "Compute_force_feed_shell_scalar");
    for (int i = 1; i <= no_spheres; i++)
        flux_file << i << " " << sphere_radius[i] << " " << flux_self[i] << " " << flux_SF[i] << "  "
                  << (i < (int)feedV.size() ? feedV[i] : 0.0) << " " << (i < (int)feedT.size() ? feedT[i] : 0.0) << "     "
                  << std::endl;
    flux_file << std::endl;
}

// Force1..3 as the reference holds them when its time loop writes the outputs (Ifluid_main.cc:278-298):
// Time_advance recomputes the force at every RK stage (Compute_force_TO_rhs), so they hold the force of the
// last stage's field; before the first step they are still zero.
IncFluid::Force_field::Force_field(IncFluid& u)
{
    for (int c = 0; c < 3; c++) f[c] = DField(u.elems_, u.esize_);
    if (!u.stepped_) return;                                   // DField starts zeroed
    DField stage[3];
    void* src[3];
    if (u.forced_) {
        for (int c = 0; c < 3; c++) src[c] = u.last_[c].data();
    } else {
        // fused step: the stage field buffer still holds the last stage's input
        for (int c = 0; c < 3; c++) {
            stage[c] = DField(u.elems_, u.esize_);
            check(th_state_download_buffer(u.state_, 1, c, stage[c].data(), 0), "th_state_download_buffer");
            src[c] = stage[c].data();
        }
    }
    void* F[3] = {f[0].data(), f[1].data(), f[2].data()};
    check(th_compute_force(u.plan_, src, F, &u.force_, nullptr), "Compute_force");
}

// Compute_force(IncVF& W) compute_force_main.cc:88-105 on the fields of the last RK stage (the stage loop
// recomputes both forces, so that is what Force1..3 and W.Force1..3 hold when the outputs are written); a
// part whose level per mode is not above MYEPS is never computed and stays 0, as before the first step.
void IncFluid::mhd_forces(DField fu[3], DField fw[3])
{
    for (int c = 0; c < 3; c++) {
        fu[c] = DField(elems_, esize_);
        fw[c] = DField(elems_, esize_);
    }
    if (!mhd_stepped_) return;
    void* const U[3] = {mlast_[0].data(), mlast_[1].data(), mlast_[2].data()};
    void* const B[3] = {mlast_[3].data(), mlast_[4].data(), mlast_[5].data()};
    void* FU[3] = {fu[0].data(), fu[1].data(), fu[2].data()};
    void* FW[3] = {fw[0].data(), fw[1].data(), fw[2].data()};
    check(th_compute_force_mhd(plan_, U, B, FU, FW, &mhd_force_, nullptr), "Compute_force(W)");
    check(th_sync(plan_), "th_sync");
}

void IncFluid::Set_force(IncVF& W, int proc, double inner_radius, double outer_radius, double p3, double p4,
                         double p5, double p6)
{
    th_force f{proc, inner_radius, outer_radius, p3, p4, p5, p6};
    // validate on a zero-length step (the entry point's argument checks) so a bad set fails here
    th_phys ph{dissipation_coefficient, W.dissipation_coefficient, 0, 0, 0};
    DField U1(elems_, esize_), U2(elems_, esize_), U3(elems_, esize_), B1(elems_, esize_), B2(elems_, esize_),
        B3(elems_, esize_);
    void* const U[3] = {U1.data(), U2.data(), U3.data()};
    void* const B[3] = {B1.data(), B2.data(), B3.data()};
    check(th_step_plain_forced_mhd(plan_, 2, U, B, 0.0, &ph, &f, nullptr), "Set_force(W)");
    check(th_sync(plan_), "th_sync");
    mhd_force_ = f;
    mhd_forced_ = true;
    mhd_stepped_ = false;
    if (mstate_) {
        th_state_destroy(mstate_);
        mstate_ = nullptr;
    }
}

void IncFluid::Set_force(int proc, double inner_radius, double outer_radius, double p3, double p4, double p5)
{
    th_force f{proc, inner_radius, outer_radius, p3, p4, p5, 0.0};
    stepped_ = false;
    if (fused_ && state_) {
        const int r = th_state_set_force(state_, &f);
        if (r == TH_OK) {
            force_ = f;
            forced_ = false;           // the fused step carries the force itself
            has_force_ = true;
            return;
        }
        if (r != TH_E_UNSUPPORTED) check(r, "Set_force");
        // the band reaches modes the fused state does not store: run on the plain stepper
    }
    // validate now (th_compute_force's checks) so a bad parameter set fails here, not mid-run
    DField F1(elems_, esize_), F2(elems_, esize_), F3(elems_, esize_);
    void* const V[3] = {V1.data(), V2.data(), V3.data()};
    void* F[3] = {F1.data(), F2.data(), F3.data()};
    check(th_compute_force(plan_, V, F, &f, nullptr), "Set_force");
    check(th_sync(plan_), "th_sync");
    force_ = f;
    forced_ = true;
    has_force_ = true;
    if (state_) {
        th_state_destroy(state_);
        state_ = nullptr;
    }
    fused_ = false;
}

void IncFluid::Time_advance()
{
    th_phys ph{dissipation_coefficient, 0, 0, 0, 0};
    if (forced_) {
        int sch = integ_scheme == "EULER" ? 0 : integ_scheme == "RK2" ? 1 : 2;
        void* const V[3] = {V1.data(), V2.data(), V3.data()};
        if (!last_[0].data())
            for (int c = 0; c < 3; c++) last_[c] = DField(elems_, esize_);
        void* const last[3] = {last_[0].data(), last_[1].data(), last_[2].data()};
        check(th_step_plain_forced(plan_, sch, V, Tdt, &ph, &force_, last), "th_step_plain_forced");
        stepped_ = true;
        check(th_sync(plan_), "th_sync");
        return;
    }
    if (fused_) {
        void* V[3] = {V1.data(), V2.data(), V3.data()};
        for (int i = 0; i < 3; i++) check(th_state_upload(state_, i, V[i], 0), "th_state_upload");
        check(th_step(plan_, state_, Tdt, &ph, 1), "th_step");
        for (int i = 0; i < 3; i++) check(th_state_download(state_, i, V[i], 0), "th_state_download");
        check(th_sync(plan_), "th_sync");
        stepped_ = has_force_;
        return;
    }
    int sch = integ_scheme == "EULER" ? 0 : integ_scheme == "RK2" ? 1 : 2;
    void* const V[3] = {V1.data(), V2.data(), V3.data()};
    check(th_step_plain(plan_, TH_NS, sch, V, Tdt, &ph), "th_step_plain");
    check(th_sync(plan_), "th_sync");
}

}  // namespace tarang
