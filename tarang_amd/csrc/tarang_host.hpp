// tarang_host.hpp -- C++ host classes mirroring the reference's hot-path interfaces, implemented
// over the C-ABI of libtarang_hip.so.  Method names and argument meanings follow the reference:
//   CVF      trunk/lib/fields/Cvf.h:82      (V1..V3 spectral, CV_Forward/Inverse_transform,
//                                            CV_Compute_totalenergy_diss, CV_Compute_shell_spectrum, CV_Dealias)
//   RVF      trunk/lib/fields/Rvf.h:92      (V1r..V3r, RV_Inverse_transform, RV_Forward_transform_RSprod)
//   NLIN     trunk/lib/IncFlow/Nlin.h:51    (nlin1..3, F (pressure), NLIN_diag_Forward_transform_derivative)
//   IncVF    trunk/lib/IncFlow/IncVF.h:75   (Compute_nlin, Compute_nlin(IncVF&), Compute_pressure,
//                                            Compute_flux, Compute_shell_tr)
//   IncFluid trunk/lib/IncFluid/IncFluid.h:82 (Compute_rhs, Time_advance, Get_dt, Tnow, Tdt)
//   EnergyTr trunk/lib/IncFlow/EnergyTr.h:97 (sphere / shell radii, flux_self, shelltoshell_self)
// Fields live in device memory in the reference layout (N1, N2, N3/2+1); errors become
// std::runtime_error carrying th_last_error().  Time_advance uses the fused dealias-pruned step;
// Compute_nlin & co. use the plain path that keeps the reference's operation sequence.
#pragma once
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include "tarang_hip.h"

namespace tarang {

void check(int rc, const char* what);

struct Grid {
    int N[4];                 // 1-based like the reference's N[]
    double kfactor[4];        // 1-based
    std::string basis_type;   // "FOUR" | "SCFT"
    std::string alias_switch; // "DEALIAS" | "ALIAS"
    std::string precision;    // "FP64" | "FP32_MIXED"
};

class Context {
public:
    explicit Context(int device = 0);
    ~Context();
    th_ctx* h = nullptr;
};

// A device field: one spectral component in the reference layout.
class DField {
public:
    DField() = default;
    DField(size_t elems, size_t esize);
    ~DField();
    DField(const DField&) = delete;
    DField& operator=(const DField&) = delete;
    DField(DField&& o) noexcept;
    DField& operator=(DField&& o) noexcept;
    void* data() const { return p_; }
    void upload(const void* host);
    void download(void* host) const;
    void copy_from(const DField& o);
    void zero();
    size_t bytes() const { return n_ * es_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0, es_ = 0;
};

class CVF {
public:
    CVF(Context& ctx, const Grid& g);
    virtual ~CVF();
    void CV_Forward_transform();                     // Cvf.cc:294 (in place, parity V1 SFT, V2/V3 CFT)
    void CV_Inverse_transform();
    void CV_Compute_totalenergy_diss();              // Cvf.cc:385-397
    void CV_Compute_shell_spectrum(std::vector<double>& result); // Cvf.cc:419-449 (n = 0, summed comps)
    void CV_Dealias();                               // Cvf.cc:699-704
    // field files (th_field_write / th_field_read): Cvf.cc:715-720 and 739-744
    void CV_output(const std::string& path, double time, bool append);
    double CV_input(const std::string& path, int snapshot = 0);   // returns the snapshot's time
    // reduced-size files (th_field_write_reduced / th_field_read_reduced): Cvf.cc:732-737 and 747-752
    void CV_output(const std::string& path, const int Nreduced[3], double time, bool append);
    double CV_input(const std::string& path, const int Nreduced[3], int snapshot = 0);
    int shell_ek_size() const;
    th_plan* plan() const { return plan_; }
    DField V1, V2, V3;
    double CV_total_energy = 0, CV_total_dissipation = 0;
    Grid grid;

protected:
    Context& ctx_;
    th_plan* plan_ = nullptr;
    size_t elems_ = 0, esize_ = 0;
    void* v(int i) const { return i == 0 ? V1.data() : i == 1 ? V2.data() : V3.data(); }
};

class RVF {
public:
    RVF(CVF& owner);
    void RV_Inverse_transform();                     // Rvf.cc:135 (V1r SFT, V2r/V3r CFT)
    void RV_Forward_transform_RSprod();              // Rvf.cc:179 (V1V2, V3V1 SFT; V2V3 CFT)
    DField V1r, V2r, V3r;

private:
    CVF& o_;
};

class NLIN {
public:
    NLIN(CVF& owner);
    void NLIN_diag_Forward_transform_derivative();   // Nlin.cc:85-94
    DField nlin1, nlin2, nlin3, F;

private:
    CVF& o_;
};

class EnergyTr {
public:
    EnergyTr(const CVF& f, int no_shells, int shell_input_scheme);   // EnergyTr.cc:44-170
    // shell_input_scheme != 0: the radii are ET_shell_radii_sector_array's entries r(0..no_shells-1), the last
    // radius INF_RADIUS (EnergyTr.cc:162-170); the scheme-1 default is r(i) = i.
    void Set_shell_radii(const std::vector<double>& r);
    std::vector<double> shell_radius, sphere_radius;
    std::vector<double> flux_self;                   // (no_spheres+1)
    std::vector<double> shelltoshell_self;           // (no_shells+1)^2 row-major
    int no_shells, no_spheres;
    // Anisotropic transfers (EnergyTr.cc:176-360) about axis globalvar_anisotropy_switch: ring shells x polar
    // sectors (ET_anisotropic_ring_switch) and cylinder shells x K_pll slabs (ET_anisotropic_cylinder_switch).
    // Shell scheme 0: radii 0, 4, 8, log-spaced, kmax/2, kmax, INF (kmax = Max_radius_inside, or
    // Anis_max_Krho_radius_inside for cylinders); otherwise r(i) = i (ET_shell_radii_sector_array's role).
    // Sectors split [0, pi/2] and slabs [0, Anis_max_Kpll] linearly (input schemes 0).
    void Init_ring_tr(int no_ring_shells, int no_sectors, int ring_scheme, int no_cylinder_shells, int no_slabs,
                      int cylinder_scheme, int axis);
    std::vector<double> ring_shell_radius, sector_angle_ring_tr, cylinder_shell_radius, cylinder_kpll_array_tr;
    int no_ring_shells = 0, no_sectors_ring_tr = 0, no_cylinder_shells = 0, no_cylinder_kpll_slabs = 0;
    int anisotropy_axis = 3;
    // 4-D tables [giver shell][giver sector][receiver shell][receiver sector], row-major
    std::vector<double> ring_to_ring_self, cylinder_ring_to_ring_self;

protected:
    const CVF* et_f_;
};

class IncVF : public CVF, public NLIN, public EnergyTr {
public:
    IncVF(Context& ctx, const Grid& g, double dissipation_coefficient, int no_shells = 32, int shell_scheme = 1);
    void Compute_nlin();                             // compute_nlin.cc:58-108 (plain path)
    void Compute_pressure();                         // compute_pressure.cc:54-62
    void Compute_flux();                             // compute_flux.cc:52-69
    void Compute_shell_tr();                         // compute_shell_tr.cc:59-79
    void Setup_Taylor_Green_field(int k0, double amp);   // ic_model.cc:58-107
    // compute_ring_tr.cc:60-291 and compute_cylin_ring_tr.cc:54-133 (Init_ring_tr sets the rings):
    void Compute_ring_tr();                          // ring_to_ring_self
    void Compute_ring_tr(IncVF& W);                  // + W.ring_to_ring_self, ring_to_ring_VF, Elsasser pair
    void Compute_cylinder_ring_tr();                 // cylinder_ring_to_ring_self
    void Compute_cylinder_ring_tr(IncVF& W);
    std::vector<double> ring_to_ring_VF, ring_to_ring_Elsasser, cylinder_ring_to_ring_VF,
        cylinder_ring_to_ring_Elsasser;
    double dissipation_coefficient;

protected:
    // one 4-D transfer table: helper H, giver source F, receiver R (ncomp 3 or 1), sign; giver shells 1..ngiv
    void ring_table(bool cyl, void* const H[3], void* const F[3], void* const R[3], int ncomp, double sign, int ngiv,
                    std::vector<double>& out);
};

// Scalar field T (CSF / IncSF, Csf.cc, IncSF.h): the temperature of Rayleigh-Benard runs.
class IncSF {
public:
    IncSF(Context& ctx, const Grid& g, double diffusion_coefficient);
    ~IncSF();
    IncSF(const IncSF&) = delete;
    IncSF& operator=(const IncSF&) = delete;
    void CS_Compute_totalenergy_diss();              // Csf.cc:293-300
    void CS_Compute_entropy();                       // Csf.cc:306-309
    th_plan* plan() const { return plan_; }
    DField F;
    double diffusion_coefficient;
    double CS_total_energy = 0, CS_total_dissipation = 0, CS_entropy = 0;

private:
    th_plan* plan_ = nullptr;
};

class IncFluid : public IncVF {
public:
    IncFluid(Context& ctx, const Grid& g, double nu, double Tinit, double Tdt_fixed, const std::string& scheme = "RK4",
             bool fused = true, int no_shells = 32, int shell_input_scheme = 1);
    ~IncFluid() override;
    void Compute_rhs();                              // compute_rhs.cc:52-60
    void Time_advance();                             // time_advance.cc:60-163
    double Get_dt();                                 // compute_dt.cc:47-59
    void Output_field(const std::string& path);      // output_field.cc:43-58 (appends a snapshot at Tnow)
    // output_field.cc:505-520: appends the modes of the Nreduced grid (N_out_reduced) at Tnow
    void Output_field_reduced(const std::string& path, const int Nreduced[3]);
    // force_field_proc 1 (energy/helicity supply) or 2 (constant energy/helicity) with
    // force_double_para (inner_radius, outer_radius, p3, p4) (IncFluid.cc:201-217).  Later steps recompute
    // the force at every stage (Compute_force_TO_rhs single_time_step.cc:351-364): on the fused step when
    // every forced mode is one the dealiased state stores, else on the plain stepper (the fused state is
    // then released).
    // p5: force_double_para(5) (the amplitude of the ABC force, force_field_proc 4).
    void Set_force(int force_field_proc, double inner_radius, double outer_radius, double p3, double p4,
                   double p5 = 0.0);
    // Init_cond() for field_input_proc = 1, init_cond_field.cc:47-86: read V from the file, then
    // Satisfy_reality_condition_field when asked, and Dealias; returns the snapshot's time (NaN if none)
    double Init_cond(const std::string& path, bool reality = false, int snapshot = 0);
    // Init_cond_reduced() for field_input_proc = 2, init_cond_field.cc:292-330: V from a file of the Nreduced grid
    // (N_in_reduced; the other modes 0), then as Init_cond
    double Init_cond_reduced(const std::string& path, const int Nreduced[3], bool reality = false, int snapshot = 0);
    // Output files, the reference's names and formats (stream precision 12):
    //   Open_output_files file_operations.cc:163-240 (<data_dir>/out/glob.d, spectrum.d, flux.d)
    //   Output_global output_energy.cc:43-81, Output_shell_spectrum :368-394, Output_flux output_ET.cc:45-74
    // The force-spectrum (spectrum.d) and force-feed (flux.d) columns come from the force of the current
    // velocity (0 for unforced runs); helicity_flux_switch = 0.
    // The streams keep the reference's formats: glob.d fixed with 12 decimals (MY_PRECISION), the other
    // files the default stream format (6 significant digits).  Every file opens with Output_prefix
    // (file_operations.cc:108-121); shell.d then holds "%% Shellradius = " and the radii 1..no_shells.
    void Open_output_files(const std::string& data_dir, const std::string& prefix_str = "");
    void Output_global();
    void Output_shell_spectrum();
    void Output_flux();
    // Output_shell_to_shell output_ET.cc:220-254 (U to U), :258-290 (U to U, T to T), :293-352 (U to U, W to W,
    // U to W, Zp to Zp, Zm to Zm): rows 1..no_shells-1, columns 1..no_shells of each matrix, printed as a
    // Blitz++ 2-D array (helicity_shell_ET_switch = 0; no mean field B0).
    void Output_shell_to_shell();
    void Output_shell_to_shell(IncSF& T);
    void Output_shell_to_shell(IncVF& W);
    // Output_ring_to_ring output_ET.cc:427-555 and Output_cylinder_ring_to_ring :627-700 (ring_to_ring.d,
    // cyl_ring_to_ring.d; opened by Open_output_files when Init_ring_tr set rings): the 4-D tables over giver and
    // receiver shells 1..n-1 and sectors (slabs) 1..m, as Blitz++ prints a 4-D array, then the force feed
    // (Compute_force_feed_ring compute_force_feed.cc:162-250: Ring_mult_all(Force, V), 0 unforced).
    void Output_ring_to_ring();
    void Output_ring_to_ring(IncSF& T);
    void Output_ring_to_ring(IncVF& W);
    void Output_cylinder_ring_to_ring();
    void Output_cylinder_ring_to_ring(IncSF& T);
    void Output_cylinder_ring_to_ring(IncVF& W);
    void Compute_ring_tr(IncSF& T);                  // ring_to_ring_self and ring_to_ring_SF
    void Compute_cylinder_ring_tr(IncSF& T);
    using IncVF::Compute_ring_tr;
    using IncVF::Compute_cylinder_ring_tr;
    std::vector<double> ring_to_ring_SF, cylinder_ring_to_ring_SF;
    std::ofstream ring_to_ring_file, cylinder_ring_to_ring_file;
    std::vector<double> shelltoshell_SF;             // T to T, (no_shells+1)^2 (Compute_shell_tr(T))
    std::vector<double> shelltoshell_mhd;            // 5 x (no_shells+1)^2: th_shell_tr_mhd's matrices
    double Tfinal = 0;                               // printed by Output_prefix
    int anisotropy_switch = 3, waveno_switch = 0;    // globalvar_anisotropy_switch / globalvar_waveno_switch
    using IncVF::Compute_flux;
    // Incompressible MHD with the magnetic field W (W.dissipation_coefficient = eta), the IMHD program's
    // step and outputs (src/IMHD, output_main.cc:311-351):
    void Time_advance(IncVF& W);                     // time_advance_VF_MHD.cc:59-180 (fused or plain)
    // force_field_proc 1 or 2 of an MHD run (Compute_force(IncVF& W) compute_force_main.cc:88-105): U takes
    // p3 / p4 and W takes p5 / p6 on the one band; later Time_advance(W) calls recompute both forces at every
    // stage on the plain stepper.  The force columns of the MHD outputs come from the forces of the last RK
    // stage's U and W.
    void Set_force(IncVF& W, int force_field_proc, double inner_radius, double outer_radius, double p3, double p4,
                   double p5, double p6);
    double Get_dt(IncVF& W);                         // compute_dt.cc:75-101
    void Compute_flux(IncVF& W);                     // compute_flux.cc:148-262 -> flux_mhd
    void Output_global(IncVF& W);                    // output_energy.cc:217-286
    void Output_shell_spectrum(IncVF& W);            // output_energy.cc:463-500
    void Output_flux(IncVF& W);                      // output_ET.cc:118-165
    std::vector<double> flux_mhd;                    // 9 x (no_spheres + 1), the rows of th_flux_mhd
    // Rayleigh-Benard (SCFT, PRLARGE with ULARGE scaling: dissipation_coefficient = sqrt(Pr/Ra),
    // T.diffusion_coefficient = 1/sqrt(Pr Ra), RB_slip_main.cc:238-239), the RB_SLIP program's step and glob.d:
    void Time_advance(IncSF& T);                     // time_advance_SF.cc:61-233 (fused or plain)
    double Get_dt(IncSF& T);                         // compute_dt.cc:66-69 (= Get_dt())
    double Get_Nusselt_no(IncSF& T);                 // nusselt.cc:53-89
    void Output_global(IncSF& T);                    // Output_global_RB output_energy.cc:156-215
    void Output_shell_spectrum(IncSF& T);            // Output_shell_spectrum_scalar output_energy.cc:410-449
    void Compute_flux(IncSF& T);                     // Compute_flux_RB compute_flux.cc:113-141 -> flux_self, flux_SF
    void Output_flux(IncSF& T);                      // output_ET.cc:81-115
    std::vector<double> flux_SF;                     // (no_spheres + 1)
    double Rayleigh = 0, Prandtl = 0;                // Ra, Pr of the RB run
    std::ofstream global_file, spectrum_file, flux_file, shell_to_shell_file;
    double Tnow, Tdt, Tdt_fixed;
    std::string integ_scheme;

private:
    struct RingFeed {
        const char* name;
        void* const* A;        // nullptr: zeros
        void* const* B;
        int ncomp;
    };
    void ring_feed(bool cyl, std::ostream& os, void* const* A, void* const* B, int ncomp);
    void ring_out(bool cyl, const std::vector<std::pair<const char*, const std::vector<double>*>>& tables,
                  const std::vector<RingFeed>& feeds);
    void ring_out_ns(bool cyl);
    void ring_out_scalar(bool cyl, IncSF& T);
    void ring_out_mhd(bool cyl, IncVF& W);
    struct Force_field {           // Force1..3 as the reference holds them when the time loop writes outputs
        explicit Force_field(IncFluid& u);
        DField f[3];
    };
    bool fused_;
    bool forced_ = false;          // the plain stepper carries the force
    bool has_force_ = false;       // Set_force was called (fused or plain)
    bool stepped_ = false;         // a forced step has run since Set_force
    bool rb_stepped_ = false;      // a fused RB step has run (its stage buffer holds the last stage)
    void rb_forces(DField fv[3], DField& ft);   // Force1..3 and T.Force after Time_advance(T)
    DField last_[3];               // plain forced stepper: the field of the last stage
    th_force force_{};
    th_force mhd_force_{};
    bool mhd_forced_ = false;      // Set_force(W, ...) was called
    bool mhd_stepped_ = false;     // a forced MHD step has run since Set_force(W, ...)
    DField mlast_[6];              // forced MHD stepper: U1..U3, B1..B3 of the last stage
    void mhd_forces(DField fu[3], DField fw[3]);   // Force1..3 and W.Force1..3 after Time_advance(W)
    th_state* state_ = nullptr;
    th_state* mstate_ = nullptr;   // fused MHD state (U, B), created by the first Time_advance(W)
    th_state* rstate_ = nullptr;   // fused RB state (V, T), created by the first Time_advance(T)
    bool state_fresh_ = false;
    double Tinit_;
};

}  // namespace tarang
