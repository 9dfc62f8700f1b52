/*
 * tarang_hip.h -- C-ABI of the MI355X-native pseudo-spectral time step (libtarang_hip.so).
 *
 * Drop-in boundary for the reference's basis-transform, nonlinear-term and energy-transfer
 * interfaces (hirdeshpharasi/tarang, paths under trunk/lib/).  Every entry point names the
 * reference interface it replaces.  All functions are extern "C", take plain pointers and
 * sizes, return an int status (0 = ok, < 0 = error; th_last_error() gives the message) and
 * never throw.
 *
 * Layout: a spectral component is the reference's C-order array A[l1][l2][l3] of shape
 * (N1, N2, N3/2+1) (trunk/lib/fields/Cvf.cc:112-114) in device memory, complex double for
 * TH_FP64 plans and complex float for TH_FP32_MIXED plans.  Real space is the same buffer with
 * z-packed reals: element (x, y, j) = (u(x,y,2j), u(x,y,2j+1)); plane j = N3/2 is padding.
 * Calls are asynchronous on the plan's HIP stream unless they return host values (reductions),
 * which synchronise.  A plan must be used by one host thread at a time.
 */
#ifndef TARANG_HIP_H
#define TARANG_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TH_OK 0
#define TH_E_ARG (-1)
#define TH_E_HIP (-2)
#define TH_E_UNSUPPORTED (-3)
#define TH_E_STATE (-4)
#define TH_E_NOMEM (-5)
#define TH_E_IO (-6)        /* a file could not be opened, written or parsed */

#define TH_FOUR 0          /* basis_type "FOUR" */
#define TH_SCFT 1          /* basis_type "SCFT" (sin/cos along x, Fourier along y, z) */
#define TH_FP64 0
#define TH_FP32_MIXED 1    /* FP32 storage/transforms; FP64 twiddles, integrating factor, sums */

#define TH_NS 0            /* incompressible Navier-Stokes (IncFluid + IncVF) */
#define TH_MHD 1           /* incompressible MHD, velocity U and magnetic field B */
#define TH_RB 2            /* Rayleigh-Benard, free-slip, PRLARGE / ULARGE scaling */

typedef struct th_ctx th_ctx;       /* device + stream + autotune cache */
typedef struct th_plan th_plan;     /* one box: basis, N, kfactor, precision, radix plans, scratch */
typedef struct th_state th_state;   /* device-resident fields of one box in the fused layout */

typedef struct {
    double nu;      /* dissipation_coefficient (IncVF), velocity */
    double eta;     /* dissipation_coefficient of W (MHD magnetic field) */
    double kappa;   /* dissipation_coefficient of the scalar (RB); ignored when Ra > 0 and Pr > 0 */
    double Ra, Pr;  /* RB: Ra, Pr; nu = sqrt(Pr/Ra), kappa = 1/sqrt(Pr Ra) (RB_slip_main.cc:238-239) */
} th_phys;

/* Energy / helicity forcing of the velocity (force_double_para of IncFluid.cc:201-217):
 * proc 1 == Compute_force_const_energy_helicity_supply compute_force_ek_hk_supply.cc:51-157
 *           (p3 = energy_supply, p4 = epsh_by_k_epse),
 * proc 2 == Compute_force_const_energy_helicity compute_force_const_ek_hk.cc:51-140
 *           (p3 = energy_level, p4 = h_by_k_E),
 * proc 3 == Compute_force_Taylor_Green compute_force_TG.cc:50-104: inner_radius = k0 (truncated to int, as
 *           force_double_para(1)), outer_radius = force_amp, p3 = forceW_amp (MHD); the fixed field
 *           amp (sin k0x cos k0y cos k0z, -cos k0x sin k0y cos k0z, 0) of Setup_Taylor_Green_force_field
 *           force_model.cc:57-88, 1 <= k0 < N/2 inside the dealiased modes; runs on the plain steppers,
 * proc 4 == Compute_force_ABC compute_force_ABC.cc:52-68: inner_radius = k0, outer_radius = A, p3 = B, p4 = C,
 *           p5 = force_amp; the fixed field amp (B cos k0y + C sin k0z, A sin k0x + C cos k0z,
 *           A cos k0x + B sin k0y) of Setup_ABC_force_field force_model.cc:122-160, same k0 limits, NS plain
 *           steppers only. */
typedef struct th_force {
    int proc;                       /* force_field_proc: 1 or 2 */
    double inner_radius, outer_radius;
    double p3, p4;
    double p5, p6;                  /* MHD (IncVF& W versions): proc 1 energy_supply_W, epsh_by_k_epse_W;
                                       proc 2 energy_level_W, h_by_k_E_W */
} th_force;

/* ---- context and plans ----
 * replaces: Init_fftw_plan_array universal_tr.cc:48, Init_fftw_plan_FOUR four_tr.cc:44-98,
 *           Init_fftw_plan_SCFT scft_tr.cc:49-74 and the process globals extern_def_vars.h:95-142 */
int th_ctx_create(th_ctx **ctx, int device, const char *tune_cache_path);
int th_ctx_destroy(th_ctx *ctx);
const char *th_last_error(void);
const char *th_version(void);
int th_plan_create(th_ctx *ctx, th_plan **plan, int basis, int N1, int N2, int N3, const double kfactor[3],
                   int precision, int dealias);
/* A plan with live fused states is released when its last state is destroyed. */
int th_plan_destroy(th_plan *plan);
size_t th_plan_field_elems(const th_plan *plan);      /* N1*N2*(N3/2+1) */
size_t th_plan_elem_bytes(const th_plan *plan);       /* 16 (FP64) or 8 (FP32_MIXED) */
int th_shell_ek_size(const th_plan *plan);            /* CV_shell_ek_size = Min_radius_outside + 1 */
int th_max_radius_inside(const th_plan *plan);        /* Max_radius_inside four_inline.h:233 */
int th_sync(th_plan *plan);

/* ---- basis transforms (in place, reference layout) ----
 * th_forward == Forward_transform_array universal_tr.cc:66-86 (ArrayFFT_FOUR four_tr.cc:199-209,
 *               ArraySFT/CFT_SCFT scft_tr.cc:230-431): normalised, padding zeroed first.
 * th_inverse == Inverse_transform_array universal_tr.cc:126-146 (ArrayIFFT_FOUR four_tr.cc:229-295,
 *               ArrayISFT/ICFT_SCFT scft_tr.cc:440-639): unnormalised, padding plane zeroed. */
int th_forward(th_plan *plan, void *A, int parity);
int th_inverse(th_plan *plan, void *A, int parity);
/* th_deriv == Xderiv/Yderiv/Zderiv universal_tr.cc:186-248 (axis 1, 2, 3); B may equal A */
int th_deriv(th_plan *plan, int axis, const void *A, void *B, int parity);
/* == Dealias_array universal_misc.cc:118-159 */
int th_dealias(th_plan *plan, void *A);
/* == Array_real_mult array_basic.cc:62-71 (C = (Re A Re B, Im A Im B)) */
int th_real_mult(th_plan *plan, const void *A, const void *B, void *C);
/* == Array_divide_ksqr_FOUR four_basic.cc:220-245 / _SCFT scft_basic.cc:247 */
int th_divide_ksqr(th_plan *plan, void *A);
/* == Array_exp_ksqr_FOUR four_basic.cc:312-337 (A *= exp(factor K^2), factor in FP64) */
int th_exp_ksqr(th_plan *plan, void *A, double factor);

/* ---- nonlinear terms, pressure, rhs: the reference's operation sequence on the GPU ----
 * th_nlin_ns  == IncVF::Compute_nlin() compute_nlin.cc:58-108 (9 transforms, dealiased)
 * th_pressure == IncVF::Compute_pressure() compute_pressure.cc:54-62
 * th_rhs      == IncFluid::Compute_rhs() compute_rhs.cc:52-60 (nlin <- -(nlin + grad p))
 * th_nlin_mhd == IncVF::Compute_nlin(IncVF&) compute_nlin_VF_MHD.cc:72-164 (U, B are modified
 *                in place by the Elsasser round trip exactly as in the reference)
 * th_nlin_rb  == IncVF::Compute_nlin_scalar(IncSF&) compute_nlin_SF.cc:80-177
 * th_et_nlin  == IncVF::EnergyTr_Compute_nlin() ET_compute_nlin.cc:85-120 with helper H, giver G */
int th_nlin_ns(th_plan *plan, void *const V[3], void *nlin[3]);
int th_pressure(th_plan *plan, void *const nlin[3], void *p);
int th_rhs(th_plan *plan, void *nlin[3], const void *p);
int th_nlin_mhd(th_plan *plan, void *U[3], void *B[3], void *nU[3], void *nB[3]);
int th_nlin_rb(th_plan *plan, void *const V[3], const void *T, void *nV[3], void *nT);
int th_et_nlin(th_plan *plan, void *const H[3], void *const G[3], void *nlin[3]);

/* ---- plain steppers on reference-layout device fields (one time step, force = 0 or RB buoyancy) ----
 * == IncFluid::Time_advance() time_advance.cc:60-163 ("EULER", "RK2", "RK4"),
 *    Time_advance(IncVF&) time_advance_VF_MHD.cc:59-180, Time_advance(IncSF&) time_advance_SF.cc:61-233.
 * fields: NS V[0..2]; MHD U[0..2], B[0..2]; RB V[0..2], T.  scheme: 0 EULER, 1 RK2, 2 RK4. */
int th_step_plain(th_plan *plan, int physics, int scheme, void *const fields[], double dt, const th_phys *p);
/* NS step with the forcing of `force` recomputed from the stage field at every stage and subtracted
 * from nlin before the pressure: Compute_force_TO_rhs single_time_step.cc:351-364 (Compute_force
 * compute_force_main.cc:44-60, Add_force add_force.cc:47-53).  FOUR plans only.  last_stage (may be
 * NULL) receives the field of the step's last stage, whose force the reference's Force1..3 still hold
 * when the time loop writes its outputs (Ifluid_main.cc:278-298). */
int th_step_plain_forced(th_plan *plan, int scheme, void *const V[3], double dt, const th_phys *p,
                         const th_force *force, void *const last_stage[3]);
/* MHD step with the forcing of both fields (force_field_proc 1,
 * Compute_force_const_energy_helicity_supply(IncVF& W) compute_force_ek_hk_supply.cc:294-420, or 2,
 * Compute_force_const_energy_helicity(IncVF& W) compute_force_const_ek_hk.cc:296-437: U forced with p3, p4
 * and W with p5, p6 on the same band, each part only when its level per mode exceeds MYEPS), computed from
 * U and B before the nonlinear term and subtracted from both nlin (Add_force(IncVF& W)
 * add_force.cc:67-76) at every stage.  FOUR plans.  last_stage (may be
 * NULL) receives U1..U3, B1..B3 of the step's last stage, the fields whose forces the reference's Force1..3
 * and W.Force1..3 hold when the MHD outputs are written (output_energy.cc:463-500, output_ET.cc:118-165). */
int th_step_plain_forced_mhd(th_plan *plan, int scheme, void *const U[3], void *const B[3], double dt,
                             const th_phys *p, const th_force *force, void *const last_stage[6]);
/* The force field itself (IncFluid::Compute_force for proc 1 / 2): F = alpha V + beta (i K x V) on the
 * band inner < |K| <= outer, zero elsewhere, dealiased when the band reaches past N/3 and with the
 * reality condition applied.  nf (may be NULL) receives the mode count of Get_number_modes_in_shell.
 * Errors: TH_E_UNSUPPORTED for other procs or an SCFT plan; TH_E_ARG for an empty band, or for proc 2
 * with |h_by_k_E| > 1, where the reference's square roots are not real. */
int th_compute_force(th_plan *plan, void *const V[3], void *F[3], const th_force *force, long *nf);
/* Both forces of an MHD run (IncFluid::Compute_force(IncVF& W) compute_force_main.cc:88-105, proc 1 / 2): FU
 * from U with p3, p4 and FW from B with p5, p6, each zero when its level per mode is not above MYEPS.
 * Under proc 2 a mode of B is forced when its modal energy is > 0, a mode of U when it is > MYEPS
 * (compute_force_const_ek_hk.cc:376-402).  Errors as th_compute_force, and TH_E_ARG for non-finite p5, p6
 * or, under proc 2, |h_by_k_E_W| > 1. */
int th_compute_force_mhd(th_plan *plan, void *const U[3], void *const B[3], void *FU[3], void *FW[3],
                         const th_force *force, long *nf);

/* ---- energies, spectra, transfers (return host doubles; synchronise) ----
 * th_energy   == Get_total_energy four_energy.cc:88-168 (B == NULL: |A|^2; else Re A conj B)
 * th_Sn       == Get_total_Sn four_energy.cc:179-228
 * th_shell_spectrum == Compute_shell_spectrum universal_energy.cc:112-153 (with the partial-shell
 *                rescale four_energy.cc:324-336); count[] = Sk_count (may be NULL)
 * th_fill_array_shell == Fill_array_shell universal_ET.cc:45-65
 * th_shell_mult_single == sum of 3 Shell_mult_single (prod_nlinV.cc:62-136)
 * th_shell_mult_all    == Shell_mult_all (dot product) universal_ET.cc:160-207
 * th_flux_ns  == IncVF::Compute_flux() compute_flux.cc:52-69 (sphere radii[1..nsph])
 * th_shell_tr_ns == IncVF::Compute_shell_tr() compute_shell_tr.cc:59-79, T is (nsh+1)^2 row-major
 * th_nusselt  == IncVF::Get_Nusselt_no nusselt.cc:53-89 (PRLARGE / ULARGE)
 * th_helicity == Compute_total_helicity_FOUR four_energy.cc:453-548 / _SCFT scft_energy.cc:448-543 (V1 taken
 *                times -i): out = {H1, H2, DH1, DH2}
 * th_shell_spectrum_helicity == Compute_shell_spectrum_helicity_FOUR four_energy.cc:551-671 / _SCFT
 *                scft_energy.cc:546-677: H1k1..3 of shell_ek_size each (K_c (Vr x Vi)_c parts,
 *                partial-shell rescale)
 * th_entropy  == Get_entropy_FOUR four_energy.cc:687-750 / Get_entropy_SCFT scft_energy.cc:679-748,
 *                spectral entropy in bits */
int th_energy(th_plan *plan, const void *A, const void *B, double *out);
int th_Sn(th_plan *plan, const void *A, double n, double *out);
int th_shell_spectrum(th_plan *plan, const void *A, const void *B, int n, double *Sk, double *count);
/* th_shell_spectrum_noalloc: no counterpart in the reference.  The same shell spectrum as th_shell_spectrum,
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): it allocates no device or host memory.  Its partial sums live in a static device array of
 * 8.4 MB that each GPU holds until the process ends; shell_ek_size above 512 is refused (E_ARG).  Calls that
 * overlap on one GPU share that array and race.  Synchronises. */
int th_shell_spectrum_noalloc(th_plan *plan, const void *A, const void *B, int n, double *Sk, double *count);
int th_helicity(th_plan *plan, void *const V[3], double *out);
int th_shell_spectrum_helicity(th_plan *plan, void *const V[3], double *H1k1, double *H1k2, double *H1k3);
/* Anisotropic spectra of V about the axis globalvar_anisotropy_switch = axis (1 x, 2 y, 3 z), FOUR 3-D:
 * th_ring_spectrum == Compute_ring_spectrum_FOUR four_energy.cc:828-992: S1k, S2k of (Max_radius_inside + 1) x (nsect + 1) doubles, row-major; nsect sectors
 *   spaced linearly over [0, pi/2] (sector_spectrum_input_scheme 0, Cvf.cc:85-106); weight |K|^n, n = 0, 2, 4.
 *   S1 is the toroidal part |(V x K)_axis|^2 / K_perp^2.  On-axis modes follow the reference: axis 1 adds
 *   |V_y|^2, |V_z|^2; axis 2 adds |V_z|^2, |V_y|^2; axis 3 adds nothing (four_energy.cc:883-897).
 * th_cylinder_ring_spectrum == Compute_cylinder_ring_spectrum_FOUR four_energy.cc:1450-1547: S1k, S2k of
 *   th_cylinder_shell_ek_size(plan, axis) x (nslab + 1), rings ceil(K_perp), nslab slabs of |K_pll| over
 *   [0, Anis_max_Kpll] (Cvf.cc:166-190); S1 = |V_axis|^2 (0 for axis 3, four_energy.cc:1493-1500).
 * Bin indices past the last column land on the next row's column 0, as the reference's unchecked
 * stores do.  rows x (bins + 1) is limited to 4096.  Synchronises. */
int th_ring_spectrum(th_plan *plan, void *const V[3], int nsect, int n, int axis, double *S1k, double *S2k);
int th_cylinder_ring_spectrum(th_plan *plan, void *const V[3], int nslab, int n, int axis, double *S1k, double *S2k);
int th_cylinder_shell_ek_size(const th_plan *plan, int axis);   /* Anis_max_Krho_radius_inside + 1 */
/* Helicity tables on the same bins: th_ring_spectrum_helicity == Compute_ring_spectrum_helicity_FOUR
 * four_energy.cc:1357-1447, th_cylinder_ring_spectrum_helicity == Compute_cylinder_ring_spectrum_helicity_FOUR
 * :1825-1912; H1k[row][col] = sum of 2 mult K.(Vr x Vi). */
int th_ring_spectrum_helicity(th_plan *plan, void *const V[3], int nsect, int axis, double *H1k);
int th_cylinder_ring_spectrum_helicity(th_plan *plan, void *const V[3], int nslab, int axis, double *H1k);
int th_entropy(th_plan *plan, void *const V[3], double *out);
/* == Get_entropy_scalar_FOUR four_energy.cc:761-826 / _SCFT scft_energy.cc:750-800 (CSF::CS_Compute_entropy
 * Csf.cc:306-309): the spectral entropy of one field in bits */
int th_entropy_scalar(th_plan *plan, const void *A, double *out);
int th_fill_array_shell(th_plan *plan, const void *A, void *B, double inner, double outer);
int th_shell_mult_single(th_plan *plan, void *const A[3], void *const B[3], double inner, double outer, double *out);
int th_shell_mult_all(th_plan *plan, void *const A[3], void *const B[3], const double *radii, int nr, double *out);
int th_flux_ns(th_plan *plan, void *const V[3], const double *radii, int nsph, double *flux);
int th_shell_tr_ns(th_plan *plan, void *const V[3], const double *radii, int nsh, double *T);
/* th_flux_mhd == IncVF::Compute_flux(IncVF& W) compute_flux.cc:148-262 for spheres radii[1..nsph];
 * out is 9 x (nsph+1), row-major, rows: 0 U<->U> (flux_self), 1 B<->B> (W.flux_self),
 * 2 U<->B> (flux_VF_in_out), 3 U<->B< (flux_VF_in_in), 4 B<->U> (W.flux_VF_in_out),
 * 5 B<->U< (W.flux_VF_in_in), 6 U>->B> (flux_VF_out_out), 7 Zp<->Zp> (flux_Elsasser),
 * 8 Zm<->Zm> (W.flux_Elsasser); column 0 is 0. */
int th_flux_mhd(th_plan *plan, void *const U[3], void *const B[3], const double *radii, int nsph, double *out);
/* th_flux_scalar == IncVF::Compute_flux_scalar(IncSF&) compute_flux.cc:89-109, the branch
 * Compute_flux_RB :113-141 takes for the PRLARGE / ULARGE scalings (passive scalar on FOUR, RB on SCFT):
 * out is 2 x (nsph+1), row 0 = flux_self (as th_flux_ns), row 1 = flux_SF = -(U.grad T<).T> per sphere.
 * th_shell_tr_scalar == IncVF::Compute_shell_tr_scalar compute_shell_tr.cc:98-117 (Compute_shell_tr_RB
 * :120-150): Tself as th_shell_tr_ns, Tsf[n][m] = -(U.grad T_n).T over shell m; both (nsh+1)^2 row-major. */
int th_flux_scalar(th_plan *plan, void *const V[3], void *T, const double *radii, int nsph, double *out);
int th_shell_tr_scalar(th_plan *plan, void *const V[3], void *T, const double *radii, int nsh, double *Tself,
                       double *Tsf);
/* th_shell_tr_mhd == IncVF::Compute_shell_tr(IncVF& W) compute_shell_tr.cc:154-260: out holds five
 * (nsh+1)^2 row-major matrices, rows n = 1..nsh-1 the giver shell: 0 U to U (shelltoshell_self),
 * 1 W to W (W.shelltoshell_self), 2 U to W (shelltoshell_VF), 3 Zp to Zp (shelltoshell_Elsasser),
 * 4 Zm to Zm (W.shelltoshell_Elsasser), Z+- = U +- B. */
int th_shell_tr_mhd(th_plan *plan, void *const U[3], void *const B[3], const double *radii, int nsh, double *out);
/* th_ring_tr == IncVF::Compute_ring_tr compute_ring_tr.cc:60-291 (cyl = 0: shells x polar sectors about
 * axis 1..3) and IncVF::Compute_cylinder_ring_tr compute_cylin_ring_tr.cc:54-133 (cyl = 1: K_perp shells x
 * K_pll slabs), FOUR 3-D.  radii[0..nr] (radii[nr] = INF_RADIUS), edges[0..nsec] (sector angles up to pi/2 or
 * K_pll edges up to Anis_max_Kpll), 2 <= nr, nsec <= 32.  For giver ring (gs, gq), gs = 1..ngiv (<= nr),
 * gq = 1..nsec: G = F on the ring (Fill_array_ring / Fill_array_cylinder_ring universal_ET.cc:70-127),
 * N = EnergyTr_Compute_nlin with helper H (ncomp 3), or its scalar form (ncomp 1: F[0] = T, Rv[0] = T), and
 * out[(gs (nsec+1) + gq) M + (rs (nsec+1) + rq)] = sign * Ring_mult_all(N, Rv) over receiver ring (rs, rq),
 * M = (nr+1)(nsec+1) (four_ET.cc:259-460).  The reference's matrices: ring_to_ring_self (H = F = Rv = U,
 * sign -1, ngiv = nr - 1), W.ring_to_ring_self (H = U, F = Rv = B, -1), ring_to_ring_VF (H = B, F = U,
 * Rv = B, +1), the Elsasser pair (H = Z-+, F = Rv = Z+-, -1) and ring_to_ring_SF (H = U, F = Rv = T, -1). */
int th_ring_tr(th_plan *plan, int cyl, int axis, void *const H[3], void *const F[3], void *const Rv[3], int ncomp,
               const double *radii, int nr, const double *edges, int nsec, int ngiv, double sign, double *out);
/* th_ring_mult_all == Ring_mult_all four_ET.cc:259-340 (cyl = 0) / Cyl_ring_mult_all :422-500 (cyl = 1):
 * out[(nr+1)(nsec+1)] = 2 mult Re(A . B*) over each receiver ring (the force feeds of compute_force_feed.cc:
 * 162-250 with A = Force, B = V). */
int th_ring_mult_all(th_plan *plan, int cyl, int axis, void *const A[3], void *const B[3], int ncomp,
                     const double *radii, int nr, const double *edges, int nsec, double *out);
/* th_ub_to_elsasser == IncVF::UB_to_Elsasser_field elsasser.cc:53-60 (back = 0: U <- U + B, B <- U - B) and
 * Elsasser_to_UB_field :62-75 (back = 1: U <- (Zp + Zm) / 2, B <- (Zp - Zm) / 2), in place; asynchronous. */
int th_ub_to_elsasser(th_plan *plan, void *const U[3], void *const B[3], int back);
int th_nusselt(th_plan *plan, void *const V[3], const void *T, double Ra, double Pr, double *Nu);

/* ---- initial conditions on reference-layout device fields ----
 * th_ic_taylor_green == IncVF::Setup_Taylor_Green_field ic_model.cc:58-107
 * th_ic_random: seeded counter-based random-phase field, projected, Hermitian on kz = 0
 *   (Satisfy_reality_array universal_misc.cc:51-109), dealiased, unit shells rescaled to
 *   E(k) = C k^4 exp(-2 (k/kp)^2) with total energy E0 (replaces the time(0)-seeded
 *   Init_cond_energy_spectrum init_cond_energy.cc:65-125)
 * th_ic_random_scalar: seeded temperature perturbation of amplitude amp, T(0,:,:) = 0
 *   (Zero_modes_RB_slip zero_mode_RB_slip.cc:50-58) */
int th_ic_taylor_green(th_plan *plan, void *V[3], int k0, double amp);
int th_ic_random(th_plan *plan, void *V[3], unsigned long long seed, double E0, double kp);
int th_ic_random_scalar(th_plan *plan, void *T, unsigned long long seed, double amp);
/* th_satisfy_reality == Satisfy_reality_array universal_misc.cc:51-109 on the kz = 0 and kz = N3/2
 * planes of A (what Init_cond applies to a field read from file when apply_realitycond_IC_switch = 1) */
int th_satisfy_reality(th_plan *plan, void *A);

/* ---- fused, dealias-pruned step on a device-resident state ----
 * The state holds the kept (2/3-rule) modes of every component in a compact layout; upload and
 * download convert from / to the reference layout (src/dst on host or device).
 * th_step == nsteps x IncFluid::Time_advance() with the RK4 integrating-factor scheme
 *   (time_advance.cc:100-155, _VF_MHD.cc:105-180, _SF.cc:105-222), fused passes, hipGraph replay.
 * RB (TH_RB, SCFT basis) == Time_advance(IncSF&) time_advance_SF.cc:105-222 with the DCT/DST x
 *   passes; dissipation sqrt(Pr/Ra), 1/sqrt(Pr Ra) when Ra, Pr > 0, else nu and kappa.
 * comp indexes: NS V1..V3 = 0..2; MHD U 0..2, B 3..5; RB V 0..2, T 3. */
int th_state_create(th_plan *plan, th_state **state, int physics);
int th_state_destroy(th_state *state);
int th_state_upload(th_state *state, int comp, const void *src, int src_on_host);
int th_state_download(th_state *state, int comp, void *dst, int dst_on_host);
/* No counterpart in the reference: which = 0 state, 1 RK stage field, 2 RK4 accumulator, each in the
 * reference layout.  After th_step the stage field is the input of the step's last RK stage, the field
 * whose force (Compute_force / Compute_force_RB) the reference's Force arrays hold after Time_advance.
 * A chained step (NS, MHD) writes no stage field; it keeps the last stage's input only for a forced state,
 * and which = 1 after a chained step of an unforced state returns E_UNSUPPORTED. */
int th_state_download_buffer(th_state *state, int which, int comp, void *dst, int dst_on_host);
int th_step(th_plan *plan, th_state *state, double dt, const th_phys *p, int nsteps);
/* Forcing of later th_step calls (force_field_proc 1 / 2 recomputed from every stage field, as
 * Compute_force_TO_rhs single_time_step.cc:351-364); NULL or proc 0 switches it off.  Single-box NS
 * states; TH_E_UNSUPPORTED when the band reaches a mode the 2/3-dealiased state does not store (then
 * use th_step_plain_forced), otherwise the errors of th_compute_force. */
int th_state_set_force(th_state *state, const th_force *force);
/* Pi(R_i) for spheres radii[1..nsh] and T(n,m) for giver shells n = 1..nsh-1 over the state's
 * velocity (Compute_flux + Compute_shell_tr), from one giver-masked fused pipeline per shell. */
int th_flux_shell_tr(th_plan *plan, th_state *state, const double *radii, int nsh, double *Pi, double *T);
/* th_flux_mhd_fused == IncVF::Compute_flux(IncVF& W) compute_flux.cc:148-262 on a fused MHD state (U, B): the
 * nine rows of th_flux_mhd (9 x (nsph+1), same order) for spheres radii[1..nsph] (radii[nsph] = INF_RADIUS),
 * built from giver shells through the fused pipeline: four helper x giver pipelines (H, G in {U, B}), each
 * binned against both receivers, the Elsasser rows from the bilinearity of N[H, G].  Synchronises. */
int th_flux_mhd_fused(th_plan *plan, th_state *state, const double *radii, int nsph, double *out);

/* Measurement only (no counterpart in the reference): runs RK stage `stage` (1..4) of the fused
 * pipeline nrep times pass by pass, timing each of the 5 passes with HIP events on the plan's
 * stream.  ms[k] = mean duration of pass k (x-inverse, y-inverse, fused z, y-forward,
 * x-forward + epilogue); bytes[k] = its algorithmic HBM bytes.  Overwrites the stage field and the
 * accumulator; the state is restored after a stage-4 run. */
int th_profile_passes(th_plan *plan, th_state *state, double dt, const th_phys *p, int stage, int nrep,
                      double ms[5], double bytes[5]);

/* Measurement only (no counterpart in the reference, whose FFTW codelets are planned in
 * Init_fftw_plan_FOUR four_tr.cc:44-98): the 16-point DFT of the line FFTs' radix-16 stages on nvec
 * FP32 complex vectors, stored point-major (x[j * nvec + v] = point j of vector v), by
 * method 0 = the vector-ALU codelet or method 1 = a DFT-as-GEMM on the FP32 matrix cores
 * (v_mfma_f32_32x32x2f32, the DFT as a real 32 x 32 matrix times 32 vectors).  y = (W / 4)^reps x with
 * W[k][n] = exp(sign 2 pi i k n / 16).  nvec must be a positive multiple of 256, 1 <= reps <= 32 (4^32 |x| stays inside the FP32 range).  The call
 * launches once (asynchronously on the context's stream); with ntimed > 0 it then times ntimed more
 * launches with HIP events and returns their mean duration in *ms (synchronises). */
int th_dft16_probe(th_ctx *ctx, int method, int sign, const void *x, void *y, long nvec, int reps, int ntimed,
                   double *ms);

/* ---- launch-shape autotuner (no counterpart in the reference, whose transforms are planned by
 * FFTW_MEASURE in Init_fftw_plan_FOUR four_tr.cc:44-98) ----
 * th_tune times every launch-shape variant of the five fused passes on the state's own buffers
 * (one RK stage; three interleaved rounds of nrep launches per variant, the best round counting),
 * keeps the fastest in the state and records it in the
 * context's tune cache (th_ctx_create's tune_cache_path), from which later states of the same
 * workload start.  var (may be NULL) receives the chosen variant per pass: x-inverse, y-inverse,
 * fused z, y-forward, x-forward.  The state's fields are unchanged.
 * th_state_set_variants sets them directly (tests, reproducing a tuned run); th_state_variants
 * reads them (a new state starts from the cache entry of its workload, else all 0). */
int th_tune(th_plan *plan, th_state *state, double dt, const th_phys *p, int nrep, int var[5]);
int th_state_set_variants(th_state *state, const int var[5]);
int th_state_variants(const th_state *state, int var[5]);

/* ---- ensemble of independent NS boxes of uneven sizes ----
 * The reference runs one grid size per process (its transform and energy routines keep static
 * scratch arrays sized for one grid, four_tr.cc:173-174, four_energy.cc:302-303), so an ensemble
 * there is a sequence of runs, each an Ifluid_main.cc time loop.  Here the boxes (FOUR basis,
 * kfactor 1, dealiased, box b of size N[3b..3b+2]) advance together: th_ensemble_step ==
 * nsteps x IncFluid::Time_advance() (RK4, time_advance.cc:100-155) for every box, box b with its
 * own dt[b] and viscosity nu[b].  Boxes of one size share each pass launch; the size classes run
 * as concurrent branches of one captured step.  th_ensemble_plan returns the plan of a box's size
 * (for initial conditions and spectra on reference-layout buffers). */
typedef struct th_ensemble th_ensemble;
int th_ensemble_create(th_ctx *ctx, th_ensemble **ens, int nbox, const int *N, int precision);
int th_ensemble_destroy(th_ensemble *ens);
int th_ensemble_plan(th_ensemble *ens, int box, th_plan **plan);
int th_ensemble_upload(th_ensemble *ens, int box, int comp, const void *src, int src_on_host);
int th_ensemble_download(th_ensemble *ens, int box, int comp, void *dst, int dst_on_host);
int th_ensemble_step(th_ensemble *ens, const double *dt, const double *nu, int nsteps);
/* th_tune for every size class of the ensemble (per-box dt and nu as in th_ensemble_step);
 * th_ensemble_variants reads the launch variants of box `box`'s size class (th_state_variants' order) */
int th_ensemble_tune(th_ensemble *ens, const double *dt, const double *nu, int nrep);
int th_ensemble_variants(const th_ensemble *ens, int box, int var[5]);

/* ---- field files (host arrays only; no GPU is touched) ----
 * th_field_write == IncFluid::Output_field output_field.cc:43-58 (CV_output Cvf.cc:715-720,
 *   Write_data_MPI array_basic.cc:814-858): writes (append = 0) or appends one snapshot, a
 *   "%% Time = t" line and ncomp components of N1 x N2 x (N3/2+1) values "(re,im) " in fixed notation with
 *   12 decimals (field_out_file: ios::fixed, MY_PRECISION = 12, file_operations.cc:229-230,
 *   extern_def_vars.h:78), each component followed by a blank line.
 * th_field_read == the reading half of IncFluid::Init_cond init_cond_field.cc:47-86 with
 *   field_input_proc = 1 (CV_input Cvf.cc:739-744, Read_data_MPI array_basic.cc:294-335): snapshot
 *   number `snapshot` (0 = first) of such a file into fields; *time = its "%% Time" value, NaN when the
 *   file has no such line.  The caller applies Satisfy_reality / Dealias as Init_cond does.
 * fields[c]: host arrays of complex double (TH_FP64) or complex float (TH_FP32_MIXED) in the
 * reference layout; N = {N1, N2, N3}. */
int th_field_write(const char *path, const int N[3], int precision, int ncomp, const void *const fields[],
                   double time, int append);
int th_field_read(const char *path, const int N[3], int precision, int ncomp, void *const fields[], int snapshot,
                  double *time);
/* Reduced-size files (field_out_reduced.d; field_input_proc = 2), Nr = {Nr1, Nr2, Nr3} even, 2 <= Nr[d] <= N[d]:
 * th_field_write_reduced == IncFluid::Output_field_reduced output_field.cc:505-520 (CV_output(file, Nreduced)
 *   Cvf.cc:732-737, Write_data_MPI(..., Nreduced, ...) array_basic.cc:997-1170): the Nr1 x Nr2 x (Nr3/2+1) modes
 *   of the reduced grid in its own layout (x rows kx = 0..Nr1/2 then -Nr1/2+1..-1 for FOUR, kx = 0..Nr1-1 for
 *   SCFT; ky = 0..Nr2/2 then -Nr2/2+1..-1; kz = 0..Nr3/2) taken from the N-grid arrays, each value "(re,im)" with
 *   no separator at the stream's default 6 significant digits, a blank line after each component.
 * th_field_read_reduced == the reading half of IncFluid::Init_cond_reduced init_cond_field.cc:292-330
 *   (CV_input(file, Nreduced) Cvf.cc:747-752, Read_data_MPI(..., Nreduced, ...) array_basic.cc:388-590): those
 *   modes into zeroed N-grid arrays, the reduced grid's Nyquist planes on both sides (ky = Nr2/2 also at
 *   -Nr2/2; FOUR: kx = Nr1/2 also at -Nr1/2), as the reference places them.  basis: TH_FOUR or TH_SCFT. */
int th_field_write_reduced(const char *path, const int N[3], const int Nr[3], int basis, int precision, int ncomp,
                           const void *const fields[], double time, int append);
int th_field_read_reduced(const char *path, const int N[3], const int Nr[3], int basis, int precision, int ncomp,
                          void *const fields[], int snapshot, double *time);

#ifdef __cplusplus
}
#endif
#endif
