/*
 * tarang_oracle.h -- FP64 CPU restatement of Tarang's pseudo-spectral time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product library (libtarang_hip.so) never links, calls or falls back to it.
 *
 * Layout (identical to the reference, trunk/lib/basis_basicfn/extern_def_vars.h:57-61 and
 * trunk/lib/fields/Cvf.cc:112-114): a spectral component is a C-order array of
 * complex doubles A[l1][l2][l3], shape (N1, N2, N3/2+1).  Real space uses the same buffer:
 * complex element (x, y, j) holds (u(x,y,2j), u(x,y,2j+1)); plane j = N3/2 is padding.
 *
 * Every function cites the reference file:line it restates.  Parity status of each
 * function (which test pins it) is listed in tests/test_oracle_*.py docstrings.
 */
#ifndef TARANG_ORACLE_H
#define TARANG_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int N1, N2, N3;          /* grid; reference N[1..3] */
    double f1, f2, f3;       /* kfactor[1..3] */
    int basis;               /* 0 = FOUR, 1 = SCFT */
    int dealias;             /* 1 = "DEALIAS" (2/3 rule), 0 = "ALIAS" */
} orc_grid;

/* ---- 1D FFT core (own mixed-radix Stockham; radices 2,3,4,5 + generic) ---- */
void orc_fft1d(int n, int sign, double *re_im_inout);        /* unnormalised, sign -1 forward */
void orc_set_threads(int nthreads);
int  orc_get_threads(void);

/* ---- wavenumbers, binning primitives (four_inline.h, scft_inline.h) ---- */
int    orc_kx(const orc_grid *g, int l1);
int    orc_ky(const orc_grid *g, int l2);
double orc_kmag(const orc_grid *g, int l1, int l2, int l3);
double orc_multiplicity(const orc_grid *g, int l1, int l2, int l3);
int    orc_min_radius_outside(const orc_grid *g);
int    orc_max_radius_inside(const orc_grid *g);
double orc_approx_modes_in_shell(const orc_grid *g, int radius);
int    orc_shell_ek_size(const orc_grid *g);
int    orc_kept(const orc_grid *g, int l1, int l2, int l3);   /* 1 if mode survives Dealias_array */

/* ---- transforms (four_tr.cc, scft_tr.cc); A is (N1,N2,N3/2+1) complex ---- */
void orc_forward(const orc_grid *g, double *A, int parity);   /* Forward_transform_array */
void orc_inverse(const orc_grid *g, double *A, int parity);   /* Inverse_transform_array */

/* ---- spectral ops ---- */
void orc_deriv(const orc_grid *g, int axis, const double *A, double *B, int parity);
void orc_dealias(const orc_grid *g, double *A);
void orc_real_mult(const orc_grid *g, const double *A, const double *B, double *C);
void orc_divide_ksqr(const orc_grid *g, double *A);
void orc_exp_ksqr(const orc_grid *g, double *A, double factor);
void orc_satisfy_reality(const orc_grid *g, double *A);

/* ---- NS nonlinear term, pressure, rhs (compute_nlin.cc, compute_pressure.cc, compute_rhs.cc) ---- */
void orc_nlin_ns(const orc_grid *g, double *const V[3], double *nlin[3]);
void orc_pressure(const orc_grid *g, double *const nlin[3], double *p);
void orc_rhs(const orc_grid *g, double *nlin[3], const double *p);
/* One RK4 integrating-factor step of NS (time_advance.cc:100-155 with the loop body of
 * Ifluid_main.cc:258-301); force = 0 (compute_force_decay.cc). */
void orc_rk4_step_ns(const orc_grid *g, double *V[3], double nu, double dt);
void orc_euler_step_ns(const orc_grid *g, double *V[3], double nu, double dt);
void orc_rk2_step_ns(const orc_grid *g, double *V[3], double nu, double dt);
/* energy/helicity forcing, force_field_proc 1 (supply) and 2 (constant level): nf, F, forced RK4 step */
int  orc_force_supply_modes(const orc_grid *g, double inner, double outer);
int  orc_force_ek_hk(const orc_grid *g, int proc, double *const V[3], double *F[3], double inner, double outer,
                     double p3, double p4);
void orc_step_ns_forced(const orc_grid *g, double *V[3], double nu, double dt, int scheme, int proc, double inner,
                        double outer, double p3, double p4, double p5, double *keep[3]);   /* scheme 0 EULER, 1 RK2, 2 RK4 */
void orc_force_abc(const orc_grid *g, double *F[3], int k0, double A, double B, double C, double amp);
void orc_rk4_stage1_ns(const orc_grid *g, double *V[3], double nu, double dt);

/* ---- MHD (compute_nlin_VF_MHD.cc, time_advance_VF_MHD.cc) ---- */
void orc_nlin_mhd(const orc_grid *g, double *U[3], double *B[3], double *nU[3], double *nB[3]);
void orc_rk4_step_mhd(const orc_grid *g, double *U[3], double *B[3], double nu, double eta, double dt);
void orc_rk4_step_mhd_forced(const orc_grid *g, double *U[3], double *B[3], double nu, double eta, double dt,
                             int proc, double inner, double outer, double p3, double p4, double p5, double p6,
                             double *keep[6]);
int orc_force_mhd(const orc_grid *g, int proc, double *const U[3], double *const B[3], double *FU[3], double *FW[3],
                  double inner, double outer, double p3, double p4, double p5, double p6);

/* ---- energies and spectra (four_energy.cc, scft_energy.cc, Cvf.cc:385-449) ---- */
double orc_energy(const orc_grid *g, const double *A, const double *B);   /* B==NULL -> A */
double orc_Sn(const orc_grid *g, const double *A, double n);
void   orc_shell_spectrum(const orc_grid *g, const double *A, const double *B, double n,
                          double *Sk, double *Sk_count);
double orc_dt(const orc_grid *g, double *const V[3], double Tdt_fixed);   /* compute_dt.cc:47-59 */

/* ---- energy transfer (EnergyTrdir, universal_ET.cc, four_ET.cc) ---- */
void   orc_fill_array_shell(const orc_grid *g, const double *A, double *B, double inner, double outer);
double orc_shell_mult_single(const orc_grid *g, const double *A, const double *B, double inner, double outer);
void   orc_shell_mult_all(const orc_grid *g, double *const A[3], double *const B[3],
                          const double *radii, int nr, double *result);
void   orc_et_nlin(const orc_grid *g, double *const H[3], double *const G[3], double *nlin[3]);
void   orc_flux_ns(const orc_grid *g, double *const V[3], const double *radii, int nsph, double *flux);
void   orc_shell_tr_ns(const orc_grid *g, double *const V[3], const double *radii, int nsh, double *T);
void   orc_helicity(const orc_grid *g, double *const V[3], double *out);
/* anisotropic spectra, axis a = 1, 2, 3: ring (Max_radius_inside+1) x (nsect+1), cylinder
 * orc_cylinder_shell_ek_size(g, a) x (nslab+1), both row-major */
void   orc_ring_spectrum(const orc_grid *g, double *const V[3], int nsect, int n, int a, double *S1, double *S2);
void   orc_ring_tr(const orc_grid *g, int cyl, int a, double *const H[3], double *const F[3], double *const R[3],
                   int ncomp, const double *radii, int nr, const double *edges, int nsec, int ngiv, double sign,
                   double *out);
int    orc_cylinder_shell_ek_size(const orc_grid *g, int a);
void   orc_aniso_helicity(const orc_grid *g, double *const V[3], int mode, int nbin, int a, double *Hk);
void   orc_cylinder_ring_spectrum(const orc_grid *g, double *const V[3], int nslab, int n, int a, double *S1,
                                  double *S2);
void   orc_shell_helicity(const orc_grid *g, double *const V[3], double *Hk1, double *Hk2, double *Hk3);
double orc_entropy(const orc_grid *g, double *const V[3]);
double orc_entropy_scalar(const orc_grid *g, const double *A);
void   orc_et_nlin_scalar(const orc_grid *g, double *const H[3], const double *G, double *nlin);
void   orc_flux_scalar(const orc_grid *g, double *const V[3], const double *T, const double *radii, int nsph,
                       double *flux_self, double *flux_sf);
void   orc_shell_tr_scalar(const orc_grid *g, double *const V[3], const double *T, const double *radii, int nsh,
                           double *Tself, double *Tsf);
void   orc_shell_tr_mhd(const orc_grid *g, double *const U[3], double *const B[3], const double *radii, int nsh,
                        double *uu, double *ww, double *uw, double *zp, double *zm);
void   orc_flux_mhd(const orc_grid *g, double *U[3], double *B[3], const double *radii, int nsph,
                    double *u2u, double *b2b, double *u2b_in_out, double *u2b_in_in,
                    double *b2u_in_out, double *b2u_in_in, double *u_out_b_out,
                    double *zp, double *zm);
void   orc_et_radii(const orc_grid *g, int scheme, int no_shells, double *radii);

/* ---- RB convection with SCFT (compute_nlin_SF.cc, time_advance_SF.cc, nusselt.cc) ---- */
void   orc_nlin_rb(const orc_grid *g, double *V[3], double *T, double *nV[3], double *nT);
void   orc_rk4_step_rb(const orc_grid *g, double *V[3], double *T, double Ra, double Pr, double dt,
                       double *keep[4]);   /* keep (may be NULL): the last stage V1..V3, T */
double orc_nusselt(const orc_grid *g, double *const V[3], const double *T, double Ra, double Pr);

/* ---- initial conditions ---- */
void orc_ic_taylor_green(const orc_grid *g, double *V[3], int k0, double amp);
void orc_ic_random(const orc_grid *g, double *V[3], unsigned long long seed, double E0,
                   double kp, int project);
void orc_ic_random_scalar(const orc_grid *g, double *T, unsigned long long seed, double amp);

#ifdef __cplusplus
}
#endif
#endif
