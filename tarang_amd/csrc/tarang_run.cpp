// tarang_run.cpp -- decaying-turbulence driver over the C++ host classes, the counterpart of the
// reference's time loop (trunk/src/IFluid/Ifluid_main.cc:258-301): per step Get_dt, Time_advance,
// energy + NaN check, then the shell spectrum and (at the end) the energy flux.
//
//   tarang_run N steps dt nu [tg|random|file=<field_in.d>] [FP64|FP32_MIXED] [fused|plain] [save=<field_out.d>]
//              [out=DIR] [force=PROC,INNER,OUTER,P3,P4[,P5,P6]] [mhd=ETA] [rb=RA,PR] [et=NSH,SCHEME]
//              [ring=NRING,NSECT,NCYL,NSLAB,AXIS,SCHEME] [file_reduced=NR1,NR2,NR3] [save_reduced=NR1,NR2,NR3]
// Output (stdout): "glob t E D" per step, then "spectrum k E(k)" lines and "flux R Pi(R)" lines.
// file=: Init_cond from a field file (field_input_proc = 1; its "%% Time" sets Tinit); save=: the
// final field is appended to a field file (Output_field), the restart point of a later run.
// file_reduced=NR1,NR2,NR3: file= holds a field of that grid (Init_cond_reduced, field_input_proc = 2);
// save_reduced=NR1,NR2,NR3: save= writes the modes of that grid (Output_field_reduced, field_out_reduced.d).
// out=DIR (any position after nu): DIR/out/glob.d every step, DIR/out/spectrum.d and flux.d at the end,
// in the reference's formats (Output_global, Output_shell_spectrum, Output_flux), and DIR/out/shell.d at the end
// (Output_shell_to_shell).
// et=NSH,SCHEME: NSH energy-transfer shells and spheres with shell_input_scheme SCHEME (0: the 2^(n/4)-like
// radii of EnergyTr.cc:136-160; else unit radii); the radii are printed as "radius i r" lines.
// ring=...: ring-to-ring transfers over NRING shells x NSECT polar sectors and cylinder-ring transfers over NCYL
// K_perp shells x NSLAB K_pll slabs about AXIS (Init_ring_tr; SCHEME 0: the 0, 4, 8, log-spaced radii, else unit
// radii; a count of 0 switches that table off); with out=DIR, DIR/out/ring_to_ring.d and cyl_ring_to_ring.d at
// the end (Output_ring_to_ring, Output_cylinder_ring_to_ring).
// force=PROC,INNER,OUTER,P3,P4: force_field_proc 1 (energy/helicity supply) or 2 (constant energy/helicity)
// with its force_double_para, recomputed at every stage; with mhd=, P5 and P6 are the magnetic field's
// parameters (energy_supply_W, epsh_by_k_epse_W or energy_level_W, h_by_k_E_W; 0 when left out).
// mhd=ETA: incompressible MHD (the IMHD program, src/IMHD/IMHD_main.cc) with magnetic diffusivity ETA; U
// from the IC choice, B random-phase (seed 2); Get_dt(W), Time_advance(W) and the W outputs; "glob t Eu Eb"
// lines, the spectrum of U and the kinetic flux.
// rb=RA,PR: Rayleigh-Benard on the SCFT basis (the RB_SLIP program, PRLARGE with ULARGE scaling: nu =
// sqrt(PR/RA), kappa = 1/sqrt(PR RA)), kfactor (pi, pi/sqrt2, pi/sqrt2), V = 0 and a temperature
// perturbation of amplitude 1e-3 (seed 4); "glob t E_V E_T" lines, "nusselt Nu" at the end, and with out=
// glob.d (Output_global_RB) every step, spectrum.d and flux.d (Output_shell_spectrum_scalar, Output_flux(T))
// at the end.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include "tarang_host.hpp"

int main(int argc, char** argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s N steps dt nu [tg|random|file=PATH] [FP64|FP32_MIXED] [fused|plain] [save=PATH] "
                        "[out=DIR] [force=PROC,INNER,OUTER,P3,P4[,P5,P6]] [mhd=ETA] [rb=RA,PR] [et=NSH,SCHEME] "
                        "[ring=NRING,NSECT,NCYL,NSLAB,AXIS,SCHEME] [file_reduced=NR1,NR2,NR3] [save_reduced=NR1,NR2,NR3]\n",
                argv[0]);
        return 2;
    }
    const int N = atoi(argv[1]), steps = atoi(argv[2]);
    const double dt = atof(argv[3]), nu = atof(argv[4]);
    const std::string ic = argc > 5 ? argv[5] : "tg";
    const std::string prec = argc > 6 ? argv[6] : "FP64";
    const std::string path = argc > 7 ? argv[7] : "fused";
    std::string save, outdir;
    int fproc = 0;
    double fpar[6] = {0, 0, 0, 0, 0, 0};
    double eta = -1, Ra = 0, Pr = 0;
    int et_nsh = N / 3 + 2, et_scheme = 1;
    int ring[6] = {0, 0, 0, 0, 3, 1};
    int nin_red[3] = {0, 0, 0}, nout_red[3] = {0, 0, 0};
    for (int a = 5; a < argc; a++) {
        if (strncmp(argv[a], "file_reduced=", 13) == 0 &&
            sscanf(argv[a] + 13, "%d,%d,%d", &nin_red[0], &nin_red[1], &nin_red[2]) != 3) {
            fprintf(stderr, "tarang_run: file_reduced=NR1,NR2,NR3 expected\n");
            return 2;
        }
        if (strncmp(argv[a], "save_reduced=", 13) == 0 &&
            sscanf(argv[a] + 13, "%d,%d,%d", &nout_red[0], &nout_red[1], &nout_red[2]) != 3) {
            fprintf(stderr, "tarang_run: save_reduced=NR1,NR2,NR3 expected\n");
            return 2;
        }
        if (strncmp(argv[a], "save=", 5) == 0) save = argv[a] + 5;
        if (strncmp(argv[a], "out=", 4) == 0) outdir = argv[a] + 4;
        if (strncmp(argv[a], "mhd=", 4) == 0) eta = atof(argv[a] + 4);
        if (strncmp(argv[a], "ring=", 5) == 0 &&
            sscanf(argv[a] + 5, "%d,%d,%d,%d,%d,%d", &ring[0], &ring[1], &ring[2], &ring[3], &ring[4], &ring[5]) != 6) {
            fprintf(stderr, "tarang_run: ring=NRING,NSECT,NCYL,NSLAB,AXIS,SCHEME expected\n");
            return 2;
        }
        if (strncmp(argv[a], "et=", 3) == 0 && sscanf(argv[a] + 3, "%d,%d", &et_nsh, &et_scheme) != 2) {
            fprintf(stderr, "tarang_run: et=NSH,SCHEME expected\n");
            return 2;
        }
        if (strncmp(argv[a], "rb=", 3) == 0 && sscanf(argv[a] + 3, "%lf,%lf", &Ra, &Pr) != 2) {
            fprintf(stderr, "tarang_run: rb=RA,PR expected\n");
            return 2;
        }
        if (strncmp(argv[a], "force=", 6) == 0) {
            const int k = sscanf(argv[a] + 6, "%d,%lf,%lf,%lf,%lf,%lf,%lf", &fproc, &fpar[0], &fpar[1], &fpar[2],
                                 &fpar[3], &fpar[4], &fpar[5]);
            if (k != 5 && k != 7) {
                fprintf(stderr, "tarang_run: force=PROC,INNER,OUTER,P3,P4[,P5,P6] expected\n");
                return 2;
            }
        }
    }
    try {
        tarang::Context ctx(0);
        if (Ra > 0 && Pr > 0) {
            const double kfx = M_PI, kfy = M_PI / sqrt(2.0);
            tarang::Grid g{{0, N, N, N}, {0, kfx, kfy, kfy}, "SCFT", "DEALIAS", prec};
            tarang::IncFluid U(ctx, g, sqrt(Pr / Ra), 0.0, dt, "RK4", true, et_nsh, et_scheme);
            U.Tfinal = steps * dt;
            if (ring[0] || ring[2]) U.Init_ring_tr(ring[0], ring[1], ring[5], ring[2], ring[3], ring[5], ring[4]);
            U.Rayleigh = Ra;
            U.Prandtl = Pr;
            tarang::IncSF T(ctx, g, 1.0 / sqrt(Pr * Ra));
            tarang::check(th_ic_random_scalar(T.plan(), T.F.data(), 4, 1e-3), "th_ic_random_scalar");
            if (!outdir.empty()) U.Open_output_files(outdir);
            for (int it = 0; it < steps; it++) {
                U.Tdt = U.Get_dt(T);
                U.Tnow += U.Tdt;
                U.Time_advance(T);
                U.CV_Compute_totalenergy_diss();
                T.CS_Compute_totalenergy_diss();
                if (isnan(U.CV_total_energy) || isnan(T.CS_total_energy)) {
                    printf("ERROR: Numerical Overflow\n");
                    return 1;
                }
                printf("glob %.6f %.15e %.15e\n", U.Tnow, U.CV_total_energy, T.CS_total_energy);
                if (!outdir.empty()) U.Output_global(T);
            }
            printf("nusselt %.15e\n", U.Get_Nusselt_no(T));
            if (!outdir.empty()) {
                U.Output_shell_spectrum(T);
                U.Output_flux(T);
                U.Output_shell_to_shell(T);
                U.Output_ring_to_ring(T);
                U.Output_cylinder_ring_to_ring(T);
            }
            return 0;
        }
        tarang::Grid g{{0, N, N, N}, {0, 1.0, 1.0, 1.0}, "FOUR", "DEALIAS", prec};
        tarang::IncFluid U(ctx, g, nu, 0.0, dt, "RK4", path == "fused", et_nsh, et_scheme);
        for (int s = 0; s <= U.no_shells; s++) printf("radius %d %.17g\n", s, U.shell_radius[s]);
        if (ic == "tg") {
            U.Setup_Taylor_Green_field(1, 1.0);
        } else if (ic.compare(0, 5, "file=") == 0) {
            // the snapshot's time becomes Tinit
            const double t = nin_red[0] ? U.Init_cond_reduced(ic.substr(5), nin_red) : U.Init_cond(ic.substr(5));
            if (!isnan(t)) U.Tnow = t;
        } else {
            void* V[3] = {U.V1.data(), U.V2.data(), U.V3.data()};
            tarang::check(th_ic_random(U.plan(), V, 1, 0.5, 4.0), "th_ic_random");
        }
        if (fproc && eta < 0) U.Set_force(fproc, fpar[0], fpar[1], fpar[2], fpar[3], fpar[4]);
        U.Tfinal = U.Tnow + steps * dt;
        if (ring[0] || ring[2]) U.Init_ring_tr(ring[0], ring[1], ring[5], ring[2], ring[3], ring[5], ring[4]);
        if (!outdir.empty()) U.Open_output_files(outdir);
        if (eta >= 0) {
            tarang::IncVF W(ctx, g, eta, et_nsh, et_scheme);
            void* B[3] = {W.V1.data(), W.V2.data(), W.V3.data()};
            tarang::check(th_ic_random(W.plan(), B, 2, 0.5, 4.0), "th_ic_random");
            if (fproc) U.Set_force(W, fproc, fpar[0], fpar[1], fpar[2], fpar[3], fpar[4], fpar[5]);
            for (int it = 0; it < steps; it++) {
                U.Tdt = U.Get_dt(W);
                U.Tnow += U.Tdt;
                U.Time_advance(W);
                U.CV_Compute_totalenergy_diss();
                W.CV_Compute_totalenergy_diss();
                if (isnan(U.CV_total_energy) || isnan(W.CV_total_energy)) {
                    printf("ERROR: Numerical Overflow\n");
                    return 1;
                }
                printf("glob %.6f %.15e %.15e\n", U.Tnow, U.CV_total_energy, W.CV_total_energy);
                if (!outdir.empty()) U.Output_global(W);
            }
            std::vector<double> Ek;
            U.CV_Compute_shell_spectrum(Ek);
            for (size_t k = 0; k < Ek.size(); k++) printf("spectrum %zu %.15e\n", k, Ek[k]);
            if (!outdir.empty()) {
                U.Output_shell_spectrum(W);
                U.Output_flux(W);
                U.Output_shell_to_shell(W);
                U.Output_ring_to_ring(W);
                U.Output_cylinder_ring_to_ring(W);
            }
            return 0;
        }
        for (int it = 0; it < steps; it++) {
            U.Tdt = U.Get_dt();
            U.Tnow += U.Tdt;
            U.Time_advance();
            U.CV_Compute_totalenergy_diss();
            if (isnan(U.CV_total_energy)) {
                printf("ERROR: Numerical Overflow\n");
                return 1;
            }
            printf("glob %.6f %.15e %.15e\n", U.Tnow, U.CV_total_energy, U.CV_total_dissipation);
            if (!outdir.empty()) U.Output_global();
        }
        std::vector<double> Ek;
        U.CV_Compute_shell_spectrum(Ek);
        for (size_t k = 0; k < Ek.size(); k++) printf("spectrum %zu %.15e\n", k, Ek[k]);
        U.Compute_flux();
        for (int s = 1; s <= U.no_spheres; s++) printf("flux %g %.15e\n", U.sphere_radius[s], U.flux_self[s]);
        if (!save.empty()) {
            if (nout_red[0]) U.Output_field_reduced(save, nout_red);
            else U.Output_field(save);
        }
        if (!outdir.empty()) {
            U.Output_shell_spectrum();
            U.Output_flux();
            U.Output_shell_to_shell();
            U.Output_ring_to_ring();
            U.Output_cylinder_ring_to_ring();
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "tarang_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
