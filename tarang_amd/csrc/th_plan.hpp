// th_plan.hpp -- the opaque th_plan / th_state structs of the C-ABI.
#pragma once
#include <vector>
#include "th_internal.hpp"

namespace th {
// scratch slots (full reference-layout fields) used by the plain path
enum {
    SC_VR0 = 0,      // Vr[0..2]   (RVF real-space copies)
    SC_VFT = 3,      // VF_temp
    SC_WR0 = 4,      // W.Vr[0..2] / giver G real space
    SC_WFT = 7,      // W.VF_temp / VF_temp2
    SC_TR = 8,       // T.Fr
    SC_SFT = 9,      // T.SF_temp
    SC_N0 = 10,      // rhs / nlin components 10..16
    SC_COPY0 = 17,   // Vcopy 17..23
    SC_TOT0 = 24,    // tot_Vrhs 24..30
    SC_P = 31,       // pressure
    SC_F1 = 32,      // Force1 (RB)
    SC_FT = 33,      // T.Force (RB)
    SC_G0 = 34,      // ET giver G1..G3 34..36
    SC_NE0 = 37,     // ET nlin 37..39
    SC_FU0 = 40,     // Force1..3 of U (forced MHD) 40..42
    SC_FW0 = 43,     // W.Force1..3 (forced MHD) 43..45
};
struct Fused;        // dealias-pruned pipeline (th_fused.hip)
}  // namespace th

struct th_plan {
    th_ctx* ctx = nullptr;
    th::Geo g;
    int prec = 0;                    // TH_FP64 / TH_FP32_MIXED
    th::TwiddleCache tw;
    th::RedScratch rs;
    std::vector<void*> scratch;
    double* radii_dev = nullptr;
    int radii_cap = 0;
    th::Fused* fused = nullptr;
    int nstates = 0;                 // fused states created on this plan and not yet destroyed
    bool dead = false;               // th_plan_destroy called while states were alive
    void* scr(int i);
    template <typename R> const th::cplx<R>* twd(int L);
    ~th_plan();
};

namespace th {
void release_fused(th_plan* p);
void plan_free(th_plan* p);           // waits for the plan's stream, frees it, uncounts it in its context
template <typename R> int plain_forward(th_plan* p, cplx<R>* A, int parity);
template <typename R> int plain_inverse(th_plan* p, cplx<R>* A, int parity);
}  // namespace th
