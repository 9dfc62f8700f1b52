// th_fused.hip -- dealias-pruned fused pipeline (placeholder until the fused passes land).
#include "../../include/tarang_hip.h"
#include "th_internal.hpp"
#include "th_plan.hpp"

using namespace th;

namespace th {
struct Fused {};
void release_fused(th_plan* p)
{
    delete p->fused;
    p->fused = nullptr;
}
}  // namespace th

struct th_state {};

extern "C" int th_state_create(th_plan*, th_state**, int) { set_error("fused path not built"); return E_UNSUPPORTED; }
extern "C" int th_state_destroy(th_state* s) { delete s; return OK; }
extern "C" int th_state_upload(th_state*, int, const void*, int) { return E_UNSUPPORTED; }
extern "C" int th_state_download(th_state*, int, void*, int) { return E_UNSUPPORTED; }
extern "C" int th_step(th_plan*, th_state*, double, const th_phys*, int) { return E_UNSUPPORTED; }
extern "C" int th_flux_shell_tr(th_plan*, th_state*, const double*, int, double*, double*) { return E_UNSUPPORTED; }
