// th_fused_f64.hip -- the fused pipeline's kernels and launchers instantiated for double storage.
#include "th_fused_kern.hpp"

namespace th {
TH_FUSED_INSTANTIATE(double)
}  // namespace th
