// th_fused_f32.hip -- the fused pipeline's kernels and launchers instantiated for float storage.
#include "th_fused_kern.hpp"

namespace th {
TH_FUSED_INSTANTIATE(float)
}  // namespace th
