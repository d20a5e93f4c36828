// Two-step sweep, fp32: the staged kernels (4-cell lanes)
// (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<float> pick_fused2_kind<float, 4, true>(int, bool, bool, bool);

}  // namespace acfdtd
