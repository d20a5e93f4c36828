// Two-step sweep, fp32: the kernels of 2-cell lanes (nz a multiple of 2, not of 4)
// (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<float> pick_fused2_kind<float, 2, false>(int, bool, bool, bool);

}  // namespace acfdtd
