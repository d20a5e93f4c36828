// Two-step sweep, fp32: the kernels of 1-cell lanes (odd nz)
// (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<float> pick_fused2_kind<float, 1, false>(int, bool, bool, bool);

}  // namespace acfdtd
