// Two-step sweep, fp64: the staged kernels (2-cell lanes)
// (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<double> pick_fused2_kind<double, 2, true>(int, bool, bool, bool);

}  // namespace acfdtd
