// Two-step sweep, fp64, 2-cell lanes, writing the divergence tape (templates in fused2.hpp and
// marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<double> pick_fused2_divt<double, 2>(int, bool, bool, bool);

}  // namespace acfdtd
