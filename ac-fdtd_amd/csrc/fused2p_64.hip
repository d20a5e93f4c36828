// Two-step sweep, fp64, 2-cell lanes on padded rows (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<double> pick_fused2_padded<double, 2>(int, bool, bool, bool);

}  // namespace acfdtd
