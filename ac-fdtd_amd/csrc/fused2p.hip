// Two-step sweep, fp32, 4-cell lanes on padded rows (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<float> pick_fused2_padded<float, 4>(int, bool, bool, bool);

}  // namespace acfdtd
