// Two-step sweep, fp64: the launchers and the kernels of 2-cell lanes
// (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template Fused2Kernel<double> pick_fused2_kind<double, 2, false>(int, bool, bool, bool);

template int64_t launch_fused2_quads<double>(const acfdtd_hip_desc &, int64_t, int64_t, int64_t,
                                             hipStream_t);

template int launch_fused2_once<double>(const acfdtd_hip_desc &, int64_t, int64_t, bool,
                                        hipStream_t);

}  // namespace acfdtd
