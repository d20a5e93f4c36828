// fp32 instantiation of the two-step sweep (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template int64_t launch_fused2_quads<float>(const acfdtd_hip_desc &, int64_t, int64_t, int64_t,
                                            hipStream_t);

template int launch_fused2_once<float>(const acfdtd_hip_desc &, int64_t, int64_t, bool,
                                        hipStream_t);

}  // namespace acfdtd
