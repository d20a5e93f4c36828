// fp32 instantiation of the two-step adjoint sweep (templates in adjoint2.hpp).

#include "adjoint2.hpp"

namespace acfdtd {

template int64_t launch_adjoint2_quads<float>(const acfdtd_hip_desc &,
                                              const acfdtd_hip_adjoint_desc &, int64_t, int64_t,
                                              hipStream_t);

}  // namespace acfdtd
