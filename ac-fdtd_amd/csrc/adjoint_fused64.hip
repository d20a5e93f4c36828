// fp64 instantiation of the fused adjoint step (templates in adjoint_fused.hpp).

#include "adjoint_fused.hpp"

namespace acfdtd {

template int64_t launch_adjoint_pairs<double>(const acfdtd_hip_desc &,
                                              const acfdtd_hip_adjoint_desc &, int64_t, int64_t,
                                              hipStream_t);

}  // namespace acfdtd
