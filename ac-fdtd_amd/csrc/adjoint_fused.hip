// fp32 instantiation of the fused adjoint step (templates in adjoint_fused.hpp).

#include "adjoint_fused.hpp"

namespace acfdtd {

int64_t adjoint_fused_waves(const acfdtd_hip_desc &d) {
    const Tiling t = fused_tiling(d);
    return (int64_t)t.tiles * t.tile_j;
}

template int64_t launch_adjoint_pairs<float>(const acfdtd_hip_desc &,
                                             const acfdtd_hip_adjoint_desc &, int64_t, int64_t,
                                             hipStream_t);

}  // namespace acfdtd
