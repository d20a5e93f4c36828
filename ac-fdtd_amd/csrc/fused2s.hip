// fp32 instantiation of the staged two-step sweep (templates in fused2.hpp and marches2.hpp).

#include "fused2.hpp"

namespace acfdtd {

template <typename real>
Fused2Kernel<real> pick_fused2_staged(int nr, bool walls, bool layer, bool comp) {
    return pick_fused2_kind<real, true>(nr, walls, layer, comp);
}

template Fused2Kernel<float> pick_fused2_staged<float>(int, bool, bool, bool);

}  // namespace acfdtd
