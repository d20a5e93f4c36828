// fp32 instantiation of the two-step adjoint sweep carrying the sponge-strength gradient
// (templates in adjoint2.hpp).

#include "adjoint2.hpp"

namespace acfdtd {

template Adjoint2Kernel<float> pick_adjoint2_sponge<float>(int, bool);

}  // namespace acfdtd
