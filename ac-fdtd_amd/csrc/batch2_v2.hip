// Batched rooms, two steps per sweep, fp32, 2-cell lanes (templates in batch2.hpp).

#include "batch2.hpp"

namespace acfdtd {

template Batch2Kernel<float> pick_batch2_kind<float, 2>(int, bool, bool, bool);

}  // namespace acfdtd
