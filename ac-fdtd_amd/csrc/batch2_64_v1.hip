// Batched rooms, two steps per sweep, fp64, 1-cell lanes (templates in batch2.hpp).

#include "batch2.hpp"

namespace acfdtd {

template Batch2Kernel<double> pick_batch2_kind<double, 1>(int, bool, bool, bool);

}  // namespace acfdtd
