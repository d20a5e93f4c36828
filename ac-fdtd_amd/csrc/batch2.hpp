// Batched rooms, two steps per sweep: the two-step sweep's workgroup (fused2_block) on a table of
// (room, tile) entries that covers every room, one launch per two steps. Each room keeps its own
// two-step tiling (its rows of nz / V lane groups, whole or in segments); the lane width V is
// the batch's. Templates only; batch2*.hip instantiate them per precision and lane width.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "acfdtd_hip.h"
#include "fused2.hpp"
#include "marches2.hpp"
#include "sweeps.hpp"

namespace acfdtd {

template <typename real>
struct BatchRoom2 {
    SweepArgs<real> args;
    Fused2IO<real> forward;   // A -> B
    Fused2IO<real> backward;  // B -> A
    Tiling2 t2;
};

struct BatchTile2 {
    int32_t room, tile;  // the room, and the workgroup's tile within it (< its t2.tiles)
};

template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(kFused2Threads, ACFDTD_FUSED2_MIN_WAVES) void batch_fused2_step(
    const BatchRoom2<real> *rooms, const BatchTile2 *tiles, const int n_tiles, const int launched,
    const int reverse, const int64_t step, const int64_t column) {
    ACFDTD_FUSED2_SHARED(real, V, NR, false);
    const unsigned n = (unsigned)launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);  // contiguous tiles per XCD, as decode_tile
    if (b >= (unsigned)n_tiles) return;   // whole workgroup, before any barrier
    const BatchTile2 tile = tiles[b];
    const BatchRoom2<real> &room = rooms[tile.room];
    fused2_block<real, V, NR, WALLS, LAYER, COMPENSATED, false>(
        room.args, reverse ? room.backward : room.forward, room.t2, (unsigned)tile.tile, step,
        column, lds, src_lds, len_lds, now_lds, rx_lds, stage);
}

template <typename real>
using Batch2Kernel = void (*)(const BatchRoom2<real> *, const BatchTile2 *, const int, const int,
                              const int, const int64_t, const int64_t);

template <typename real, int V, int NR>
Batch2Kernel<real> pick_batch2_nr(bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (comp) {
            if (walls) return layer ? batch_fused2_step<real, V, NR, true, true, true>
                                    : batch_fused2_step<real, V, NR, true, false, true>;
            return layer ? batch_fused2_step<real, V, NR, false, true, true>
                         : batch_fused2_step<real, V, NR, false, false, true>;
        }
    }
    (void)comp;
    if (walls) return layer ? batch_fused2_step<real, V, NR, true, true, false>
                            : batch_fused2_step<real, V, NR, true, false, false>;
    return layer ? batch_fused2_step<real, V, NR, false, true, false>
                 : batch_fused2_step<real, V, NR, false, false, false>;
}

template <typename real, int V>
Batch2Kernel<real> pick_batch2_kind(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_batch2_nr<real, V, 0>(walls, layer, comp);
        case 1: return pick_batch2_nr<real, V, 1>(walls, layer, comp);
        case 2: return pick_batch2_nr<real, V, 2>(walls, layer, comp);
        case 3: return pick_batch2_nr<real, V, 3>(walls, layer, comp);
        default: return pick_batch2_nr<real, V, 4>(walls, layer, comp);
    }
}

// Each (precision, lane width) family is instantiated in a translation unit of its own.
extern template Batch2Kernel<float> pick_batch2_kind<float, 4>(int, bool, bool, bool);
extern template Batch2Kernel<float> pick_batch2_kind<float, 2>(int, bool, bool, bool);
extern template Batch2Kernel<float> pick_batch2_kind<float, 1>(int, bool, bool, bool);
extern template Batch2Kernel<double> pick_batch2_kind<double, 2>(int, bool, bool, bool);
extern template Batch2Kernel<double> pick_batch2_kind<double, 1>(int, bool, bool, bool);

template <typename real>
Batch2Kernel<real> pick_batch2(int vec, int nr, bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) return pick_batch2_kind<real, 4>(nr, walls, layer, comp);
    }
    if (vec == 2) return pick_batch2_kind<real, 2>(nr, walls, layer, comp);
    return pick_batch2_kind<real, 1>(nr, walls, layer, comp);
}

}  // namespace acfdtd
