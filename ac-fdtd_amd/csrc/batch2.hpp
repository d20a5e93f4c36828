// Batched rooms, two steps per sweep: the two-step sweep's workgroup (fused2_block) on a table of
// (room, tile) entries that covers every room, one launch per two steps. Each room keeps its own
// two-step tiling (its rows of row-length / V lane groups, whole or in segments); the lane width
// V is the batch's, and rooms whose nz it does not divide come with padded rows (z_pitch), so
// the kernels take padded rows. Templates only; batch2*.hip instantiate them per precision and
// lane width.

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

// A pointer read from device memory is generic (flat) to the compiler, and flat accesses also
// count in the LDS counter, so each LDS wait would wait for them. The field pointers all point
// to global memory: the empty asm hands each one back as a global pointer (the compiler
// folds a plain cast pair away), and the accesses through it become global ones. The pointers
// are the same for the whole workgroup, hence scalar registers.
template <typename T>
__device__ __forceinline__ T *global_ptr(T *p) {
    __attribute__((address_space(1))) T *g;
    asm("" : "=s"(g) : "0"((__attribute__((address_space(1))) T *)p));
    return (T *)g;
}

template <typename real>
__device__ __forceinline__ void global_ptrs(SweepArgs<real> &a) {
    a.wall_table = global_ptr(a.wall_table);
    a.p = global_ptr(a.p);
    a.vx = global_ptr(a.vx);
    a.vy = global_ptr(a.vy);
    a.vz = global_ptr(a.vz);
    a.psi = global_ptr(a.psi);
    for (int d = 0; d < 3; ++d) {
        a.cell[d] = global_ptr(a.cell[d]);
        a.face[d] = global_ptr(a.face[d]);
    }
    a.source_cell = global_ptr(a.source_cell);
    a.source_length = global_ptr(a.source_length);
    a.source_signal = global_ptr(a.source_signal);
    a.receiver_cell = global_ptr(a.receiver_cell);
    a.recording = global_ptr(a.recording);
    a.wall_tape = global_ptr(a.wall_tape);
}

template <typename real>
__device__ __forceinline__ void global_ptrs(Fused2IO<real> &io) {
    io.p_in = global_ptr(io.p_in);
    io.vx_in = global_ptr(io.vx_in);
    io.vy_in = global_ptr(io.vy_in);
    io.vz_in = global_ptr(io.vz_in);
    io.psi_in = global_ptr(io.psi_in);
    io.p_out = global_ptr(io.p_out);
    io.vx_out = global_ptr(io.vx_out);
    io.vy_out = global_ptr(io.vy_out);
    io.vz_out = global_ptr(io.vz_out);
    io.psi_out = global_ptr(io.psi_out);
}

template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(kFused2Threads, ACFDTD_FUSED2_MIN_WAVES) void batch_fused2_step(
    const BatchRoom2<real> *__restrict__ rooms, const BatchTile2 *__restrict__ tiles,
    const int n_tiles, const int launched, const int reverse, const int64_t step,
    const int64_t column) {
    ACFDTD_FUSED2_SHARED(real, V, NR, false);
    const unsigned n = (unsigned)launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);  // contiguous tiles per XCD, as decode_tile
    if (b >= (unsigned)n_tiles) return;   // whole workgroup, before any barrier
    const BatchTile2 tile = tiles[b];
    const BatchRoom2<real> &room = rooms[tile.room];
    SweepArgs<real> args = room.args;
    Fused2IO<real> io = reverse ? room.backward : room.forward;
    global_ptrs(args);
    global_ptrs(io);
    fused2_block<real, V, NR, WALLS, LAYER, COMPENSATED, false, true>(
        args, io, room.t2, (unsigned)tile.tile, step, column, lds, src_lds, len_lds, now_lds,
        rx_lds, stage);
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
