// Batched rooms: many rooms of different sizes advance in one launch per step.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "acfdtd_hip.h"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {
//
// Many rooms of different sizes advance in one launch per step: a table of tiles
// (room, z-block, y-block, x-chunk) covers every room, each workgroup looks its tile up and
// runs the single-room march on that room's arguments, which sit in device memory (one
// BatchRoom per room, read through wave-uniform scalar loads). The XCD remap applies to the
// global tile index, so each XCD owns a contiguous run of tiles of one or a few rooms.

template <typename real>
struct BatchRoom {
    SweepArgs<real> args;
    FusedIO<real> forward;   // A -> B
    FusedIO<real> backward;  // B -> A
};

struct BatchTile {
    int32_t room, kb, jb, xc;
};

__device__ inline int batch_slot(int n_tiles, int launched) {
    const unsigned n = (unsigned)launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);
    return b < (unsigned)n_tiles ? (int)b : -1;
}

template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(512) void batch_fused_step(const BatchRoom<real> *rooms,
                                                        const BatchTile *tiles, const int n_tiles,
                                                        const int launched, const int reverse,
                                                        const int64_t step, const int64_t column) {
    const int slot = batch_slot(n_tiles, launched);
    if (slot < 0) return;
    const BatchTile t = tiles[slot];
    const BatchRoom<real> &room = rooms[t.room];
    const SweepArgs<real> &a = room.args;
    const int k = t.kb * 64 + (int)threadIdx.x;
    const int j = t.jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = t.xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    const FusedIO<real> &io = reverse ? room.backward : room.forward;
    fused_march<real, NR, WALLS, LAYER, COMPENSATED>(
        a, i0, i1, j, k, step, column, io.p_in, io.vx_in, io.vy_in, io.vz_in, io.p_out,
        io.vx_out, io.vy_out, io.vz_out, a.psi, a.wall_table, a.cell[0], a.face[0], a.cell[1],
        a.face[1], a.cell[2], a.face[2], a.recording);
}

template <typename real, bool LAYER, bool FUSED_FMA>
__global__ __launch_bounds__(512) void batch_velocity_sweep(const BatchRoom<real> *rooms,
                                                            const BatchTile *tiles,
                                                            const int n_tiles, const int launched) {
    const int slot = batch_slot(n_tiles, launched);
    if (slot < 0) return;
    const BatchTile t = tiles[slot];
    const SweepArgs<real> &a = rooms[t.room].args;
    const int k = t.kb * 64 + (int)threadIdx.x;
    const int j = t.jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = t.xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    velocity_march<real, LAYER, FUSED_FMA>(a, i0, i1, j, k, a.p, a.vx, a.vy, a.vz, a.cell[0],
                                           a.face[0], a.cell[1], a.face[1], a.cell[2],
                                           a.face[2]);
}

template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(512) void batch_pressure_sweep(const BatchRoom<real> *rooms,
                                                            const BatchTile *tiles,
                                                            const int n_tiles, const int launched,
                                                            const int64_t step,
                                                            const int64_t column) {
    const int slot = batch_slot(n_tiles, launched);
    if (slot < 0) return;
    const BatchTile t = tiles[slot];
    const SweepArgs<real> &a = rooms[t.room].args;
    const int k = t.kb * 64 + (int)threadIdx.x;
    const int j = t.jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = t.xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    pressure_march<real, NR, WALLS, LAYER, COMPENSATED>(
        a, i0, i1, j, k, step, column, a.p, a.psi, a.vx, a.vy, a.vz, a.wall_table, a.cell[0],
        a.cell[1], a.cell[2], a.recording);
}

template <typename real>
using BatchFused = void (*)(const BatchRoom<real> *, const BatchTile *, const int, const int,
                            const int, const int64_t, const int64_t);
template <typename real>
using BatchVelocity = void (*)(const BatchRoom<real> *, const BatchTile *, const int, const int);
template <typename real>
using BatchPressure = void (*)(const BatchRoom<real> *, const BatchTile *, const int, const int,
                               const int64_t, const int64_t);

template <typename real, int NR>
void pick_batch_nr(bool walls, bool layer, bool comp, BatchFused<real> &f, BatchPressure<real> &p) {
    const int key = (walls ? 1 : 0) | (layer ? 2 : 0) | (comp ? 4 : 0);
    switch (key) {
#define ACFDTD_BATCH_CASE(K, W, L, C)                      \
    case K:                                                \
        f = batch_fused_step<real, NR, W, L, C>;           \
        p = batch_pressure_sweep<real, NR, W, L, C>;       \
        return;
        ACFDTD_BATCH_CASE(0, false, false, false)
        ACFDTD_BATCH_CASE(1, true, false, false)
        ACFDTD_BATCH_CASE(2, false, true, false)
        ACFDTD_BATCH_CASE(3, true, true, false)
        ACFDTD_BATCH_CASE(4, false, false, true)
        ACFDTD_BATCH_CASE(5, true, false, true)
        ACFDTD_BATCH_CASE(6, false, true, true)
        default:
        ACFDTD_BATCH_CASE(7, true, true, true)
#undef ACFDTD_BATCH_CASE
    }
}

template <typename real>
void pick_batch(int nr, bool walls, bool layer, bool comp, BatchFused<real> &f,
                BatchPressure<real> &p, BatchVelocity<real> &v) {
    switch (nr) {
        case 0: pick_batch_nr<real, 0>(walls, layer, comp, f, p); break;
        case 1: pick_batch_nr<real, 1>(walls, layer, comp, f, p); break;
        case 2: pick_batch_nr<real, 2>(walls, layer, comp, f, p); break;
        case 3: pick_batch_nr<real, 3>(walls, layer, comp, f, p); break;
        default: pick_batch_nr<real, 4>(walls, layer, comp, f, p); break;
    }
    if (layer) v = comp ? batch_velocity_sweep<real, true, true> : batch_velocity_sweep<real, true, false>;
    else v = comp ? batch_velocity_sweep<real, false, true> : batch_velocity_sweep<real, false, false>;
}

}  // namespace acfdtd

// Host handle of a batch: where its tables live (in the caller's workspace) and what kernels
// its rooms need.
struct acfdtd_hip_batch {
    int n_rooms;
    int n_tiles;
    int launched;
    int tile_j;
    int real_bytes;
    int n_relax;
    bool walls, layer, comp, swap;
    void *rooms;  // device: n_rooms BatchRoom<real>
    void *tiles;  // device: n_tiles BatchTile
};

namespace acfdtd {

inline size_t batch_room_bytes(int real_bytes) {
    return real_bytes == 8 ? sizeof(BatchRoom<double>) : sizeof(BatchRoom<float>);
}

inline size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

int batch_validate(const acfdtd_hip_desc *rooms, int32_t n_rooms, int64_t *n_tiles) {
    if (rooms == nullptr || n_rooms < 1) return fail(ACFDTD_HIP_EINVAL, "a batch needs rooms");
    int64_t total = 0;
    for (int r = 0; r < n_rooms; ++r) {
        const int status = validate(&rooms[r]);
        if (status != ACFDTD_HIP_OK)
            return fail(status, "room " + std::to_string(r) + ": " + g_last_error);
        const acfdtd_hip_desc &a = rooms[0], &b = rooms[r];
        if (b.real_bytes != a.real_bytes || b.n_relax != a.n_relax || b.has_layer != a.has_layer ||
            b.flags != a.flags || b.tile_j != a.tile_j ||
            has_swap(b) != has_swap(a))
            return fail(ACFDTD_HIP_EINVAL, "rooms of a batch must share precision, relaxation "
                                           "count, sponge presence, flags, tile_j and swap use");
        if (b.wall_tape) return fail(ACFDTD_HIP_EINVAL, "batched rooms do not take wall tapes");
        total += tiling_for(b).tiles;
    }
    if (total > (int64_t)INT32_MAX - 8) return fail(ACFDTD_HIP_EINVAL, "too many tiles");
    *n_tiles = total;
    return ACFDTD_HIP_OK;
}

template <typename real>
std::string pack_rooms(const acfdtd_hip_desc *rooms, int n_rooms) {
    std::string blob(sizeof(BatchRoom<real>) * (size_t)n_rooms, '\0');
    for (int r = 0; r < n_rooms; ++r) {
        BatchRoom<real> room;
        std::memset(&room, 0, sizeof(room));
        fill_args(rooms[r], room.args);
        if (has_swap(rooms[r])) {
            room.forward = fused_io<real>(rooms[r], false);
            room.backward = fused_io<real>(rooms[r], true);
        }
        std::memcpy(&blob[sizeof(room) * (size_t)r], &room, sizeof(room));
    }
    return blob;
}

std::string pack_tiles(const acfdtd_hip_desc *rooms, int n_rooms) {
    std::string blob;
    for (int r = 0; r < n_rooms; ++r) {
        const Tiling t = tiling_for(rooms[r]);
        for (int xc = 0; xc < t.nchunk; ++xc)
            for (int jb = 0; jb < t.nblk_j; ++jb)
                for (int kb = 0; kb < t.nblk_k; ++kb) {
                    const BatchTile tile{r, kb, jb, xc};
                    blob.append(reinterpret_cast<const char *>(&tile), sizeof(tile));
                }
    }
    return blob;
}

template <typename real>
int batch_steps(const acfdtd_hip_batch &b, int64_t step_offset, int64_t n_steps, int64_t column,
                hipStream_t stream) {
    BatchFused<real> fk;
    BatchPressure<real> pk;
    BatchVelocity<real> vk;
    pick_batch<real>(b.n_relax, b.walls, b.layer, b.comp, fk, pk, vk);
    const BatchRoom<real> *rooms = static_cast<const BatchRoom<real> *>(b.rooms);
    const BatchTile *tiles = static_cast<const BatchTile *>(b.tiles);
    const dim3 block(64, b.tile_j), grid(b.launched);
    int64_t n = 0;
    if (b.swap)
        for (; n + 1 < n_steps; n += 2) {
            hipLaunchKernelGGL(fk, grid, block, 0, stream, rooms, tiles, b.n_tiles, b.launched, 0,
                               step_offset + n, column + n);
            hipLaunchKernelGGL(fk, grid, block, 0, stream, rooms, tiles, b.n_tiles, b.launched, 1,
                               step_offset + n + 1, column + n + 1);
        }
    for (; n < n_steps; ++n) {
        hipLaunchKernelGGL(vk, grid, block, 0, stream, rooms, tiles, b.n_tiles, b.launched);
        hipLaunchKernelGGL(pk, grid, block, 0, stream, rooms, tiles, b.n_tiles, b.launched,
                           step_offset + n, column + n);
    }
    return check_hip(hipGetLastError(), "batch launch");
}

}  // namespace acfdtd

using namespace acfdtd;

extern "C" size_t acfdtd_hip_batch_workspace_bytes(const acfdtd_hip_desc *rooms, int32_t n_rooms) {
    int64_t n_tiles = 0;
    if (batch_validate(rooms, n_rooms, &n_tiles) != ACFDTD_HIP_OK) return 0;
    return align_up(batch_room_bytes(rooms[0].real_bytes) * (size_t)n_rooms, 256) +
           sizeof(BatchTile) * (size_t)n_tiles;
}

extern "C" int acfdtd_hip_batch_create(const acfdtd_hip_desc *rooms, int32_t n_rooms,
                                       void *workspace, size_t workspace_bytes, void *stream,
                                       acfdtd_hip_batch **out) {
    if (out == nullptr) return fail(ACFDTD_HIP_EINVAL, "output handle pointer is NULL");
    *out = nullptr;
    int64_t n_tiles = 0;
    int status = batch_validate(rooms, n_rooms, &n_tiles);
    if (status != ACFDTD_HIP_OK) return status;
    const size_t room_bytes = align_up(batch_room_bytes(rooms[0].real_bytes) * (size_t)n_rooms, 256);
    if (workspace == nullptr || workspace_bytes < room_bytes + sizeof(BatchTile) * (size_t)n_tiles)
        return fail(ACFDTD_HIP_EINVAL, "batch workspace is NULL or too small");
    acfdtd_hip_batch *b = new acfdtd_hip_batch{};
    b->n_rooms = n_rooms;
    b->n_tiles = (int)n_tiles;
    b->launched = ceil_div(n_tiles, 8) * 8;
    b->tile_j = tiling_for(rooms[0]).tile_j;
    b->real_bytes = rooms[0].real_bytes;
    b->n_relax = rooms[0].n_relax;
    b->layer = rooms[0].has_layer != 0;
    b->comp = (rooms[0].flags & ACFDTD_HIP_COMPENSATED) != 0;
    b->swap = has_swap(rooms[0]);
    b->walls = false;
    for (int r = 0; r < n_rooms; ++r) b->walls = b->walls || rooms[r].wall_mask != 0;
    b->rooms = workspace;
    b->tiles = static_cast<char *>(workspace) + room_bytes;
    const std::string tiles = pack_tiles(rooms, n_rooms);
    hipStream_t s = static_cast<hipStream_t>(stream);
    status = check_hip(hipMemcpyAsync(b->tiles, tiles.data(), tiles.size(), hipMemcpyHostToDevice, s),
                       "batch tile upload");
    if (status == ACFDTD_HIP_OK) status = acfdtd_hip_batch_update(b, rooms, n_rooms, stream);
    if (status != ACFDTD_HIP_OK) {
        delete b;
        return status;
    }
    *out = b;
    return ACFDTD_HIP_OK;
}

extern "C" int acfdtd_hip_batch_update(acfdtd_hip_batch *batch, const acfdtd_hip_desc *rooms,
                                       int32_t n_rooms, void *stream) {
    if (batch == nullptr) return fail(ACFDTD_HIP_EINVAL, "batch handle is NULL");
    int64_t n_tiles = 0;
    const int status = batch_validate(rooms, n_rooms, &n_tiles);
    if (status != ACFDTD_HIP_OK) return status;
    if (n_rooms != batch->n_rooms || n_tiles != batch->n_tiles ||
        rooms[0].real_bytes != batch->real_bytes)
        return fail(ACFDTD_HIP_EINVAL, "rooms do not match the batch they update");
    const std::string blob = rooms[0].real_bytes == 8 ? pack_rooms<double>(rooms, n_rooms)
                                                      : pack_rooms<float>(rooms, n_rooms);
    // The upload is synchronous so the host blob may be freed on return.
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st = check_hip(hipMemcpyAsync(batch->rooms, blob.data(), blob.size(), hipMemcpyHostToDevice, s),
                       "batch room upload");
    if (st == ACFDTD_HIP_OK) st = check_hip(hipStreamSynchronize(s), "batch room upload");
    return st;
}

extern "C" int acfdtd_hip_batch_run(acfdtd_hip_batch *batch, int64_t step_offset, int64_t n_steps,
                                    int64_t recording_column, void *stream) {
    if (batch == nullptr) return fail(ACFDTD_HIP_EINVAL, "batch handle is NULL");
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (batch->real_bytes == 8) return batch_steps<double>(*batch, step_offset, n_steps, recording_column, s);
    return batch_steps<float>(*batch, step_offset, n_steps, recording_column, s);
}

extern "C" void acfdtd_hip_batch_destroy(acfdtd_hip_batch *batch) { delete batch; }
