// Batched rooms: many rooms of different sizes advance in one launch per step.
//
// A table of tiles (room, plane block, x-chunk) covers every room; each workgroup looks its
// tile up and runs the fused step's march on that room's arguments, which sit in device memory
// (one BatchRoom per room, read through wave-uniform scalar loads; its field pointers are
// declared global, see global_ptr). The lanes of a room flatten
// its (y, z / V) plane as in the single-room fused kernel, with one V for the whole batch (the
// widest every room's row length allows; a room given padded rows, z_pitch, takes lanes as
// wide as the pitch allows, its last lane per row partly padding). The XCD remap applies to the
// global tile index, so each XCD owns a contiguous run of tiles of one or a few rooms. Steps go
// in fused pairs (A -> B -> A); an odd trailing step runs room by room through the two in-place
// sweeps.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "acfdtd_hip.h"
#include "batch2.hpp"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {

template <typename real>
struct BatchRoom {
    SweepArgs<real> args;
    FusedIO<real> forward;   // A -> B
    FusedIO<real> backward;  // B -> A
};

struct BatchTile {
    int32_t room, jb, xc, pad;
};

template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(512) void batch_fused_step(const BatchRoom<real> *__restrict__ rooms,
                                                        const BatchTile *__restrict__ tiles,
                                                        const int n_tiles, const int launched,
                                                        const int reverse, const int64_t step,
                                                        const int64_t column) {
    const unsigned n = (unsigned)launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);
    if (b >= (unsigned)n_tiles) return;
    const BatchTile t = tiles[b];
    const BatchRoom<real> &room = rooms[t.room];
    SweepArgs<real> a = room.args;
    global_ptrs(a);
    int j, k0;
    if (!flat_lane(a.tiles, a.ny, a.zp, t.jb, j, k0)) return;
    const int i0 = t.xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    FusedIO<real> io = reverse ? room.backward : room.forward;
    io.p_in = global_ptr(io.p_in);
    io.vx_in = global_ptr(io.vx_in);
    io.vy_in = global_ptr(io.vy_in);
    io.vz_in = global_ptr(io.vz_in);
    io.p_out = global_ptr(io.p_out);
    io.vx_out = global_ptr(io.vx_out);
    io.vy_out = global_ptr(io.vy_out);
    io.vz_out = global_ptr(io.vz_out);
    fused_march_vec<real, V, NR, WALLS, LAYER, COMPENSATED, true>(
        a, i0, i1, j, k0, step, column, io.p_in, io.vx_in, io.vy_in, io.vz_in, io.p_out,
        io.vx_out, io.vy_out, io.vz_out, a.psi, a.wall_table, a.cell[0], a.face[0], a.cell[1],
        a.face[1], a.cell[2], a.face[2], a.recording);
}

template <typename real>
using BatchKernel = void (*)(const BatchRoom<real> *, const BatchTile *, const int, const int,
                             const int, const int64_t, const int64_t);

template <typename real, int V, int NR>
BatchKernel<real> pick_batch_nr(bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (comp) {
            if (walls) return layer ? batch_fused_step<real, V, NR, true, true, true>
                                    : batch_fused_step<real, V, NR, true, false, true>;
            return layer ? batch_fused_step<real, V, NR, false, true, true>
                         : batch_fused_step<real, V, NR, false, false, true>;
        }
    }
    (void)comp;
    if (walls) return layer ? batch_fused_step<real, V, NR, true, true, false>
                            : batch_fused_step<real, V, NR, true, false, false>;
    return layer ? batch_fused_step<real, V, NR, false, true, false>
                 : batch_fused_step<real, V, NR, false, false, false>;
}

template <typename real, int V>
BatchKernel<real> pick_batch_v(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_batch_nr<real, V, 0>(walls, layer, comp);
        case 1: return pick_batch_nr<real, V, 1>(walls, layer, comp);
        case 2: return pick_batch_nr<real, V, 2>(walls, layer, comp);
        case 3: return pick_batch_nr<real, V, 3>(walls, layer, comp);
        default: return pick_batch_nr<real, V, 4>(walls, layer, comp);
    }
}

template <typename real>
BatchKernel<real> pick_batch(int vec, int nr, bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) return pick_batch_v<real, 4>(nr, walls, layer, comp);
    }
    if (vec == 2) return pick_batch_v<real, 2>(nr, walls, layer, comp);
    return pick_batch_v<real, 1>(nr, walls, layer, comp);
}

// The batch's lane width: the widest V that divides every room's row length (nz or z_pitch).
inline int batch_vec(const acfdtd_hip_desc *rooms, int n_rooms) {
    int vec = 4;
    for (int r = 0; r < n_rooms; ++r) vec = std::min(vec, fused_vec(rooms[r]));
    return vec;
}

inline Tiling batch_room_tiling(const acfdtd_hip_desc &d, int vec) { return flat_tiling(d, vec); }

inline size_t batch_room_bytes(int real_bytes) {
    return real_bytes == 8 ? sizeof(BatchRoom<double>) : sizeof(BatchRoom<float>);
}

inline size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

int batch_validate(const acfdtd_hip_desc *rooms, int32_t n_rooms, int64_t *n_tiles) {
    if (rooms == nullptr || n_rooms < 1) return fail(ACFDTD_HIP_EINVAL, "a batch needs rooms");
    const int vec = batch_vec(rooms, n_rooms);
    int64_t total = 0;
    for (int r = 0; r < n_rooms; ++r) {
        const int status = validate(&rooms[r], true);
        if (status != ACFDTD_HIP_OK)
            return fail(status, "room " + std::to_string(r) + ": " + g_last_error);
        const acfdtd_hip_desc &a = rooms[0], &b = rooms[r];
        if (b.real_bytes != a.real_bytes || b.n_relax != a.n_relax || b.has_layer != a.has_layer ||
            b.flags != a.flags || b.tile_j != a.tile_j)
            return fail(ACFDTD_HIP_EINVAL, "rooms of a batch must share precision, relaxation "
                                           "count, sponge presence, flags and tile_j");
        if (!has_swap(b)) return fail(ACFDTD_HIP_EINVAL, "batched rooms need swap buffers");
        if (b.wall_tape || b.div_tape || b.layer_tape)
            return fail(ACFDTD_HIP_EINVAL, "batched rooms take no wall tapes, divergence tapes or layer tapes");
        total += batch_room_tiling(b, vec).tiles;
    }
    if (total > (int64_t)INT32_MAX - 8) return fail(ACFDTD_HIP_EINVAL, "too many tiles");
    *n_tiles = total;
    return ACFDTD_HIP_OK;
}

template <typename real>
std::string pack_rooms(const acfdtd_hip_desc *rooms, int n_rooms, int vec) {
    std::string blob(sizeof(BatchRoom<real>) * (size_t)n_rooms, '\0');
    for (int r = 0; r < n_rooms; ++r) {
        BatchRoom<real> room;
        std::memset(&room, 0, sizeof(room));
        fill_args(rooms[r], room.args);
        room.args.tiles = batch_room_tiling(rooms[r], vec);
        room.forward = fused_io<real>(rooms[r], false);
        room.backward = fused_io<real>(rooms[r], true);
        std::memcpy(&blob[sizeof(room) * (size_t)r], &room, sizeof(room));
    }
    return blob;
}

// A room's descriptor as the batched two-step sweep tiles it with lanes of `vec` cells: the
// room's own tile_j and chunk_x, or, where they are 0, per lane width the shape measured best on
// the 64-room workload's rooms of that width (among 4 to 16 rows and 16 to 32 planes): 4-cell
// lanes 12 rows and 24-plane marches (157.7 Gcell/s, next 146.5), 2-cell lanes 8 rows and 32
// planes (125.5, next 118.7), 1-cell lanes 8 rows and 32 planes (108.0, within 2 % of the
// others). The single-room default march length does not apply: it sizes the marches for the
// workgroups of one room, and a batch launch holds every room's.
inline acfdtd_hip_desc batch2_desc(const acfdtd_hip_desc &d, int vec) {
    acfdtd_hip_desc t = d;
    if (t.tile_j <= 0) t.tile_j = vec == 4 ? 6 : 4;
    if (t.chunk_x <= 0) t.chunk_x = vec == 4 ? 24 : 32;
    return t;
}

// Whether every room may advance two steps per sweep (fused2_applies) with the batch's lane
// width; then the number of (room, tile) entries and the widest room workgroup.
inline bool batch_two_step(const acfdtd_hip_desc *rooms, int n_rooms, int vec, int64_t *n_tiles,
                           int *threads) {
    int64_t total = 0;
    int widest = 64;
    for (int r = 0; r < n_rooms; ++r) {
        Tiling2 t2;
        if (!fused2_applies(rooms[r]) || !make_tiling2(batch2_desc(rooms[r], vec), vec, t2)) return false;
        total += t2.tiles;
        widest = std::max(widest, t2.threads);
    }
    if (total > (int64_t)INT32_MAX - 8) return false;
    *n_tiles = total;
    *threads = widest;
    return true;
}

inline size_t batch_room2_bytes(int real_bytes) {
    return real_bytes == 8 ? sizeof(BatchRoom2<double>) : sizeof(BatchRoom2<float>);
}

template <typename real>
std::string pack_rooms2(const acfdtd_hip_desc *rooms, int n_rooms, int vec) {
    std::string blob(sizeof(BatchRoom2<real>) * (size_t)n_rooms, '\0');
    for (int r = 0; r < n_rooms; ++r) {
        BatchRoom2<real> room;
        std::memset(&room, 0, sizeof(room));
        fill_args(rooms[r], room.args);
        fused2_io<real>(rooms[r], room.forward, room.backward);
        make_tiling2(batch2_desc(rooms[r], vec), vec, room.t2);
        std::memcpy(&blob[sizeof(room) * (size_t)r], &room, sizeof(room));
    }
    return blob;
}

std::string pack_tiles2(const acfdtd_hip_desc *rooms, int n_rooms, int vec) {
    std::string blob;
    for (int r = 0; r < n_rooms; ++r) {
        Tiling2 t2;
        make_tiling2(batch2_desc(rooms[r], vec), vec, t2);
        for (int b = 0; b < t2.tiles; ++b) {
            const BatchTile2 tile{r, b};
            blob.append(reinterpret_cast<const char *>(&tile), sizeof(tile));
        }
    }
    return blob;
}

std::string pack_tiles(const acfdtd_hip_desc *rooms, int n_rooms, int vec) {
    std::string blob;
    for (int r = 0; r < n_rooms; ++r) {
        const Tiling t = batch_room_tiling(rooms[r], vec);
        for (int xc = 0; xc < t.nchunk; ++xc)
            for (int jb = 0; jb < t.nblk_j; ++jb) {
                const BatchTile tile{r, jb, xc, 0};
                blob.append(reinterpret_cast<const char *>(&tile), sizeof(tile));
            }
    }
    return blob;
}

}  // namespace acfdtd

// Host handle of a batch: where its tables live (in the caller's workspace), what kernels its
// rooms need, and the rooms' descriptors for steps run room by room.
struct acfdtd_hip_batch {
    int n_rooms;
    int n_tiles;
    int launched;
    int tile_j;
    int vec;
    int real_bytes;
    int n_relax;
    bool walls, layer, comp;
    void *rooms;  // device: n_rooms BatchRoom<real>
    void *tiles;  // device: n_tiles BatchTile
    // Two steps per sweep, when every room allows it: its own room and tile tables.
    bool two_step;
    int n_tiles2, launched2, threads2;
    void *rooms2;  // device: n_rooms BatchRoom2<real>
    void *tiles2;  // device: n_tiles2 BatchTile2
    std::vector<acfdtd_hip_desc> descs;
};

namespace acfdtd {

template <typename real>
int batch_pairs(const acfdtd_hip_batch &b, int64_t step_offset, int64_t n_pairs, int64_t column,
                hipStream_t stream) {
    const BatchKernel<real> k = pick_batch<real>(b.vec, b.n_relax, b.walls, b.layer, b.comp);
    const BatchRoom<real> *rooms = static_cast<const BatchRoom<real> *>(b.rooms);
    const BatchTile *tiles = static_cast<const BatchTile *>(b.tiles);
    const dim3 block(64, b.tile_j), grid(b.launched);
    for (int64_t p = 0; p < n_pairs; ++p) {
        const int64_t n = 2 * p;
        hipLaunchKernelGGL(k, grid, block, 0, stream, rooms, tiles, b.n_tiles, b.launched, 0,
                           step_offset + n, column + n);
        hipLaunchKernelGGL(k, grid, block, 0, stream, rooms, tiles, b.n_tiles, b.launched, 1,
                           step_offset + n + 1, column + n + 1);
    }
    return check_hip(hipGetLastError(), "batch launch");
}

// Two-step sweeps for the leading multiple of four of n_steps (A -> B, B -> A); returns the
// number of steps done.
template <typename real>
int64_t batch_quads(const acfdtd_hip_batch &b, int64_t step_offset, int64_t n_steps,
                    int64_t column, hipStream_t stream) {
    const Batch2Kernel<real> k = pick_batch2<real>(b.vec, b.n_relax, b.walls, b.layer, b.comp);
    const BatchRoom2<real> *rooms = static_cast<const BatchRoom2<real> *>(b.rooms2);
    const BatchTile2 *tiles = static_cast<const BatchTile2 *>(b.tiles2);
    const dim3 block(b.threads2), grid(b.launched2);
    int64_t n = 0;
    for (; n + 3 < n_steps; n += 4) {
        hipLaunchKernelGGL(k, grid, block, 0, stream, rooms, tiles, b.n_tiles2, b.launched2, 0,
                           step_offset + n, column + n);
        hipLaunchKernelGGL(k, grid, block, 0, stream, rooms, tiles, b.n_tiles2, b.launched2, 1,
                           step_offset + n + 2, column + n + 2);
    }
    return n;
}

// Sizes of the workspace parts: one-step room table, tile table, and the two-step ones.
struct BatchLayout {
    size_t rooms, tiles, rooms2, tiles2;
    bool two_step;
    int64_t n_tiles2;
    int threads2;
    size_t total() const { return rooms + tiles + rooms2 + tiles2; }
};

inline BatchLayout batch_layout(const acfdtd_hip_desc *rooms, int n_rooms, int64_t n_tiles) {
    BatchLayout l{};
    const int vec = batch_vec(rooms, n_rooms);
    l.rooms = align_up(batch_room_bytes(rooms[0].real_bytes) * (size_t)n_rooms, 256);
    l.tiles = align_up(sizeof(BatchTile) * (size_t)n_tiles, 256);
    l.two_step = batch_two_step(rooms, n_rooms, vec, &l.n_tiles2, &l.threads2);
    if (l.two_step) {
        l.rooms2 = align_up(batch_room2_bytes(rooms[0].real_bytes) * (size_t)n_rooms, 256);
        l.tiles2 = sizeof(BatchTile2) * (size_t)l.n_tiles2;
    }
    return l;
}

}  // namespace acfdtd

using namespace acfdtd;

extern "C" size_t acfdtd_hip_batch_workspace_bytes(const acfdtd_hip_desc *rooms, int32_t n_rooms) {
    int64_t n_tiles = 0;
    if (batch_validate(rooms, n_rooms, &n_tiles) != ACFDTD_HIP_OK) return 0;
    return batch_layout(rooms, n_rooms, n_tiles).total();
}

extern "C" int acfdtd_hip_batch_create(const acfdtd_hip_desc *rooms, int32_t n_rooms,
                                       void *workspace, size_t workspace_bytes, void *stream,
                                       acfdtd_hip_batch **out) {
    if (out == nullptr) return fail(ACFDTD_HIP_EINVAL, "output handle pointer is NULL");
    *out = nullptr;
    int64_t n_tiles = 0;
    int status = batch_validate(rooms, n_rooms, &n_tiles);
    if (status != ACFDTD_HIP_OK) return status;
    const BatchLayout layout = batch_layout(rooms, n_rooms, n_tiles);
    if (workspace == nullptr || workspace_bytes < layout.total())
        return fail(ACFDTD_HIP_EINVAL, "batch workspace is NULL or too small");
    acfdtd_hip_batch *b = new acfdtd_hip_batch{};
    b->n_rooms = n_rooms;
    b->n_tiles = (int)n_tiles;
    b->launched = ceil_div(n_tiles, 8) * 8;
    b->vec = batch_vec(rooms, n_rooms);
    b->tile_j = batch_room_tiling(rooms[0], b->vec).tile_j;
    b->real_bytes = rooms[0].real_bytes;
    b->n_relax = rooms[0].n_relax;
    b->layer = rooms[0].has_layer != 0;
    b->comp = (rooms[0].flags & ACFDTD_HIP_COMPENSATED) != 0;
    b->walls = false;
    for (int r = 0; r < n_rooms; ++r) b->walls = b->walls || rooms[r].wall_mask != 0;
    char *base = static_cast<char *>(workspace);
    b->rooms = base;
    b->tiles = base + layout.rooms;
    b->two_step = layout.two_step;
    b->n_tiles2 = (int)layout.n_tiles2;
    b->launched2 = ceil_div(layout.n_tiles2, 8) * 8;
    b->threads2 = layout.threads2;
    b->rooms2 = layout.two_step ? base + layout.rooms + layout.tiles : nullptr;
    b->tiles2 = layout.two_step ? base + layout.rooms + layout.tiles + layout.rooms2 : nullptr;
    const std::string tiles = pack_tiles(rooms, n_rooms, b->vec);
    hipStream_t s = static_cast<hipStream_t>(stream);
    status = check_hip(hipMemcpyAsync(b->tiles, tiles.data(), tiles.size(), hipMemcpyHostToDevice, s),
                       "batch tile upload");
    std::string tiles2;
    if (status == ACFDTD_HIP_OK && b->two_step) {
        tiles2 = pack_tiles2(rooms, n_rooms, b->vec);
        status = check_hip(hipMemcpyAsync(b->tiles2, tiles2.data(), tiles2.size(),
                                          hipMemcpyHostToDevice, s),
                           "batch tile upload");
    }
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
        rooms[0].real_bytes != batch->real_bytes || batch_vec(rooms, n_rooms) != batch->vec)
        return fail(ACFDTD_HIP_EINVAL, "rooms do not match the batch they update");
    const BatchLayout layout = batch_layout(rooms, n_rooms, n_tiles);
    if (layout.two_step != batch->two_step ||
        (layout.two_step && (layout.n_tiles2 != batch->n_tiles2 || layout.threads2 != batch->threads2)))
        return fail(ACFDTD_HIP_EINVAL, "rooms do not match the batch they update");
    const std::string blob = rooms[0].real_bytes == 8
                                 ? pack_rooms<double>(rooms, n_rooms, batch->vec)
                                 : pack_rooms<float>(rooms, n_rooms, batch->vec);
    std::string blob2;
    if (batch->two_step)
        blob2 = rooms[0].real_bytes == 8 ? pack_rooms2<double>(rooms, n_rooms, batch->vec)
                                         : pack_rooms2<float>(rooms, n_rooms, batch->vec);
    batch->descs.assign(rooms, rooms + n_rooms);
    // Synchronous, so the host blobs may be freed on return.
    hipStream_t s = static_cast<hipStream_t>(stream);
    int st = check_hip(hipMemcpyAsync(batch->rooms, blob.data(), blob.size(), hipMemcpyHostToDevice, s),
                       "batch room upload");
    if (st == ACFDTD_HIP_OK && batch->two_step)
        st = check_hip(hipMemcpyAsync(batch->rooms2, blob2.data(), blob2.size(),
                                      hipMemcpyHostToDevice, s),
                       "batch room upload");
    if (st == ACFDTD_HIP_OK) st = check_hip(hipStreamSynchronize(s), "batch room upload");
    return st;
}

extern "C" int acfdtd_hip_batch_run(acfdtd_hip_batch *batch, int64_t step_offset, int64_t n_steps,
                                    int64_t recording_column, void *stream) {
    if (batch == nullptr) return fail(ACFDTD_HIP_EINVAL, "batch handle is NULL");
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    for (const acfdtd_hip_desc &d : batch->descs)
        if (d.n_receivers > 0 && recording_column + n_steps > d.recording_stride)
            return fail(ACFDTD_HIP_EINVAL, "a room's recording is too short for this many steps");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t done = 0;
    if (batch->two_step) {
        done = batch->real_bytes == 8
                   ? batch_quads<double>(*batch, step_offset, n_steps, recording_column, s)
                   : batch_quads<float>(*batch, step_offset, n_steps, recording_column, s);
        const int status = check_hip(hipGetLastError(), "batch two-step launch");
        if (status != ACFDTD_HIP_OK) return status;
    }
    const int64_t rest = n_steps - done;
    const int64_t pairs = rest / 2;
    int status = batch->real_bytes == 8
                     ? batch_pairs<double>(*batch, step_offset + done, pairs, recording_column + done, s)
                     : batch_pairs<float>(*batch, step_offset + done, pairs, recording_column + done, s);
    if (status != ACFDTD_HIP_OK || rest % 2 == 0) return status;
    // An odd last step: the two in-place sweeps, room by room (they take padded rows too).
    for (const acfdtd_hip_desc &d : batch->descs) {
        status = d.real_bytes == 8
                     ? inplace_steps<double>(d, step_offset + n_steps - 1, 1,
                                             recording_column + n_steps - 1, s)
                     : inplace_steps<float>(d, step_offset + n_steps - 1, 1,
                                            recording_column + n_steps - 1, s);
        if (status != ACFDTD_HIP_OK) return status;
    }
    return ACFDTD_HIP_OK;
}

extern "C" int acfdtd_hip_batch_two_step(const acfdtd_hip_batch *batch) {
    return batch != nullptr && batch->two_step;
}

extern "C" void acfdtd_hip_batch_destroy(acfdtd_hip_batch *batch) { delete batch; }
