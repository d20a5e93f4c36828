// Shared pieces of the sweep kernels: tiling, field layout, kernel arguments, small numerics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"

namespace acfdtd {

// Message of the last failure on this thread (defined in sweeps.hip).
extern thread_local std::string g_last_error;
int fail(int code, const std::string &message);
int check_hip(hipError_t status, const char *what);

// Workgroup tiling of a room: 64 lanes along z, tile_j rows along y, chunk planes along x.
// `launched` rounds the tile count up to a multiple of the 8 XCDs so that the remap in
// decode_tile is a bijection; the surplus workgroups exit at once.
struct Tiling {
    int vec;  // consecutive z cells per lane
    int tile_j;
    int chunk;
    int nblk_k;
    int nblk_j;
    int nchunk;
    int tiles;
    int launched;
};

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

inline Tiling make_tiling(int nx, int ny, int nz, int tile_j, int chunk, int vec = 1) {
    Tiling t{};
    t.vec = vec;
    t.tile_j = tile_j > 0 ? tile_j : 4;
    t.nblk_k = ceil_div(nz, 64 * vec);
    t.nblk_j = ceil_div(ny, t.tile_j);
    if (chunk <= 0) {
        // Enough workgroups to fill 256 CUs several times over, but never fewer than 4 planes
        // per march, so the one re-read halo plane stays a small fraction of the traffic.
        const int columns = t.nblk_k * t.nblk_j;
        const int wanted = ceil_div(4096, columns);
        chunk = ceil_div(nx, wanted > 0 ? wanted : 1);
        if (chunk < 4) chunk = 4;
    }
    t.chunk = chunk;
    t.nchunk = ceil_div(nx, chunk);
    t.tiles = t.nblk_k * t.nblk_j * t.nchunk;
    t.launched = ceil_div(t.tiles, 8) * 8;
    return t;
}

inline Tiling tiling_for(const acfdtd_hip_desc &d) {
    return make_tiling(d.nx, d.ny, d.nz, d.tile_j, d.chunk_x);
}

// Elements per row of the cell fields: nz, or the padded row length z_pitch.
inline int row_cells(const acfdtd_hip_desc &d) { return d.z_pitch > 0 ? d.z_pitch : d.nz; }

// z cells per lane of the fused step: as many as the rows allow (4 bytes x vec per access),
// 4 for fp32 and 2 for fp64 when the row length is a multiple, else fewer. With padded rows the
// last lane of a row may hold fewer cells of the room than its width (see lane_consts).
inline int fused_vec(const acfdtd_hip_desc &d) {
    const int widest = d.real_bytes == 4 ? 4 : 2;
    for (int v = widest; v > 1; v /= 2)
        if (row_cells(d) % v == 0) return v;
    return 1;
}

// Tiling of the fused kernels: the (y, z / vec) plane is flattened and cut into workgroups of
// 64 * tile_j lanes (see flat_lane), so nblk_k is 1 and nblk_j counts those workgroups.
inline Tiling flat_tiling(const acfdtd_hip_desc &d, int vec) {
    const int64_t lanes = (int64_t)d.ny * (row_cells(d) / vec);
    const int tile_j = d.tile_j > 0 ? d.tile_j : 4;
    // Four planes per march measured fastest at 256^3 and 384^3 (fewer halo re-reads pay less
    // than the extra workgroups gain).
    const int chunk = d.chunk_x > 0 ? d.chunk_x : 4;
    return make_tiling(d.nx, ceil_div(lanes, 64 * tile_j) * tile_j, 64 * vec, tile_j, chunk, vec);
}

inline Tiling fused_tiling(const acfdtd_hip_desc &d) { return flat_tiling(d, fused_vec(d)); }

// Consecutive workgroup ids go to the 8 XCDs in turn; give XCD x the contiguous tile range
// [x * launched / 8, (x + 1) * launched / 8). Tiles are ordered z-block, then y-block, then
// x-chunk.
__device__ inline bool decode_tile(const Tiling &t, int &kb, int &jb, int &xc) {
    const unsigned n = (unsigned)t.launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);
    if (b >= (unsigned)t.tiles) return false;
    kb = (int)(b % (unsigned)t.nblk_k);
    b /= (unsigned)t.nblk_k;
    jb = (int)(b % (unsigned)t.nblk_j);
    xc = (int)(b / (unsigned)t.nblk_j);
    return true;
}

// Offsets of cell (i, j, k) and of the lower y / z faces of that cell in the staggered arrays.
struct Layout {
    int64_t plane, plane_y, plane_z, jk, jk_z;
    __device__ Layout(int nx, int ny, int nz, int j, int k)
        : plane((int64_t)ny * nz),
          plane_y((int64_t)(ny + 1) * nz),
          plane_z((int64_t)ny * (nz + 1)),
          jk((int64_t)j * nz + k),
          jk_z((int64_t)j * (nz + 1) + k) {
        (void)nx;
    }
    // Rows of zp cells (p, psi, vx, vy) and of zpz faces (vz): the layout above when zp = nz and
    // zpz = nz + 1, a padded one when both are the row length z_pitch.
    struct Rows {};
    __device__ Layout(Rows, int ny, int zp, int zpz, int j, int k)
        : plane((int64_t)ny * zp),
          plane_y((int64_t)(ny + 1) * zp),
          plane_z((int64_t)ny * zpz),
          jk((int64_t)j * zp + k),
          jk_z((int64_t)j * zpz + k) {}
    __device__ int64_t cell(int i) const { return (int64_t)i * plane + jk; }
    __device__ int64_t face_y(int i) const { return (int64_t)i * plane_y + jk; }
    __device__ int64_t face_z(int i) const { return (int64_t)i * plane_z + jk_z; }
};

// Wall class of index i on an axis of n cells: bit 0 low face, bit 1 high face.
__device__ inline int wall_class(int i, int n) { return (i == 0 ? 1 : 0) | (i == n - 1 ? 2 : 0); }

// A load of read-only data (never written during the launch) at a wavefront-uniform address,
// through the scalar data cache: it counts in the scalar memory counter, so a wait for its value
// does not also wait for the wavefront's vector memory operations in flight (vector loads and
// stores complete in issue order, so a vector load issued late is only awaited together with
// everything issued before it, such as the next plane's inputs).
template <typename T>
__device__ inline T uniform_load(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// The cells a layer tape holds (the absorbing layer's frame) and where each lies in a row:
// whole planes i < xl or i >= nx - xh; in the other planes whole rows j < yl or j >= ny - yh,
// and in the other rows the cells k < zl or k >= nz - zh (zl, zh multiples of 4, so the cells
// of a lane of 1, 2 or 4 are all in or all out when nz is a multiple of the lane width).
struct LayerFrame {
    int nx, ny, nz, xl, xh, yl, yh, zl, zh;
    int64_t whole, frame, cells, stride;
    __host__ __device__ LayerFrame() : nx(0), ny(0), nz(0), xl(0), xh(0), yl(0), yh(0), zl(0),
                                       zh(0), whole(0), frame(0), cells(0), stride(0) {}
    __host__ __device__ LayerFrame(int nx_, int ny_, int nz_, const int32_t *lo, const int32_t *hi)
        : nx(nx_), ny(ny_), nz(nz_) {
        auto clampn = [](int v, int n) { return v < 0 ? 0 : (v > n ? n : v); };
        xl = clampn(lo[0], nx);
        xh = clampn(hi[0], nx);
        if (xl + xh >= nx) xl = nx, xh = 0;
        yl = clampn(lo[1], ny);
        yh = clampn(hi[1], ny);
        zl = (clampn(lo[2], nz) + 3) / 4 * 4;
        zh = (clampn(hi[2], nz) + 3) / 4 * 4;
        if (zl + zh >= nz) zl = nz, zh = 0, yl = ny, yh = 0;
        if (yl + yh >= ny) yl = ny, yh = 0;
        whole = (int64_t)ny * nz;
        frame = (int64_t)(yl + yh) * nz + (int64_t)(ny - yl - yh) * (zl + zh);
        cells = plane_base(nx);
        const int64_t far = (int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny;
        stride = (4 * cells + far + 3) / 4 * 4;
    }
    __host__ __device__ bool plane_whole(int i) const { return i < xl || i >= nx - xh; }
    // The first frame cell of plane i.
    __host__ __device__ int64_t plane_base(int i) const {
        if (i <= xl) return (int64_t)i * whole;
        const int mid = nx - xl - xh;
        const int in_mid = i - xl < mid ? i - xl : mid;
        const int past = i - (nx - xh) > 0 ? i - (nx - xh) : 0;
        return (int64_t)xl * whole + (int64_t)in_mid * frame + (int64_t)past * whole;
    }
    // Position of cell (j, k) within a plane that is not whole, or -1 outside the frame.
    __host__ __device__ int64_t in_frame(int j, int k) const {
        const int mid = ny - yl - yh, zw = zl + zh;
        if (j < yl) return (int64_t)j * nz + k;
        if (j >= yl + mid) return (int64_t)yl * nz + (int64_t)mid * zw + (int64_t)(j - yl - mid) * nz + k;
        const int64_t row = (int64_t)yl * nz + (int64_t)(j - yl) * zw;
        if (k < zl) return row + k;
        if (k >= nz - zh) return row + zl + (k - (nz - zh));
        return -1;
    }
    // Offset of cell (i, j, k) in a row's p part (-1 outside the frame), given in_frame(j, k).
    __host__ __device__ int64_t offset(int i, int j, int k, int64_t frame_pos) const {
        if (plane_whole(i)) return plane_base(i) + (int64_t)j * nz + k;
        return frame_pos < 0 ? -1 : plane_base(i) + frame_pos;
    }
    // Parts of a row: p, vx, vy, vz of the frame cells, then the far faces.
    __host__ __device__ int64_t far_x() const { return 4 * cells; }
    __host__ __device__ int64_t far_y() const { return 4 * cells + (int64_t)ny * nz; }
    __host__ __device__ int64_t far_z() const { return 4 * cells + (int64_t)ny * nz + (int64_t)nx * nz; }
};

template <typename real>
struct SweepArgs {
    Tiling tiles;
    int nx, ny, nz;
    // Row lengths: cells (p, psi, vx, vy) and vz faces; nz and nz + 1 unless rows are padded.
    int zp, zpz;
    int n_sources, n_receivers;
    real vel, pres, src, relax_scale;
    real decay[ACFDTD_HIP_MAX_RELAXATION];
    real mean[ACFDTD_HIP_MAX_RELAXATION];
    real gain[ACFDTD_HIP_MAX_RELAXATION];
    uint64_t wall_mask;
    const real *wall_table;
    real *p, *vx, *vy, *vz, *psi;
    const real *cell[3];
    const real *face[3];
    const int32_t *source_cell;
    const int64_t *source_length;
    const real *source_signal;
    int64_t source_stride;
    const int32_t *receiver_cell;
    real *recording;
    int64_t recording_stride;
    real *wall_tape;
    int64_t wall_tape_stride;
    int64_t wall_tape_step0;
    // Tape rows start 16-byte aligned (aligned base, stride a multiple of 4 elements), so a
    // lane's V cells of an x or y face row move as one vector access.
    int wall_tape_rows16;
    // Divergence tape: row n - div_tape_step0 of nx*ny*nz values holds step n's divergences.
    real *div_tape;
    int64_t div_tape_step0;
    // Layer tape: row n - layer_tape_step0 holds the frame's fields after step n.
    real *layer_tape;
    int64_t layer_tape_step0;
    int64_t layer_tape_stride;
    LayerFrame layer;
};

// The layout of a room's fields from its arguments (padded rows or not).
template <typename real>
__device__ inline Layout room_layout(const SweepArgs<real> &a, int j, int k) {
    return Layout(Layout::Rows{}, a.ny, a.zp, a.zpz, j, k);
}

// Offsets of the six face planes in one row of the boundary tape, and the row length.
struct TapePlanes {
    int64_t x_lo, x_hi, y_lo, y_hi, z_lo, z_hi, stride;
    __host__ __device__ TapePlanes(int nx, int ny, int nz) {
        const int64_t yz = (int64_t)ny * nz, xz = (int64_t)nx * nz, xy = (int64_t)nx * ny;
        x_lo = 0;
        x_hi = yz;
        y_lo = 2 * yz;
        y_hi = 2 * yz + xz;
        z_lo = 2 * (yz + xz);
        z_hi = z_lo + xy;
        stride = 2 * (yz + xz + xy);
    }
};

// Write v on every face plane cell (i, j, k) lies on (wall code `code`, nonzero).
template <typename real>
__device__ inline void tape_store(real *row, const TapePlanes &t, int nx, int ny, int nz, int i,
                                  int j, int k, int code, real v) {
    (void)nx;
    const int cx = code & 3, cy = (code >> 2) & 3, cz = (code >> 4) & 3;
    const int64_t yz = (int64_t)j * nz + k, xz = (int64_t)i * nz + k, xy = (int64_t)i * ny + j;
    if (cx & 1) row[t.x_lo + yz] = v;
    if (cx & 2) row[t.x_hi + yz] = v;
    if (cy & 1) row[t.y_lo + xz] = v;
    if (cy & 2) row[t.y_hi + xz] = v;
    if (cz & 1) row[t.z_lo + xy] = v;
    if (cz & 2) row[t.z_hi + xy] = v;
}

// Read the value of cell (i, j, k) from the first face plane it lies on.
template <typename real>
__device__ inline real tape_load(const real *row, const TapePlanes &t, int ny, int nz, int i,
                                 int j, int k, int code) {
    const int cx = code & 3, cy = (code >> 2) & 3;
    if (cx & 1) return row[t.x_lo + (int64_t)j * nz + k];
    if (cx & 2) return row[t.x_hi + (int64_t)j * nz + k];
    if (cy & 1) return row[t.y_lo + (int64_t)i * nz + k];
    if (cy & 2) return row[t.y_hi + (int64_t)i * nz + k];
    if (code & 16) return row[t.z_lo + (int64_t)i * ny + j];
    return row[t.z_hi + (int64_t)i * ny + j];
}

template <typename real>
void fill_args(const acfdtd_hip_desc &d, SweepArgs<real> &a) {
    a.tiles = tiling_for(d);
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.zp = row_cells(d);
    a.zpz = d.z_pitch > 0 ? d.z_pitch : d.nz + 1;
    a.n_sources = d.n_sources;
    a.n_receivers = d.n_receivers;
    a.vel = (real)d.vel_coef;
    a.pres = (real)d.pres_coef;
    a.src = (real)d.src_coef;
    a.relax_scale = (real)d.relax_scale;
    for (int r = 0; r < ACFDTD_HIP_MAX_RELAXATION; ++r) {
        a.decay[r] = (real)d.relax_decay[r];
        a.mean[r] = (real)d.relax_mean[r];
        a.gain[r] = (real)d.relax_gain[r];
    }
    a.wall_mask = d.wall_mask;
    a.wall_table = static_cast<const real *>(d.wall_table);
    a.p = static_cast<real *>(d.p);
    a.vx = static_cast<real *>(d.vx);
    a.vy = static_cast<real *>(d.vy);
    a.vz = static_cast<real *>(d.vz);
    a.psi = static_cast<real *>(d.psi);
    for (int q = 0; q < 3; ++q) {
        a.cell[q] = static_cast<const real *>(d.cell[q]);
        a.face[q] = static_cast<const real *>(d.face[q]);
    }
    a.source_cell = d.source_cell;
    a.source_length = d.source_length;
    a.source_signal = static_cast<const real *>(d.source_signal);
    a.source_stride = d.source_stride;
    a.receiver_cell = d.receiver_cell;
    a.recording = static_cast<real *>(d.recording);
    a.recording_stride = d.recording_stride;
    a.wall_tape = static_cast<real *>(d.wall_tape);
    a.wall_tape_stride = d.wall_tape_stride;
    a.wall_tape_step0 = d.wall_tape_step0;
    a.wall_tape_rows16 = d.wall_tape != nullptr &&
                         reinterpret_cast<uintptr_t>(d.wall_tape) % 16 == 0 &&
                         d.wall_tape_stride % 4 == 0;
    a.div_tape = static_cast<real *>(d.div_tape);
    a.div_tape_step0 = d.div_tape_step0;
    a.layer_tape = static_cast<real *>(d.layer_tape);
    a.layer_tape_step0 = d.layer_tape_step0;
    a.layer_tape_stride = d.layer_tape_stride;
    a.layer = LayerFrame(d.nx, d.ny, d.nz, d.layer_lo, d.layer_hi);
}

// Whether the descriptor's divergence tape holds rows for steps lo .. hi (true without a tape).
inline bool div_tape_covers(const acfdtd_hip_desc &d, int64_t lo, int64_t hi) {
    return d.div_tape == nullptr || hi < lo ||
           (lo >= d.div_tape_step0 && hi - d.div_tape_step0 < d.div_tape_rows);
}

// The same for the layer tape.
inline bool layer_tape_covers(const acfdtd_hip_desc &d, int64_t lo, int64_t hi) {
    return d.layer_tape == nullptr || hi < lo ||
           (lo >= d.layer_tape_step0 && hi - d.layer_tape_step0 < d.layer_tape_rows);
}

// `padded`: the caller's kernels take padded rows (z_pitch > 0); the other entry points refuse
// them.
inline int validate(const acfdtd_hip_desc *d, const bool padded = false) {
    if (d == nullptr) return fail(ACFDTD_HIP_EINVAL, "descriptor is NULL");
    if (d->abi_version != ACFDTD_HIP_ABI_VERSION)
        return fail(ACFDTD_HIP_EABI, "descriptor ABI version " + std::to_string(d->abi_version) +
                                         " does not match library version " +
                                         std::to_string(ACFDTD_HIP_ABI_VERSION));
    if (d->struct_size != sizeof(acfdtd_hip_desc))
        return fail(ACFDTD_HIP_EABI, "descriptor is " + std::to_string(d->struct_size) +
                                         " bytes, library expects " +
                                         std::to_string(sizeof(acfdtd_hip_desc)));
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(ACFDTD_HIP_EINVAL, "cell counts must be >= 1");
    if (d->reserved != 0) return fail(ACFDTD_HIP_EINVAL, "reserved descriptor field must be 0");
    if (d->z_pitch != 0) {
        if (!padded)
            return fail(ACFDTD_HIP_EINVAL,
                        "padded rows (z_pitch) are taken by acfdtd_hip_run and the batch API only");
        if (d->z_pitch <= d->nz)
            return fail(ACFDTD_HIP_EINVAL, "z_pitch must be 0 or greater than nz");
        if (d->wall_tape) return fail(ACFDTD_HIP_EINVAL, "padded rows take no wall tape");
        if (d->div_tape) return fail(ACFDTD_HIP_EINVAL, "padded rows take no divergence tape");
        if (d->layer_tape) return fail(ACFDTD_HIP_EINVAL, "padded rows take no layer tape");
    }
    if (d->layer_tape) {
        if (!d->has_layer) return fail(ACFDTD_HIP_EINVAL, "a layer tape needs an absorbing layer");
        const LayerFrame f(d->nx, d->ny, d->nz, d->layer_lo, d->layer_hi);
        if (d->layer_tape_stride < f.stride || d->layer_tape_stride % 4 != 0 ||
            d->layer_tape_rows < 0 || d->layer_tape_step0 < 0 ||
            reinterpret_cast<uintptr_t>(d->layer_tape) % 16 != 0)
            return fail(ACFDTD_HIP_EINVAL,
                        "layer tape: 16-byte aligned, stride a multiple of 4 of at least "
                        "acfdtd_hip_layer_tape_stride, rows from a step >= 0");
    }
    if (d->div_tape && (d->n_relax < 1 || d->div_tape_rows < 0 || d->div_tape_step0 < 0))
        return fail(ACFDTD_HIP_EINVAL,
                    "a divergence tape needs relaxation states and rows from a step >= 0");
    if ((int64_t)d->ny * row_cells(*d) > INT32_MAX)
        return fail(ACFDTD_HIP_EINVAL, "ny * nz must fit in 31 bits");
    if (d->real_bytes != 4 && d->real_bytes != 8)
        return fail(ACFDTD_HIP_EINVAL, "real_bytes must be 4 or 8");
    if (d->n_relax < 0 || d->n_relax > ACFDTD_HIP_MAX_RELAXATION)
        return fail(ACFDTD_HIP_EINVAL, "n_relax out of range");
    if ((d->flags & ~(ACFDTD_HIP_COMPENSATED | ACFDTD_HIP_STAGED)) != 0)
        return fail(ACFDTD_HIP_EINVAL, "unknown flags");
    if ((d->flags & ACFDTD_HIP_COMPENSATED) && d->real_bytes != 4)
        return fail(ACFDTD_HIP_EINVAL, "compensated summation is an fp32 mode");
    if (d->tile_j < 0 || d->tile_j > 8) return fail(ACFDTD_HIP_EINVAL, "tile_j must be 0 to 8");
    if (d->chunk_x < 0) return fail(ACFDTD_HIP_EINVAL, "chunk_x must be >= 0");
    if (d->time_block < 0 || d->time_block > 2)
        return fail(ACFDTD_HIP_EINVAL, "time_block must be 0, 1 or 2");
    if (!d->p || !d->vx || !d->vy || !d->vz) return fail(ACFDTD_HIP_EINVAL, "field pointer is NULL");
    if (d->n_relax > 0 && !d->psi) return fail(ACFDTD_HIP_EINVAL, "relaxation states are NULL");
    if (d->wall_mask != 0 && !d->wall_table) return fail(ACFDTD_HIP_EINVAL, "wall table is NULL");
    if (d->has_layer)
        for (int q = 0; q < 3; ++q)
            if (!d->cell[q] || !d->face[q]) return fail(ACFDTD_HIP_EINVAL, "sponge profile is NULL");
    if (d->n_sources < 0 || d->n_receivers < 0)
        return fail(ACFDTD_HIP_EINVAL, "negative source or receiver count");
    if (d->n_sources > 0 &&
        (!d->source_cell || !d->source_length || !d->source_signal || d->source_stride < 0))
        return fail(ACFDTD_HIP_EINVAL, "source arrays are NULL");
    if (d->n_receivers > 0 && (!d->receiver_cell || !d->recording))
        return fail(ACFDTD_HIP_EINVAL, "receiver arrays are NULL");
    if (d->wall_tape && (d->wall_tape_stride < TapePlanes(d->nx, d->ny, d->nz).stride ||
                         d->wall_tape_rows < 0 || d->wall_tape_step0 < 0))
        return fail(ACFDTD_HIP_EINVAL, "wall tape stride, rows or first step out of range");
    const Tiling t = tiling_for(*d);
    if ((int64_t)t.nblk_k * t.nblk_j * t.nchunk > (int64_t)INT32_MAX - 8)
        return fail(ACFDTD_HIP_EINVAL, "grid too large for one launch");
    return ACFDTD_HIP_OK;
}

// ------------------------------------------------------------------------------ numerics --

__device__ inline float fma_real(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_real(double a, double b, double c) { return __builtin_fma(a, b, c); }

// v - vel * d, either as the reference's two roundings or as one fma.
template <typename real, bool FUSED>
__device__ inline real face_update(real v, real vel, real d) {
    if (FUSED) return fma_real(-vel, d, v);
    return v - vel * d;
}

// s += t with the rounding error of the addition accumulated into e (Knuth's TwoSum).
template <typename real>
__device__ inline void two_sum_acc(real &s, real t, real &e) {
    const real x = s + t;
    const real tb = x - s;
    const real sa = x - tb;
    e += (s - sa) + (t - tb);
    s = x;
}

__device__ inline double wave_sum(double v) {
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
    return v;
}

// Sum three values over a (64, blockDim.y) workgroup, in a fixed order; result in thread 0.
__device__ inline void block_sum3(double &a, double &b, double &c, double (*lds)[8]) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (threadIdx.x == 0) {
        lds[0][threadIdx.y] = a;
        lds[1][threadIdx.y] = b;
        lds[2][threadIdx.y] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        for (unsigned w = 1; w < blockDim.y; ++w) {
            a += lds[0][w];
            b += lds[1][w];
            c += lds[2][w];
        }
    }
}

}  // namespace acfdtd
