// The pressure-only (second-order wave equation) scheme on gfx950: one sweep per step.
//
// The reference plans this variant without implementing it (PLAN.md:52-56, :116); the
// arithmetic is stated in include/acfdtd_hip.h (acfdtd_hip_wave_desc) and restated by the CPU
// oracle, which it reproduces bit for bit (the library is built with -ffp-contract=off).
//
// Mapping to CDNA4. The (y, z) plane is flattened into lane groups of V consecutive z cells
// (V = 4 in fp32 and 2 in fp64 when nz allows, so every access of a lane's own cells is one
// 16-byte load or store), numbered row after row so that wavefronts fill with no idle lanes
// whatever nz is. A workgroup of 64 * tile_j lanes marches a chunk of planes along x and keeps
// the x neighbours in registers (the next plane's vector becomes the current one), so each
// cell of p^n crosses HBM once per step; the y neighbours are the rows of the lanes nz / V
// places away and the z neighbours the edge cells of the adjacent lanes, re-read through
// L1/L2. p^{n-1} is read once and p^{n+1} written over it. That is 3 elements per
// cell-update (12 B fp32), an HBM-bound stencil with no MFMA work. Workgroup ids are remapped
// so each of the 8 XCDs owns a contiguous range of tiles (neighbouring rows share one L2).
// Sources (in insertion order) and receivers are handled by the lanes that own their cells,
// on the values still in registers.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
#include "sweeps.hpp"

namespace acfdtd {

namespace {

template <typename real, int V>
struct alignas(sizeof(real) * V) Cells {
    real c[V];
};

template <typename real, int V>
__device__ __forceinline__ Cells<real, V> load_cells(const real *__restrict__ base, int64_t at) {
    return *reinterpret_cast<const Cells<real, V> *>(base + at);
}

struct WaveTiling {
    int vec;
    int lanes_per_row;  // nz / vec
    int64_t lanes;      // ny * lanes_per_row
    int threads;        // 64 * tile_j
    int blocks_per_plane;
    int chunk;
    int nchunk;
    int tiles;
    int launched;
};

template <typename real>
struct WaveArgs {
    WaveTiling t;
    int nx, ny, nz;
    real courant2;
    const real *wall_table;
    // Sources and receivers (cells as i and j * nz + k), handled by the lanes that own them.
    int n_sources, n_receivers;
    real src;
    const int32_t *source_cell;
    const int64_t *source_length;
    const real *source_signal;
    int64_t source_stride;
    const int32_t *receiver_cell;
    real *recording;
    int64_t recording_stride;
};

__device__ __forceinline__ int axis_class(int i, int n) { return (i == 0 ? 1 : 0) | (i == n - 1 ? 2 : 0); }

template <typename real, int V, bool WALLS>
__global__ __launch_bounds__(512) void wave_step(const WaveArgs<real> a, const real *__restrict__ p,
                                                 real *__restrict__ q, const int64_t step,
                                                 const int64_t column) {
    const WaveTiling &t = a.t;
    unsigned b = blockIdx.x;
    b = (b & 7u) * ((unsigned)t.launched >> 3) + (b >> 3);
    if (b >= (unsigned)t.tiles) return;
    const int blk = (int)(b % (unsigned)t.blocks_per_plane);
    const int xc = (int)(b / (unsigned)t.blocks_per_plane);
    const int64_t g = (int64_t)blk * t.threads + threadIdx.x;
    if (g >= t.lanes) return;
    const int j = (int)(g / t.lanes_per_row);
    const int k0 = (int)(g % t.lanes_per_row) * V;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = xc * t.chunk;
    const int i1 = min(i0 + t.chunk, nx);
    const int64_t plane = (int64_t)ny * nz;
    const int64_t row = (int64_t)j * nz + k0;
    const int jk0 = j * nz + k0;
    // Neighbour offsets within a plane; a missing neighbour is read from the lane's own cells
    // (in bounds) and its flux replaced by 0.
    const bool has_ylo = j > 0, has_yhi = j + 1 < ny;
    const bool has_zlo = k0 > 0, has_zhi = k0 + V < nz;
    const int64_t y_lo = has_ylo ? row - nz : row;
    const int64_t y_hi = has_yhi ? row + nz : row;
    const int64_t z_lo = has_zlo ? row - 1 : row;
    const int64_t z_hi = has_zhi ? row + V : row;
    const int code_y = axis_class(j, ny) << 2;
    const real zero = real(0), two = real(2), lam2 = a.courant2;
    Cells<real, V> below = load_cells<real, V>(p, (int64_t)(i0 > 0 ? i0 - 1 : i0) * plane + row);
    Cells<real, V> here = load_cells<real, V>(p, (int64_t)i0 * plane + row);
    // Whether a source or receiver lies in this workgroup's march (worked out after the first
    // loads are issued, so the table reads overlap them): lane g owns the flattened
    // cells j * nz + k in [g V, g V + V), so the workgroup owns one contiguous range of them.
    const int jk_first = blk * t.threads * V, jk_last = jk_first + t.threads * V;
    bool events = false;
    for (int s = 0; s < a.n_sources; ++s) {
        const int is = a.source_cell[2 * s], jk = a.source_cell[2 * s + 1];
        events = events || (is >= i0 && is < i1 && jk >= jk_first && jk < jk_last);
    }
    for (int r = 0; r < a.n_receivers; ++r) {
        const int ir = a.receiver_cell[2 * r], jk = a.receiver_cell[2 * r + 1];
        events = events || (ir >= i0 && ir < i1 && jk >= jk_first && jk < jk_last);
    }

    for (int i = i0; i < i1; ++i) {
        const int64_t base = (int64_t)i * plane;
        const bool has_xhi = i + 1 < nx;
        const Cells<real, V> above = load_cells<real, V>(p, (has_xhi ? base + plane : base) + row);
        const Cells<real, V> south = load_cells<real, V>(p, base + y_lo);
        const Cells<real, V> north = load_cells<real, V>(p, base + y_hi);
        const real z_before = p[base + z_lo];
        const real z_after = p[base + z_hi];
        const Cells<real, V> old = load_cells<real, V>(q, base + row);
        const int code_xy = WALLS ? (axis_class(i, nx) | code_y) : 0;
        Cells<real, V> out;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const real c = here.c[m];
            const real fx_hi = has_xhi ? above.c[m] - c : zero;
            const real fx_lo = i > 0 ? c - below.c[m] : zero;
            const real fy_hi = has_yhi ? north.c[m] - c : zero;
            const real fy_lo = has_ylo ? c - south.c[m] : zero;
            const real next_z = m + 1 < V ? here.c[m + 1] : z_after;
            const real prev_z = m > 0 ? here.c[m - 1] : z_before;
            const bool z_hi_face = m + 1 < V || has_zhi;
            const bool z_lo_face = m > 0 || has_zlo;
            const real fz_hi = z_hi_face ? next_z - c : zero;
            const real fz_lo = z_lo_face ? c - prev_z : zero;
            const real lap = ((fx_hi - fx_lo) + (fy_hi - fy_lo)) + (fz_hi - fz_lo);
            if (WALLS) {
                const int code = code_xy | (axis_class(k0 + m, nz) << 4);
                const real f = a.wall_table[code];
                const real keep = a.wall_table[64 + code];
                out.c[m] = ((c * two - keep * old.c[m]) + lam2 * lap) * f;
            } else {
                out.c[m] = (c * two - old.c[m]) + lam2 * lap;
            }
        }
        // Sources in insertion order, then receivers, in the few workgroups that hold one.
        const bool plane_events = events;
        for (int s = 0; plane_events && s < a.n_sources; ++s) {
            if (a.source_cell[2 * s] != i) continue;
            const int at = a.source_cell[2 * s + 1] - jk0;
            if (at < 0 || at >= V) continue;
            const int64_t length = a.source_length[s];
            const real *signal = a.source_signal + (int64_t)s * a.source_stride;
            const real now = step < length ? a.src * signal[step] : zero;
            const real before = (step >= 1 && step - 1 < length) ? a.src * signal[step - 1] : zero;
#pragma unroll
            for (int m = 0; m < V; ++m)
                if (m == at) out.c[m] = out.c[m] + (now - before);
        }
        *reinterpret_cast<Cells<real, V> *>(q + base + row) = out;
        for (int r = 0; plane_events && r < a.n_receivers; ++r) {
            if (a.receiver_cell[2 * r] != i) continue;
            const int at = a.receiver_cell[2 * r + 1] - jk0;
            if (at < 0 || at >= V) continue;
            real value = out.c[0];
#pragma unroll
            for (int m = 1; m < V; ++m)
                if (m == at) value = out.c[m];
            a.recording[(int64_t)r * a.recording_stride + column] = value;
        }
        below = here;
        here = above;
    }
}

int wave_vec(const acfdtd_hip_wave_desc &d) {
    const int widest = d.real_bytes == 4 ? 4 : 2;
    for (int v = widest; v > 1; v /= 2)
        if (d.nz % v == 0) return v;
    return 1;
}

WaveTiling wave_tiling(const acfdtd_hip_wave_desc &d) {
    WaveTiling t{};
    t.vec = wave_vec(d);
    t.lanes_per_row = d.nz / t.vec;
    t.lanes = (int64_t)d.ny * t.lanes_per_row;
    t.threads = 64 * (d.tile_j > 0 ? d.tile_j : 4);
    t.blocks_per_plane = ceil_div(t.lanes, t.threads);
    int chunk = d.chunk_x;
    if (chunk <= 0) {
        // About 8192 workgroups of 4 wavefronts (32 per CU), never marches shorter than 4
        // planes (the plane before a march is read by two marches). Measured at 256^3, walls,
        // fp32 / fp64, microseconds per step by marching length 4, 8, 16, 32, 64 planes:
        // 31.2, 32.1, 32.2, 46.7, 80.0 / 54.7, 58.1, 59.6, 56.2, 79.7: the short marches'
        // occupancy outweighs their re-read planes.
        const int wanted = ceil_div(8192, t.blocks_per_plane);
        chunk = ceil_div(d.nx, wanted > 0 ? wanted : 1);
        if (chunk < 4) chunk = 4;
    }
    t.chunk = chunk;
    t.nchunk = ceil_div(d.nx, chunk);
    t.tiles = t.blocks_per_plane * t.nchunk;
    t.launched = ceil_div(t.tiles, 8) * 8;
    return t;
}

int validate_wave(const acfdtd_hip_wave_desc *d) {
    if (d == nullptr) return fail(ACFDTD_HIP_EINVAL, "wave descriptor is NULL");
    if (d->abi_version != ACFDTD_HIP_ABI_VERSION)
        return fail(ACFDTD_HIP_EABI, "wave descriptor ABI version " + std::to_string(d->abi_version) +
                                         " does not match library version " +
                                         std::to_string(ACFDTD_HIP_ABI_VERSION));
    if (d->struct_size != sizeof(acfdtd_hip_wave_desc))
        return fail(ACFDTD_HIP_EABI, "wave descriptor is " + std::to_string(d->struct_size) +
                                         " bytes, library expects " +
                                         std::to_string(sizeof(acfdtd_hip_wave_desc)));
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(ACFDTD_HIP_EINVAL, "cell counts must be >= 1");
    if ((int64_t)d->ny * d->nz > INT32_MAX) return fail(ACFDTD_HIP_EINVAL, "ny * nz must fit in 31 bits");
    if (d->real_bytes != 4 && d->real_bytes != 8)
        return fail(ACFDTD_HIP_EINVAL, "real_bytes must be 4 or 8");
    if (d->tile_j < 0 || d->tile_j > 8) return fail(ACFDTD_HIP_EINVAL, "tile_j must be 0 to 8");
    if (d->chunk_x < 0) return fail(ACFDTD_HIP_EINVAL, "chunk_x must be >= 0");
    if (!d->p || !d->p_prev) return fail(ACFDTD_HIP_EINVAL, "field pointer is NULL");
    if (d->p == d->p_prev) return fail(ACFDTD_HIP_EINVAL, "p and p_prev must be distinct buffers");
    if (d->wall_mask != 0 && !d->wall_table) return fail(ACFDTD_HIP_EINVAL, "wall table is NULL");
    if (d->n_sources < 0 || d->n_receivers < 0)
        return fail(ACFDTD_HIP_EINVAL, "negative source or receiver count");
    if (d->n_sources > 0 &&
        (!d->source_cell || !d->source_length || !d->source_signal || d->source_stride < 0))
        return fail(ACFDTD_HIP_EINVAL, "source arrays are NULL");
    if (d->n_receivers > 0 && (!d->receiver_cell || !d->recording))
        return fail(ACFDTD_HIP_EINVAL, "receiver arrays are NULL");
    const uintptr_t align = (uintptr_t)wave_vec(*d) * d->real_bytes;
    if (reinterpret_cast<uintptr_t>(d->p) % align || reinterpret_cast<uintptr_t>(d->p_prev) % align)
        return fail(ACFDTD_HIP_EINVAL, "fields must be aligned to a lane's cells");
    return ACFDTD_HIP_OK;
}

template <typename real, int V>
void launch_wave_step(const WaveArgs<real> &a, bool walls, const real *p, real *q, int64_t step,
                      int64_t column, hipStream_t s) {
    if (walls)
        hipLaunchKernelGGL((wave_step<real, V, true>), dim3(a.t.launched), dim3(a.t.threads), 0, s, a,
                           p, q, step, column);
    else
        hipLaunchKernelGGL((wave_step<real, V, false>), dim3(a.t.launched), dim3(a.t.threads), 0, s,
                           a, p, q, step, column);
}

template <typename real>
int wave_run(const acfdtd_hip_wave_desc &d, int64_t step_offset, int64_t n_steps, int64_t column,
             hipStream_t s) {
    WaveArgs<real> a{};
    a.t = wave_tiling(d);
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.courant2 = (real)d.courant2;
    a.wall_table = static_cast<const real *>(d.wall_table);
    a.n_sources = d.n_sources;
    a.n_receivers = d.n_receivers;
    a.src = (real)d.src_coef;
    a.source_cell = d.source_cell;
    a.source_length = d.source_length;
    a.source_signal = static_cast<const real *>(d.source_signal);
    a.source_stride = d.source_stride;
    a.receiver_cell = d.receiver_cell;
    a.recording = static_cast<real *>(d.recording);
    a.recording_stride = d.recording_stride;
    const bool walls = d.wall_mask != 0;
    real *current = static_cast<real *>(d.p);
    real *previous = static_cast<real *>(d.p_prev);
    for (int64_t n = 0; n < n_steps; ++n) {
        const int64_t step = step_offset + n, col = column + n;
        if (a.t.vec == 4) {
            if constexpr (sizeof(real) == 4)
                launch_wave_step<real, 4>(a, walls, current, previous, step, col, s);
        } else if (a.t.vec == 2) {
            launch_wave_step<real, 2>(a, walls, current, previous, step, col, s);
        } else {
            launch_wave_step<real, 1>(a, walls, current, previous, step, col, s);
        }
        real *newest = previous;
        previous = current;
        current = newest;
    }
    return check_hip(hipGetLastError(), "wave launch");
}

}  // namespace

}  // namespace acfdtd

using namespace acfdtd;

extern "C" size_t acfdtd_hip_wave_struct_size(void) { return sizeof(acfdtd_hip_wave_desc); }

extern "C" int acfdtd_hip_wave_run(const acfdtd_hip_wave_desc *desc, int64_t step_offset,
                                   int64_t n_steps, int64_t recording_column, void *stream) {
    const int status = validate_wave(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + n_steps > desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording is too short for this many steps");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return wave_run<double>(*desc, step_offset, n_steps, recording_column, s);
    return wave_run<float>(*desc, step_offset, n_steps, recording_column, s);
}
