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
#include <type_traits>

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

// Two steps per sweep. A workgroup owns `rows` rows (or a segment of `seg` lane groups of them)
// and also computes p^{n+1} on one halo row each side in y (and one halo lane group each side
// in z for segments), so that the second step at every owned cell finds its neighbours' p^{n+1}
// in LDS. Marching along x, iteration s
//   A. computes p^{n+1} at plane s (the halo planes i0 - 1 and i1 included) from p^n of plane
//      s in LDS (in-plane neighbours; a lane on the block's edge takes its outside neighbour
//      from global memory) and the lane's own p^n of planes s - 1, s, s + 1 in registers;
//   B. publishes it in LDS and issues the loads of p^{n-1} and the outside neighbour of plane
//      s + 2 (used two iterations later);
//   C. computes p^{n+2} at plane s - 1 from that plane's p^{n+1} in LDS and the lane's own
//      p^{n+1} of planes s - 2, s - 1, s;
//   D. publishes its p^n of plane s + 1 in LDS, issues the load of p^n of plane s + 3 (first
//      used as an x neighbour two iterations later); one barrier.
// Two slots for each field and one barrier per plane suffice: a slot written in iteration s
// was last read in iteration s - 1, before the barrier that ends it, and is read next in
// iteration s + 1, after the barrier that ends s. Each p^{n+2} and owned p^{n+1} is written
// once; p^n and p^{n-1} are read once per workgroup (the halo rows and the two halo planes a
// second time, mostly from L2): 4 elements per two cell-updates.
struct Wave2Tiling {
    int vec;
    int groups;  // lane groups per row, nz / vec
    int rows;    // owned rows per workgroup
    int pitch;   // lanes per block row: a whole row, or a segment with a halo lane each side
    int zhalo;   // 0 (whole rows) or 1 (segments)
    int seg;     // owned lane groups per block row
    int nseg;
    int threads;  // (rows + 2) * pitch
    int blocks_per_plane;
    int chunk;
    int nchunk;
    int tiles;
    int launched;
};

template <typename real>
struct Wave2Args {
    Wave2Tiling t;
    WaveArgs<real> w;
};

template <typename real, int V>
__device__ __forceinline__ void add_sources(const WaveArgs<real> &a, Cells<real, V> &out, int i,
                                            int jk0, int64_t step) {
    const real zero = real(0);
    for (int s = 0; s < a.n_sources; ++s) {
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
}

// Receivers read back the lane's own cells just stored at `cells` (picking one of the lane's
// registers by a run-time index would put them in scratch memory).
template <typename real, int V>
__device__ __forceinline__ void record(const WaveArgs<real> &a, const real *cells, int i, int jk0,
                                       int64_t column) {
    for (int r = 0; r < a.n_receivers; ++r) {
        if (a.receiver_cell[2 * r] != i) continue;
        const int at = a.receiver_cell[2 * r + 1] - jk0;
        if (at < 0 || at >= V) continue;
        a.recording[(int64_t)r * a.recording_stride + column] = cells[at];
    }
}

// LDS layout of one plane of a field: lane groups at index tid + pitch (a block row of room
// above and below for the outside rows), then the outside cells of segment edges by block row.
constexpr int kWaveRoom = 768;  // (rows + 4) * pitch <= 512 + 2 * pitch, pitch <= 128
constexpr int kWaveEdges = 256; // rows + 2 <= 170 when segments are used

template <typename real, int V, bool WALLS>
__global__ __launch_bounds__(512) void wave_step2(const Wave2Args<real> args,
                                                  const real *__restrict__ p,
                                                  const real *__restrict__ p_prev,
                                                  real *__restrict__ q2, real *__restrict__ q1,
                                                  const int64_t step, const int64_t column) {
    constexpr int kCells = kWaveRoom * V;
    __shared__ alignas(16) real pn_lds[2][kCells + 2 * kWaveEdges];
    __shared__ alignas(16) real p1_lds[2][kCells];
    __shared__ real walls_lds[WALLS ? 128 : 1];
    const Wave2Tiling &t = args.t;
    const WaveArgs<real> &a = args.w;
    unsigned b = blockIdx.x;
    b = (b & 7u) * ((unsigned)t.launched >> 3) + (b >> 3);
    if (b >= (unsigned)t.tiles) return;  // the whole workgroup leaves together
    const int blk = (int)(b % (unsigned)t.blocks_per_plane);
    const int xc = (int)(b / (unsigned)t.blocks_per_plane);
    const int rb = blk / t.nseg, sg = blk % t.nseg;
    const int tid = threadIdx.x;
    const int br = tid / t.pitch, pos = tid % t.pitch;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int j = rb * t.rows - 1 + br;
    const int kg = sg * t.seg - t.zhalo + pos;
    const bool valid = j >= 0 && j < ny && kg >= 0 && kg < t.groups;
    const bool y_edge = br == 0 || br == t.rows + 1;
    const bool z_edge = pos < t.zhalo || pos >= t.zhalo + t.seg;
    // The block's corner lanes only publish p^n: nobody reads their p^{n+1}.
    const bool computes = valid && !(y_edge && z_edge);
    const bool owner = valid && !y_edge && !z_edge;
    const int k0 = kg * V;
    const int i0 = xc * t.chunk;
    const int i1 = min(i0 + t.chunk, nx);
    const int s0 = i0 > 0 ? i0 - 1 : 0;
    const int64_t plane = (int64_t)ny * nz;
    const int jk0 = j * nz + k0;
    const int64_t row = valid ? (int64_t)jk0 : 0;
    const bool has_ylo = j > 0, has_yhi = j + 1 < ny;
    const bool has_zlo = k0 > 0, has_zhi = k0 + V < nz;
    // A computing lane has at most one neighbour outside the block: a row (y edge) or one cell
    // (z edge of a segment). It loads it and publishes it in LDS with its own p^n.
    const bool out_south = computes && br == 0 && has_ylo;
    const bool out_north = computes && br == t.rows + 1 && has_yhi;
    const bool out_before = computes && pos == 0 && t.zhalo && has_zlo;
    const bool out_after = computes && pos == t.pitch - 1 && t.zhalo && has_zhi;
    // Loaded as the whole neighbouring row or lane group (aligned), the lane's own otherwise.
    const int64_t out_at = out_south   ? row - nz
                           : out_north ? row + nz
                           : out_before ? row - V
                           : out_after  ? row + V
                                        : row;
    // Neighbour addresses in a plane's LDS (in elements). A missing neighbour (a wall) is the
    // cell itself, so its difference is c - c = +0, the wall's zero flux, with no select.
    const int own = (tid + t.pitch) * V;
    const int a_ylo = has_ylo ? own - t.pitch * V : own;
    const int a_yhi = has_yhi ? own + t.pitch * V : own;
    const int a_zlo = !has_zlo ? own : pos > 0 ? own - 1 : kCells + br;
    const int a_zhi = !has_zhi ? own + V - 1 : pos < t.pitch - 1 ? own + V : kCells + kWaveEdges + br;
    const int a_out = out_south ? a_ylo : out_north ? a_yhi : out_before ? a_zlo : a_zhi;
    const int code_y = axis_class(j, ny) << 2;
    const real zero = real(0), lam2 = a.courant2;
    // The wall table in LDS and the lane's factors on planes inside the x walls in registers:
    // a global load in the march would make every wait for it also wait for the loads issued
    // ahead of it (loads complete in order).
    if (WALLS)
        for (int e = tid; e < 128; e += t.threads) walls_lds[e] = a.wall_table[e];
    real f_in[V], keep_in[V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const int code = code_y | (axis_class(k0 + m, nz) << 4);
        f_in[m] = WALLS && valid ? a.wall_table[code] : real(1);
        keep_in[m] = WALLS && valid ? a.wall_table[64 + code] : real(1);
    }

    // One step at plane x for the lane's V cells: `mid` is the field there, `lo` and `hi` at
    // planes x - 1 and x + 1, `lds` the plane in LDS. XIN: x is inside the x walls.
    auto advance = [&](auto xin, const Cells<real, V> &lo, const Cells<real, V> &mid,
                       const Cells<real, V> &hi, const real *lds, const Cells<real, V> &old,
                       int x) {
        constexpr bool XIN = decltype(xin)::value;
        const Cells<real, V> south = *reinterpret_cast<const Cells<real, V> *>(lds + a_ylo);
        const Cells<real, V> north = *reinterpret_cast<const Cells<real, V> *>(lds + a_yhi);
        const real z_before = lds[a_zlo];
        const real z_after = lds[a_zhi];
        Cells<real, V> res;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const real c = mid.c[m];
            const real fx_hi = (XIN || x + 1 < nx) ? hi.c[m] - c : zero;
            const real fx_lo = (XIN || x > 0) ? c - lo.c[m] : zero;
            const real fy_hi = north.c[m] - c;
            const real fy_lo = c - south.c[m];
            const real fz_hi = (m + 1 < V ? mid.c[m + 1] : z_after) - c;
            const real fz_lo = c - (m > 0 ? mid.c[m - 1] : z_before);
            const real lap = ((fx_hi - fx_lo) + (fy_hi - fy_lo)) + (fz_hi - fz_lo);
            if (!WALLS) {
                res.c[m] = (c * real(2) - old.c[m]) + lam2 * lap;
            } else {
                real f = f_in[m], keep = keep_in[m];
                if (!XIN) {
                    const int code = axis_class(x, nx) | code_y | (axis_class(k0 + m, nz) << 4);
                    f = walls_lds[code];
                    keep = walls_lds[64 + code];
                }
                res.c[m] = ((c * real(2) - keep * old.c[m]) + lam2 * lap) * f;
            }
        }
        return res;
    };
    auto publish_pn = [&](real *lds, const Cells<real, V> &mine_pn, const Cells<real, V> &o) {
        if (valid) *reinterpret_cast<Cells<real, V> *>(lds + own) = mine_pn;
        if (out_south || out_north)
            *reinterpret_cast<Cells<real, V> *>(lds + a_out) = o;
        else if (out_before)
            lds[a_out] = o.c[V - 1];
        else if (out_after)
            lds[a_out] = o.c[0];
    };

    auto at_plane = [&](int x) { return (int64_t)min(x, nx - 1) * plane; };

    // Whether a source or receiver lies in this lane's column of cells (any plane).
    bool mine = false;
    if (valid) {
        for (int e = 0; e < a.n_sources; ++e) {
            const int at = a.source_cell[2 * e + 1] - jk0;
            mine = mine || (at >= 0 && at < V);
        }
        for (int e = 0; e < a.n_receivers; ++e) {
            const int at = a.receiver_cell[2 * e + 1] - jk0;
            mine = mine || (at >= 0 && at < V);
        }
    }

    // Register rings indexed by u = s - s0: p^n of plane s0 + u in pn[u mod 4] (planes s - 1 ..
    // s + 2 live in iteration s), p^{n-1} and the outside neighbour of plane s0 + u in
    // old[u mod 2] and out[u mod 2]. The march is unrolled by four so every index is a constant
    // and a load in flight is never copied between registers (a copy would wait for it). Every
    // lane issues every load (clamped to the grid), so the waits count the same on all paths.
    Cells<real, V> pn[4], old[2], out[2];
    pn[3] = load_cells<real, V>(p, at_plane(s0 > 0 ? s0 - 1 : 0) + row);
    pn[0] = load_cells<real, V>(p, at_plane(s0) + row);
    pn[1] = load_cells<real, V>(p, at_plane(s0 + 1) + row);
    pn[2] = load_cells<real, V>(p, at_plane(s0 + 2) + row);
    old[0] = load_cells<real, V>(p_prev, at_plane(s0) + row);
    old[1] = load_cells<real, V>(p_prev, at_plane(s0 + 1) + row);
    out[0] = load_cells<real, V>(p, at_plane(s0) + out_at);
    out[1] = load_cells<real, V>(p, at_plane(s0 + 1) + out_at);
    publish_pn(pn_lds[0], pn[0], out[0]);
    __syncthreads();

    Cells<real, V> p1_lo{}, p1_mid{};  // the lane's p^{n+1} of planes s - 2, s - 1
    for (int u0 = 0;; u0 += 4) {
#pragma unroll
        for (int ph = 0; ph < 4; ++ph) {
            const int s = s0 + u0 + ph;
            if (s > i1) return;  // uniform: the whole workgroup leaves after the same barrier
            const int cur = ph & 1;
            const Cells<real, V> &pn_m1 = pn[(ph + 3) & 3];
            const Cells<real, V> &pn_0 = pn[ph];
            const Cells<real, V> &pn_p1 = pn[(ph + 1) & 3];
            Cells<real, V> p1{};
            if (s < nx && computes) {
                // A: step 1 at plane s.
                p1 = s > 0 && s + 1 < nx
                         ? advance(std::true_type{}, pn_m1, pn_0, pn_p1, pn_lds[cur], old[ph & 1], s)
                         : advance(std::false_type{}, pn_m1, pn_0, pn_p1, pn_lds[cur], old[ph & 1], s);
                // Halo lanes add the sources too: their p^{n+1} feeds the owners' second step.
                if (mine) add_sources<real, V>(a, p1, s, jk0, step);
                if (owner && s >= i0 && s < i1) {
                    real *const cells = q1 + (int64_t)s * plane + row;
                    *reinterpret_cast<Cells<real, V> *>(cells) = p1;
                    if (mine) record<real, V>(a, cells, s, jk0, column);
                }
                // B: publish it.
                *reinterpret_cast<Cells<real, V> *>(p1_lds[cur] + own) = p1;
            }
            // B: p^{n-1} and the outside neighbour of plane s + 2, into the slots just freed.
            old[ph & 1] = load_cells<real, V>(p_prev, at_plane(s + 2) + row);
            const Cells<real, V> out_next = out[(ph + 1) & 1];
            out[ph & 1] = load_cells<real, V>(p, at_plane(s + 2) + out_at);
            const int r = s - 1;
            if (r >= i0 && owner) {
                // C: step 2 at plane r (the x neighbour above is p1 of plane s, absent at nx).
                Cells<real, V> next =
                    r > 0 && s < nx
                        ? advance(std::true_type{}, p1_lo, p1_mid, p1, p1_lds[cur ^ 1], pn_m1, r)
                        : advance(std::false_type{}, p1_lo, p1_mid, p1, p1_lds[cur ^ 1], pn_m1, r);
                if (mine) add_sources<real, V>(a, next, r, jk0, step + 1);
                real *const cells = q2 + (int64_t)r * plane + row;
                *reinterpret_cast<Cells<real, V> *>(cells) = next;
                if (mine) record<real, V>(a, cells, r, jk0, column + 1);
            }
            // D: p^n (and the outside neighbour) of plane s + 1 for the next step 1; p^n of
            // plane s + 3 into the slot of plane s - 1, whose last use was step 2 above.
            if (s + 1 < nx) publish_pn(pn_lds[cur ^ 1], pn_p1, out_next);
            pn[(ph + 3) & 3] = load_cells<real, V>(p, at_plane(s + 3) + row);
            __syncthreads();
            p1_lo = p1_mid;
            p1_mid = p1;
        }
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

Wave2Tiling wave2_tiling(const acfdtd_hip_wave_desc &d) {
    Wave2Tiling t{};
    t.vec = wave_vec(d);
    t.groups = d.nz / t.vec;
    const int want = d.tile_j > 0 ? 2 * d.tile_j : 6;
    if ((want + 2) * t.groups <= 512) {
        // Whole rows; by default as many as fit 512 lanes (no more than the grid has).
        t.pitch = t.seg = t.groups;
        t.zhalo = 0;
        t.nseg = 1;
        t.rows = want;
        if (d.tile_j <= 0) {
            const int fit = 512 / t.groups - 2;
            if (fit > t.rows) t.rows = fit;
            if (t.rows > d.ny) t.rows = d.ny;
        }
    } else {
        // Row segments of equal width with one halo lane group each side.
        const int widest = 512 / (want + 2) - 2;
        t.nseg = ceil_div(t.groups, widest);
        t.seg = ceil_div(t.groups, t.nseg);
        t.pitch = t.seg + 2;
        t.zhalo = 1;
        t.rows = d.tile_j > 0 ? want : 512 / t.pitch - 2;
    }
    t.threads = (t.rows + 2) * t.pitch;
    t.blocks_per_plane = ceil_div(d.ny, t.rows) * t.nseg;
    int chunk = d.chunk_x;
    if (chunk <= 0) {
        const int wanted = ceil_div(2048, t.blocks_per_plane);
        chunk = ceil_div(d.nx, wanted > 0 ? wanted : 1);
        if (chunk < 8) chunk = 8;
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
    if (d->time_block < 0 || d->time_block > 2) return fail(ACFDTD_HIP_EINVAL, "time_block must be 0, 1 or 2");
    if (d->time_block == 2) {
        const void *f[4] = {d->p, d->p_prev, d->p_swap, d->p_prev_swap};
        if (!f[2] || !f[3]) return fail(ACFDTD_HIP_EINVAL, "time_block 2 needs both swap buffers");
        for (int x = 0; x < 4; ++x)
            for (int y = x + 1; y < 4; ++y)
                if (f[x] == f[y])
                    return fail(ACFDTD_HIP_EINVAL, "p, p_prev and the swap buffers must be distinct");
        if (reinterpret_cast<uintptr_t>(f[2]) % align || reinterpret_cast<uintptr_t>(f[3]) % align)
            return fail(ACFDTD_HIP_EINVAL, "swap buffers must be aligned to a lane's cells");
    }
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

template <typename real, int V>
void launch_wave_step2(const Wave2Args<real> &a, bool walls, const real *p, const real *p_prev,
                       real *q2, real *q1, int64_t step, int64_t column, hipStream_t s) {
    if (walls)
        hipLaunchKernelGGL((wave_step2<real, V, true>), dim3(a.t.launched), dim3(a.t.threads), 0, s,
                           a, p, p_prev, q2, q1, step, column);
    else
        hipLaunchKernelGGL((wave_step2<real, V, false>), dim3(a.t.launched), dim3(a.t.threads), 0,
                           s, a, p, p_prev, q2, q1, step, column);
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
    int64_t n = 0;
    if (d.time_block == 2) {
        // Pairs of two-step sweeps, (p, p_prev) -> (p_swap, p_prev_swap) -> (p, p_prev): after
        // each four steps the newest field is back in p, as after four single steps.
        Wave2Args<real> a2{};
        a2.t = wave2_tiling(d);
        a2.w = a;
        real *swap[2] = {static_cast<real *>(d.p_swap), static_cast<real *>(d.p_prev_swap)};
        real *home[2] = {current, previous};
        const int64_t paired = n_steps / 4 * 4;
        for (; n < paired; n += 2) {
            real *const *from = (n % 4 == 0) ? home : swap;
            real *const *to = (n % 4 == 0) ? swap : home;
            const int64_t step = step_offset + n, col = column + n;
            if (a2.t.vec == 4) {
                if constexpr (sizeof(real) == 4)
                    launch_wave_step2<real, 4>(a2, walls, from[0], from[1], to[0], to[1], step, col, s);
            } else if (a2.t.vec == 2) {
                launch_wave_step2<real, 2>(a2, walls, from[0], from[1], to[0], to[1], step, col, s);
            } else {
                launch_wave_step2<real, 1>(a2, walls, from[0], from[1], to[0], to[1], step, col, s);
            }
        }
    }
    for (; n < n_steps; ++n) {
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
