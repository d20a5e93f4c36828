// The fourth-order-in-space staggered scheme on gfx950: a velocity sweep and a pressure sweep
// per step, the reference C loop's two-sweep structure (acfdtd.c:110-178) with the spatial
// differences widened to the standard fourth-order staggered stencil
//
//     D u = c1 (u[i] - u[i-1]) - c2 (u[i+1] - u[i-2]),    c1 = 9/8, c2 = 1/24,
//
// which the reference lists as an option and does not implement (PLAN.md:17, :116). Rigid
// walls by images: the pressure is extended evenly and the normal velocity oddly past each wall
// face (p[-1] = p[0], v[-1] = -v[1], and likewise at the far walls), which keeps the discrete
// energy conserved and the rigid box's cosine modes exact. Sponge, air absorption, sources and
// receivers are the first-order sweeps' (acfdtd_hip_desc). The CPU oracle restates the same
// arithmetic, which these kernels reproduce bit for bit (built with -ffp-contract=off).
//
// Mapping to CDNA4: each lane holds V consecutive z cells (V = 4 in fp32, 2 in fp64, when nz
// allows; else one), so every field access is a 16-byte vector access, and lanes are numbered
// across the flattened (y, z / V) plane as in the fused first-order step (no idle lanes
// whatever nz is). A workgroup of 64 x tile_j lanes marches a chunk of planes along x with the
// x stencil (four planes) in a register queue, so each field crosses HBM once per sweep; the y
// stencil points (two rows or faces each side) are re-read through L1/L2, and of the z stencil
// points only the two or three past the lane's own cells are loaded (one word each). Bytes per
// cell-update are those of the two-sweep first-order step, 12 + 2R elements: the wider stencil
// costs cache traffic and arithmetic, not HBM traffic. The one-cell kernels (64 lanes along z
// times tile_j rows) take the grids whose nz no wider lane divides.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {

namespace {

// Cell index mirrored into [0, n) across the wall faces (clamped for n = 1, where no interior
// face reads it).
__device__ __forceinline__ int cell_image(int i, int n) {
    if (i < 0) i = -1 - i;
    if (i >= n) i = 2 * n - 1 - i;
    return min(max(i, 0), n - 1);
}

// Face index mirrored into [0, n] and the sign of the odd extension.
__device__ __forceinline__ int face_image(int g, int n, bool &negate) {
    negate = g < 0 || g > n;
    if (g < 0) g = -g;
    if (g > n) g = 2 * n - g;
    return g;
}

template <typename real>
__device__ __forceinline__ real signed_value(real v, bool negate) {
    return v * (negate ? real(-1) : real(1));
}

template <typename real>
__device__ __forceinline__ real stencil(real c1, real c2, real lo2, real lo, real hi, real hi2) {
    return c1 * (hi - lo) - c2 * (hi2 - lo2);
}

template <typename real, bool LAYER>
__global__ __launch_bounds__(512) void velocity_sweep4(const SweepArgs<real> a, const real c1,
                                                       const real c2) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, nx);
    const Layout L(nx, ny, nz, j, k);
    const real *__restrict__ p = a.p;
    real *__restrict__ vx = a.vx;
    real *__restrict__ vy = a.vy;
    real *__restrict__ vz = a.vz;
    // y and z stencil points of the lower face of cell (., j, k), as offsets from the cell.
    const int64_t y_m2 = (int64_t)(cell_image(j - 2, ny) - j) * nz;
    const int64_t y_p1 = (int64_t)(cell_image(j + 1, ny) - j) * nz;
    const int64_t z_m2 = cell_image(k - 2, nz) - k;
    const int64_t z_p1 = cell_image(k + 1, nz) - k;
    const real vel = a.vel;
    real fyc = 1, fzc = 1, fyf = 1, fzf = 1;
    if (LAYER) {
        fyc = a.cell[1][j];
        fzc = a.cell[2][k];
        fyf = a.face[1][j];
        fzf = a.face[2][k];
    }
    real p_m2 = p[L.cell(cell_image(i0 - 2, nx))];
    real p_m1 = p[L.cell(cell_image(i0 - 1, nx))];
    real p_c = p[L.cell(i0)];
    real p_p1 = p[L.cell(cell_image(i0 + 1, nx))];
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real p_p2 = p[L.cell(cell_image(i + 2, nx))];
        real fxc = 1, fxf = 1;
        if (LAYER) {
            fxc = a.cell[0][i];
            fxf = a.face[0][i];
        }
        if (LAYER || i > 0) {
            real v = vx[c];
            if (i > 0) v = v - vel * stencil(c1, c2, p_m2, p_m1, p_c, p_p1);
            if (LAYER) v = ((v * fxf) * fyc) * fzc;
            vx[c] = v;
        }
        const int64_t cy = L.face_y(i);
        if (LAYER || j > 0) {
            real v = vy[cy];
            if (j > 0) v = v - vel * stencil(c1, c2, p[c + y_m2], p[c - nz], p_c, p[c + y_p1]);
            if (LAYER) v = ((v * fxc) * fyf) * fzc;
            vy[cy] = v;
        }
        const int64_t cz = L.face_z(i);
        if (LAYER || k > 0) {
            real v = vz[cz];
            if (k > 0) v = v - vel * stencil(c1, c2, p[c + z_m2], p[c - 1], p_c, p[c + z_p1]);
            if (LAYER) v = ((v * fxc) * fyc) * fzf;
            vz[cz] = v;
        }
        if (LAYER) {
            // The far wall faces, never updated, only damped (as the second-order sweep).
            if (i == nx - 1) {
                const int64_t t = c + L.plane;
                vx[t] = ((vx[t] * a.face[0][nx]) * fyc) * fzc;
            }
            if (j == ny - 1) {
                const int64_t t = cy + nz;
                vy[t] = ((vy[t] * fxc) * a.face[1][ny]) * fzc;
            }
            if (k == nz - 1) {
                const int64_t t = cz + 1;
                vz[t] = ((vz[t] * fxc) * fyc) * a.face[2][nz];
            }
        }
        p_m2 = p_m1;
        p_m1 = p_c;
        p_c = p_p1;
        p_p1 = p_p2;
    }
}

template <typename real, int NR, bool LAYER>
__global__ __launch_bounds__(512) void pressure_sweep4(const SweepArgs<real> a, const real c1,
                                                       const real c2, const int64_t step,
                                                       const int64_t column) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, nx);
    const Layout L(nx, ny, nz, j, k);
    const int64_t n_cells = L.plane * nx;
    const int jk = j * nz + k;
    real *__restrict__ p = a.p;
    real *__restrict__ psi = a.psi;
    const real *__restrict__ vx = a.vx;
    const real *__restrict__ vy = a.vy;
    const real *__restrict__ vz = a.vz;
    // y faces j - 1 and j + 2, z faces k - 1 and k + 2, mirrored, as offsets from the cell's
    // lower face, with the signs of the odd extension.
    bool ny_m1, ny_p2, nz_m1, nz_p2;
    const int64_t y_m1 = (int64_t)(face_image(j - 1, ny, ny_m1) - j) * nz;
    const int64_t y_p2 = (int64_t)(face_image(j + 2, ny, ny_p2) - j) * nz;
    const int64_t z_m1 = face_image(k - 1, nz, nz_m1) - k;
    const int64_t z_p2 = face_image(k + 2, nz, nz_p2) - k;
    real fyc = 1, fzc = 1;
    if (LAYER) {
        fyc = a.cell[1][j];
        fzc = a.cell[2][k];
    }
    // x faces i - 1, i, i + 1, i + 2 of the march, mirrored and signed.
    bool neg;
    int g = face_image(i0 - 1, nx, neg);
    real x_m1 = signed_value(vx[L.cell(g)], neg);
    real x_0 = vx[L.cell(i0)];
    real x_1 = vx[L.cell(i0 + 1)];
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        g = face_image(i + 2, nx, neg);
        const real x_2 = signed_value(vx[L.cell(g)], neg);
        const int64_t cy = L.face_y(i);
        const int64_t cz = L.face_z(i);
        const real dx = stencil(c1, c2, x_m1, x_0, x_1, x_2);
        const real dy = stencil(c1, c2, signed_value(vy[cy + y_m1], ny_m1), vy[cy], vy[cy + nz],
                                signed_value(vy[cy + y_p2], ny_p2));
        const real dz = stencil(c1, c2, signed_value(vz[cz + z_m1], nz_m1), vz[cz], vz[cz + 1],
                                signed_value(vz[cz + z_p2], nz_p2));
        const real div = (dx + dy) + dz;
        real pn = p[c];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            real *state = psi + (int64_t)r * n_cells + c;
            const real target = div * a.gain[r];
            const real offset = *state - target;
            *state = offset * a.decay[r] + target;
            pn = pn + offset * a.mean[r];
        }
        if (NR > 0) pn = pn + div * a.relax_scale;
        pn = pn - div * a.pres;
        if (LAYER) pn = ((pn * a.cell[0][i]) * fyc) * fzc;
        for (int s = 0; s < a.n_sources; ++s) {
            if (a.source_cell[2 * s] == i && a.source_cell[2 * s + 1] == jk &&
                step < a.source_length[s])
                pn = pn + a.src * a.source_signal[s * a.source_stride + step];
        }
        p[c] = pn;
        for (int r = 0; r < a.n_receivers; ++r) {
            if (a.receiver_cell[2 * r] == i && a.receiver_cell[2 * r + 1] == jk)
                a.recording[r * a.recording_stride + column] = pn;
        }
        x_m1 = x_0;
        x_0 = x_1;
        x_1 = x_2;
    }
}

// The velocity sweep on lanes of V z cells (vector accesses): the arithmetic of
// velocity_sweep4, cell for cell.
template <typename real, int V, bool LAYER>
__global__ __launch_bounds__(512) void velocity_sweep4_vec(const SweepArgs<real> a, const real c1,
                                                           const real c2) {
    int kb, jb, xc, j, k0;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    if (!flat_lane(a.tiles, a.ny, a.nz, jb, j, k0)) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, nx);
    const Layout L(nx, ny, nz, j, k0);
    const real *__restrict__ p = a.p;
    real *__restrict__ vx = a.vx;
    real *__restrict__ vy = a.vy;
    real *__restrict__ vz = a.vz;
    // y stencil points of the lower y face of the lane's cells, as offsets from the cells.
    const int64_t y_m2 = (int64_t)(cell_image(j - 2, ny) - j) * nz;
    const int64_t y_p1 = (int64_t)(cell_image(j + 1, ny) - j) * nz;
    // z stencil points past the lane's own cells: cells k0 - 2, k0 - 1 and k0 + V, mirrored.
    const int64_t z_m2 = cell_image(k0 - 2, nz) - k0;
    const int64_t z_m1 = cell_image(k0 - 1, nz) - k0;
    const int64_t z_pv = cell_image(k0 + V, nz) - k0;
    const real vel = a.vel;
    real fyc = 1, fyf = 1, fzc[V], fzf[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        fzc[e] = 1;
        fzf[e] = 1;
    }
    if (LAYER) {
        fyc = a.cell[1][j];
        fyf = a.face[1][j];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            fzc[e] = a.cell[2][k0 + e];
            fzf[e] = a.face[2][k0 + e];
        }
    }
    real p_m2[V], p_m1[V], p_c[V], p_p1[V];
    vload<real, V>(p + L.cell(cell_image(i0 - 2, nx)), p_m2);
    vload<real, V>(p + L.cell(cell_image(i0 - 1, nx)), p_m1);
    vload<real, V>(p + L.cell(i0), p_c);
    vload<real, V>(p + L.cell(cell_image(i0 + 1, nx)), p_p1);
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        real p_p2[V];
        vload<real, V>(p + L.cell(cell_image(i + 2, nx)), p_p2);
        real fxc = 1, fxf = 1;
        if (LAYER) {
            fxc = uniform_load(a.cell[0] + i);
            fxf = uniform_load(a.face[0] + i);
        }
        if (LAYER || i > 0) {
            real v[V];
            vload<real, V>(vx + c, v);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                real w = v[e];
                if (i > 0) w = w - vel * stencil(c1, c2, p_m2[e], p_m1[e], p_c[e], p_p1[e]);
                if (LAYER) w = ((w * fxf) * fyc) * fzc[e];
                v[e] = w;
            }
            vstore<real, V>(vx + c, v);
        }
        const int64_t cy = L.face_y(i);
        if (LAYER || j > 0) {
            real v[V], q_m2[V], q_m1[V], q_p1[V];
            vload<real, V>(vy + cy, v);
            if (j > 0) {
                vload<real, V>(p + c + y_m2, q_m2);
                vload<real, V>(p + c - nz, q_m1);
                vload<real, V>(p + c + y_p1, q_p1);
            }
#pragma unroll
            for (int e = 0; e < V; ++e) {
                real w = v[e];
                if (j > 0) w = w - vel * stencil(c1, c2, q_m2[e], q_m1[e], p_c[e], q_p1[e]);
                if (LAYER) w = ((w * fxc) * fyf) * fzc[e];
                v[e] = w;
            }
            vstore<real, V>(vy + cy, v);
        }
        const int64_t cz = L.face_z(i);
        {
            // p of cells k0 - 2 .. k0 + V; face k0 + e takes points e .. e + 3.
            real zc[V + 3];
            zc[0] = p[c + z_m2];
            zc[1] = p[c + z_m1];
#pragma unroll
            for (int e = 0; e < V; ++e) zc[2 + e] = p_c[e];
            zc[V + 2] = p[c + z_pv];
            real v[V];
            vload<real, V>(vz + cz, v);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                real w = v[e];
                if (k0 + e > 0) w = w - vel * stencil(c1, c2, zc[e], zc[e + 1], zc[e + 2], zc[e + 3]);
                if (LAYER) w = ((w * fxc) * fyc) * fzf[e];
                v[e] = w;
            }
            vstore<real, V>(vz + cz, v);
        }
        if (LAYER) {
            // The far wall faces, never updated, only damped (as the second-order sweep).
            if (i == nx - 1) {
                const real f = uniform_load(a.face[0] + nx);
                real v[V];
                vload<real, V>(vx + c + L.plane, v);
#pragma unroll
                for (int e = 0; e < V; ++e) v[e] = ((v[e] * f) * fyc) * fzc[e];
                vstore<real, V>(vx + c + L.plane, v);
            }
            if (j == ny - 1) {
                const real f = uniform_load(a.face[1] + ny);
                real v[V];
                vload<real, V>(vy + cy + nz, v);
#pragma unroll
                for (int e = 0; e < V; ++e) v[e] = ((v[e] * fxc) * f) * fzc[e];
                vstore<real, V>(vy + cy + nz, v);
            }
            if (k0 + V == nz) vz[cz + V] = ((vz[cz + V] * fxc) * fyc) * uniform_load(a.face[2] + nz);
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            p_m2[e] = p_m1[e];
            p_m1[e] = p_c[e];
            p_c[e] = p_p1[e];
            p_p1[e] = p_p2[e];
        }
    }
}

// The pressure sweep on lanes of V z cells: the arithmetic of pressure_sweep4, cell for cell.
template <typename real, int V, int NR, bool LAYER>
__global__ __launch_bounds__(512) void pressure_sweep4_vec(const SweepArgs<real> a, const real c1,
                                                           const real c2, const int64_t step,
                                                           const int64_t column) {
    int kb, jb, xc, j, k0;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    if (!flat_lane(a.tiles, a.ny, a.nz, jb, j, k0)) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, nx);
    const Layout L(nx, ny, nz, j, k0);
    const int64_t n_cells = L.plane * nx;
    const int jk0 = j * nz + k0;
    real *__restrict__ p = a.p;
    real *__restrict__ psi = a.psi;
    const real *__restrict__ vx = a.vx;
    const real *__restrict__ vy = a.vy;
    const real *__restrict__ vz = a.vz;
    // y faces j - 1 and j + 2, mirrored, as offsets from the cells' lower face, with the signs
    // of the odd extension; z faces k0 - 1 and k0 + V + 1 likewise (faces k0 .. k0 + V are
    // the lane's own and its top face, all inside the room).
    bool ny_m1, ny_p2, nz_m1, nz_pv;
    const int64_t y_m1 = (int64_t)(face_image(j - 1, ny, ny_m1) - j) * nz;
    const int64_t y_p2 = (int64_t)(face_image(j + 2, ny, ny_p2) - j) * nz;
    const int64_t z_m1 = face_image(k0 - 1, nz, nz_m1) - k0;
    const int64_t z_pv = face_image(k0 + V + 1, nz, nz_pv) - k0;
    real fyc = 1, fzc[V];
#pragma unroll
    for (int e = 0; e < V; ++e) fzc[e] = 1;
    if (LAYER) {
        fyc = a.cell[1][j];
#pragma unroll
        for (int e = 0; e < V; ++e) fzc[e] = a.cell[2][k0 + e];
    }
    // x faces i - 1, i, i + 1, i + 2 of the march, mirrored and signed.
    bool neg;
    int g = face_image(i0 - 1, nx, neg);
    real x_m1[V], x_0[V], x_1[V];
    vload<real, V>(vx + L.cell(g), x_m1);
#pragma unroll
    for (int e = 0; e < V; ++e) x_m1[e] = signed_value(x_m1[e], neg);
    vload<real, V>(vx + L.cell(i0), x_0);
    vload<real, V>(vx + L.cell(i0 + 1), x_1);
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        g = face_image(i + 2, nx, neg);
        real x_2[V];
        vload<real, V>(vx + L.cell(g), x_2);
#pragma unroll
        for (int e = 0; e < V; ++e) x_2[e] = signed_value(x_2[e], neg);
        const int64_t cy = L.face_y(i);
        const int64_t cz = L.face_z(i);
        real y_lo[V], y_0[V], y_1[V], y_hi[V];
        vload<real, V>(vy + cy + y_m1, y_lo);
        vload<real, V>(vy + cy, y_0);
        vload<real, V>(vy + cy + nz, y_1);
        vload<real, V>(vy + cy + y_p2, y_hi);
        // vz of faces k0 - 1 .. k0 + V + 1; cell k0 + e takes points e .. e + 3.
        real zf[V + 3];
        zf[0] = signed_value(vz[cz + z_m1], nz_m1);
        {
            real w[V];
            vload<real, V>(vz + cz, w);
#pragma unroll
            for (int e = 0; e < V; ++e) zf[1 + e] = w[e];
        }
        zf[V + 1] = vz[cz + V];
        zf[V + 2] = signed_value(vz[cz + z_pv], nz_pv);
        real pn[V];
        vload<real, V>(p + c, pn);
        real states[NR > 0 ? NR : 1][V];
#pragma unroll
        for (int r = 0; r < NR; ++r) vload<real, V>(psi + (int64_t)r * n_cells + c, states[r]);
        const real fxc = LAYER ? uniform_load(a.cell[0] + i) : real(1);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real dx = stencil(c1, c2, x_m1[e], x_0[e], x_1[e], x_2[e]);
            const real dy = stencil(c1, c2, signed_value(y_lo[e], ny_m1), y_0[e], y_1[e],
                                    signed_value(y_hi[e], ny_p2));
            const real dz = stencil(c1, c2, zf[e], zf[e + 1], zf[e + 2], zf[e + 3]);
            const real div = (dx + dy) + dz;
            real pe = pn[e];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const real target = div * a.gain[r];
                const real offset = states[r][e] - target;
                states[r][e] = offset * a.decay[r] + target;
                pe = pe + offset * a.mean[r];
            }
            if (NR > 0) pe = pe + div * a.relax_scale;
            pe = pe - div * a.pres;
            if (LAYER) pe = ((pe * fxc) * fyc) * fzc[e];
            pn[e] = pe;
        }
        for (int s = 0; s < a.n_sources; ++s) {
            const int e = a.source_cell[2 * s + 1] - jk0;
            if (a.source_cell[2 * s] == i && e >= 0 && e < V && step < a.source_length[s])
                lane_add<real, V>(pn, e, a.src * a.source_signal[s * a.source_stride + step]);
        }
        vstore<real, V>(p + c, pn);
#pragma unroll
        for (int r = 0; r < NR; ++r) vstore<real, V>(psi + (int64_t)r * n_cells + c, states[r]);
        for (int r = 0; r < a.n_receivers; ++r) {
            const int e = a.receiver_cell[2 * r + 1] - jk0;
            if (a.receiver_cell[2 * r] == i && e >= 0 && e < V)
                a.recording[r * a.recording_stride + column] = lane_get<real, V>(pn, e);
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            x_m1[e] = x_0[e];
            x_0[e] = x_1[e];
            x_1[e] = x_2[e];
        }
    }
}

template <typename real>
using Velocity4 = void (*)(const SweepArgs<real>, const real, const real);
template <typename real>
using Pressure4 = void (*)(const SweepArgs<real>, const real, const real, const int64_t,
                           const int64_t);

template <typename real, int NR>
Pressure4<real> pick_pressure4_nr(bool layer) {
    return layer ? pressure_sweep4<real, NR, true> : pressure_sweep4<real, NR, false>;
}

template <typename real>
Pressure4<real> pick_pressure4(int nr, bool layer) {
    switch (nr) {
        case 0: return pick_pressure4_nr<real, 0>(layer);
        case 1: return pick_pressure4_nr<real, 1>(layer);
        case 2: return pick_pressure4_nr<real, 2>(layer);
        case 3: return pick_pressure4_nr<real, 3>(layer);
        default: return pick_pressure4_nr<real, 4>(layer);
    }
}

template <typename real, int V, int NR>
Pressure4<real> pick_pressure4_vec_nr(bool layer) {
    return layer ? pressure_sweep4_vec<real, V, NR, true> : pressure_sweep4_vec<real, V, NR, false>;
}

template <typename real, int V>
Pressure4<real> pick_pressure4_vec(int nr, bool layer) {
    switch (nr) {
        case 0: return pick_pressure4_vec_nr<real, V, 0>(layer);
        case 1: return pick_pressure4_vec_nr<real, V, 1>(layer);
        case 2: return pick_pressure4_vec_nr<real, V, 2>(layer);
        case 3: return pick_pressure4_vec_nr<real, V, 3>(layer);
        default: return pick_pressure4_vec_nr<real, V, 4>(layer);
    }
}

// The kernels of a lane width (4 exists for fp32 only: 16 bytes).
template <typename real>
void pick_vec(int vec, int nr, bool layer, Velocity4<real> &vk, Pressure4<real> &pk) {
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) {
            vk = layer ? velocity_sweep4_vec<real, 4, true> : velocity_sweep4_vec<real, 4, false>;
            pk = pick_pressure4_vec<real, 4>(nr, layer);
            return;
        }
    }
    vk = layer ? velocity_sweep4_vec<real, 2, true> : velocity_sweep4_vec<real, 2, false>;
    pk = pick_pressure4_vec<real, 2>(nr, layer);
}

// Tiling of the V-wide sweeps: the fused step's flattened (y, z / V) plane, by default in
// workgroups of 8 wavefronts marching 16 planes (the x stencil re-reads three planes of p, or
// of vx, at each march end; 256^3 fp32 with sponge and air measured 67 Gcell/s, against 63 to
// 66 for 1 to 4 wavefronts or 8 planes, and 51 for 32 planes).
inline Tiling fourth_vec_tiling(const acfdtd_hip_desc &d, int vec) {
    acfdtd_hip_desc e = d;
    if (e.tile_j <= 0) e.tile_j = 8;
    if (e.chunk_x <= 0) e.chunk_x = 16;
    return flat_tiling(e, vec);
}

template <typename real>
int fourth_run(const acfdtd_hip_desc &d, double c1, double c2, int64_t step_offset,
               int64_t n_steps, int64_t column, hipStream_t s) {
    SweepArgs<real> a;
    fill_args(d, a);
    Velocity4<real> vk = d.has_layer ? velocity_sweep4<real, true> : velocity_sweep4<real, false>;
    Pressure4<real> pk = pick_pressure4<real>(d.n_relax, d.has_layer != 0);
    const int vec = fused_vec(d);
    if (vec > 1) {
        a.tiles = fourth_vec_tiling(d, vec);
        pick_vec<real>(vec, d.n_relax, d.has_layer != 0, vk, pk);
    }
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    for (int64_t n = 0; n < n_steps; ++n) {
        hipLaunchKernelGGL(vk, grid, block, 0, s, a, (real)c1, (real)c2);
        hipLaunchKernelGGL(pk, grid, block, 0, s, a, (real)c1, (real)c2, step_offset + n,
                           column + n);
    }
    return check_hip(hipGetLastError(), "fourth-order launch");
}

}  // namespace

}  // namespace acfdtd

using namespace acfdtd;

extern "C" int acfdtd_hip_fourth_run(const acfdtd_hip_desc *desc, double c1, double c2,
                                     int64_t step_offset, int64_t n_steps,
                                     int64_t recording_column, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (desc->wall_mask != 0)
        return fail(ACFDTD_HIP_EINVAL, "the fourth-order sweeps take rigid walls only (images)");
    if (desc->flags != 0)
        return fail(ACFDTD_HIP_EINVAL, "the fourth-order sweeps take no flags");
    if (desc->wall_tape || desc->div_tape || desc->layer_tape)
        return fail(ACFDTD_HIP_EINVAL, "the fourth-order sweeps write no tape");
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + n_steps > desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording is too short for this many steps");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8)
        return fourth_run<double>(*desc, c1, c2, step_offset, n_steps, recording_column, s);
    return fourth_run<float>(*desc, c1, c2, step_offset, n_steps, recording_column, s);
}
