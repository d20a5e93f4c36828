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
// Mapping to CDNA4: as the first-order two sweeps, a workgroup is 64 lanes along z (one
// wavefront per row, unit-stride and coalesced) times tile_j rows along y, marching a chunk of
// planes along x with the x stencil (four planes) in a register queue, so each field crosses
// HBM once per sweep; the y and z stencil points (two cells or faces each side) are re-read
// through L1/L2. Bytes per cell-update are those of the two-sweep first-order step, 12 + 2R
// elements: the wider stencil costs cache traffic and arithmetic, not HBM traffic.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
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

template <typename real>
int fourth_run(const acfdtd_hip_desc &d, double c1, double c2, int64_t step_offset,
               int64_t n_steps, int64_t column, hipStream_t s) {
    SweepArgs<real> a;
    fill_args(d, a);
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    const Velocity4<real> vk = d.has_layer ? velocity_sweep4<real, true> : velocity_sweep4<real, false>;
    const Pressure4<real> pk = pick_pressure4<real>(d.n_relax, d.has_layer != 0);
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
    if (desc->wall_tape || desc->div_tape)
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
