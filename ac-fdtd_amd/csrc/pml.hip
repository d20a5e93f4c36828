// The split-field perfectly matched layer on gfx950: a velocity sweep and a pressure sweep per
// step (include/acfdtd_hip.h, acfdtd_hip_pml_desc, states the arithmetic; the CPU oracle
// restates it and these kernels reproduce it bit for bit, built with -ffp-contract=off).
//
// Mapping, two ways with the same arithmetic term by term (so the same bits):
// * the fused step (pml_fused_step, when the descriptor has the swap set): one sweep per step
//   from one set of fields to the other, in pairs A -> B -> A, V z cells per lane (4 fp32, 2
//   fp64 when nz allows: 16-byte accesses), lanes flattened over the (y, z / V) plane, a chunk
//   of planes marched along x with the x neighbour in registers. Each cell computes the new
//   velocity on its six faces from the unmodified inputs (a face shared with a neighbour gets
//   the same bits in both) and its owner stores it, as in the first-order fused step;
// * the two in-place sweeps (pml_velocity, pml_pressure; an odd last step, or no swap set):
//   the first-order two sweeps' structure (acfdtd.c:110-178), 64 lanes along z per wavefront.
// Outside the layer a cell does the reference's update and touches no split part; in the layer
// it reads and writes p_y and p_z too (updated in place, like the relaxation states of the
// fused step). So the split parts cost HBM traffic only in the layer, and the layer's per-axis
// coefficients are 1D arrays read through the scalar and L1 caches.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {

namespace {

template <typename real>
struct PMLArgs {
    const real *cell_a[3], *cell_b[3];
    const real *face_a[3], *face_b[3];
    real *p_y, *p_z;
};

template <typename real>
__global__ __launch_bounds__(512) void pml_velocity(const SweepArgs<real> a, const PMLArgs<real> m) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    const Layout L(a.nx, a.ny, a.nz, j, k);
    const real *__restrict__ p = a.p;
    real *__restrict__ vx = a.vx;
    real *__restrict__ vy = a.vy;
    real *__restrict__ vz = a.vz;
    const real vel = a.vel;
    const real ay = m.face_a[1][j], by = m.face_b[1][j];
    const real az = m.face_a[2][k], bz = m.face_b[2][k];
    real p_prev = i0 > 0 ? p[L.cell(i0 - 1)] : real(0);
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real pc = p[c];
        if (i > 0) vx[c] = m.face_a[0][i] * vx[c] - m.face_b[0][i] * (vel * (pc - p_prev));
        if (j > 0) {
            const int64_t cy = L.face_y(i);
            vy[cy] = ay * vy[cy] - by * (vel * (pc - p[c - a.nz]));
        }
        if (k > 0) {
            const int64_t cz = L.face_z(i);
            vz[cz] = az * vz[cz] - bz * (vel * (pc - p[c - 1]));
        }
        p_prev = pc;
    }
}

template <typename real>
__global__ __launch_bounds__(512) void pml_pressure(const SweepArgs<real> a, const PMLArgs<real> m,
                                                    const int64_t step, const int64_t column) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    const Layout L(a.nx, a.ny, a.nz, j, k);
    const int jk = j * a.nz + k;
    real *__restrict__ p = a.p;
    const real *__restrict__ vx = a.vx;
    const real *__restrict__ vy = a.vy;
    const real *__restrict__ vz = a.vz;
    const real pres = a.pres, one = real(1);
    const real ay = m.cell_a[1][j], by = m.cell_b[1][j];
    const real az = m.cell_a[2][k], bz = m.cell_b[2][k];
    const bool layer_yz = ay != one || az != one;
    real vx_lo = vx[L.cell(i0)];
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real vx_hi = vx[c + L.plane];
        const int64_t cy = L.face_y(i);
        const int64_t cz = L.face_z(i);
        const real dx = vx_hi - vx_lo;
        const real dy = vy[cy + a.nz] - vy[cy];
        const real dz = vz[cz + 1] - vz[cz];
        vx_lo = vx_hi;
        const real ax = m.cell_a[0][i];
        real pn;
        if (layer_yz || ax != one) {
            const real bx = m.cell_b[0][i];
            const real py = m.p_y[c], pz = m.p_z[c];
            const real px = (p[c] - py) - pz;
            const real nx_ = ax * px - bx * (pres * dx);
            const real ny_ = ay * py - by * (pres * dy);
            const real nz_ = az * pz - bz * (pres * dz);
            m.p_y[c] = ny_;
            m.p_z[c] = nz_;
            pn = (nx_ + ny_) + nz_;
        } else {
            pn = p[c] - ((dx + dy) + dz) * pres;
        }
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
    }
}

// One fused PML step of the lane's V cells over planes [i0, i1): pml_velocity's face updates
// on all six faces of each cell (the owner stores each face: the lower x, y and z faces, and the
// upper ones on the far walls), then pml_pressure's update, sources and receivers.
template <typename real, int V>
__device__ __forceinline__ void pml_march(const SweepArgs<real> &a, const PMLArgs<real> &m,
                                          const FusedIO<real> &io, const int i0, const int i1,
                                          const int j, const int k0, const int64_t step,
                                          const int64_t column) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const Layout L(nx, ny, nz, j, k0);
    const int jk0 = j * nz + k0;
    const real vel = a.vel, pres = a.pres, one = real(1);
    const real *__restrict__ p_in = io.p_in;
    const real *__restrict__ vx_in = io.vx_in;
    const real *__restrict__ vy_in = io.vy_in;
    const real *__restrict__ vz_in = io.vz_in;
    real *__restrict__ p_out = io.p_out;
    real *__restrict__ vx_out = io.vx_out;
    real *__restrict__ vy_out = io.vy_out;
    real *__restrict__ vz_out = io.vz_out;
    real *__restrict__ p_y = m.p_y;
    real *__restrict__ p_z = m.p_z;

    // Neighbours that exist, and clamped offsets to read in their place when they do not.
    const bool has_ylo = j > 0, has_yhi = j + 1 < ny;
    const bool has_zlo = k0 > 0, has_zhi = k0 + V < nz;
    const int64_t dy_lo = has_ylo ? nz : 0, dy_hi = has_yhi ? nz : 0;
    const int64_t dz_lo = has_zlo ? 1 : 0, dz_hi = has_zhi ? V : V - 1;
    // Face coefficients of the lane's y faces (rows j, j + 1) and z faces (k0 .. k0 + V), cell
    // coefficients of its row and cells.
    const real fay_lo = m.face_a[1][j], fby_lo = m.face_b[1][j];
    const real fay_hi = m.face_a[1][j + 1], fby_hi = m.face_b[1][j + 1];
    const real cay = m.cell_a[1][j], cby = m.cell_b[1][j];
    real faz[V + 1], fbz[V + 1], caz[V], cbz[V];
    bool layer_yz[V];
    bool any_yz = false;
#pragma unroll
    for (int e = 0; e <= V; ++e) {
        faz[e] = m.face_a[2][k0 + e];
        fbz[e] = m.face_b[2][k0 + e];
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
        caz[e] = m.cell_a[2][k0 + e];
        cbz[e] = m.cell_b[2][k0 + e];
        layer_yz[e] = cay != one || caz[e] != one;
        any_yz = any_yz || layer_yz[e];
    }
    // Planes of the lane's sources and receivers (none when ev_lo > ev_hi).
    int ev_lo = nx, ev_hi = -1;
    for (int s = 0; s < a.n_sources; ++s) {
        const int e = a.source_cell[2 * s + 1] - jk0;
        if (e >= 0 && e < V) {
            ev_lo = min(ev_lo, a.source_cell[2 * s]);
            ev_hi = max(ev_hi, a.source_cell[2 * s]);
        }
    }
    for (int r = 0; r < a.n_receivers; ++r) {
        const int e = a.receiver_cell[2 * r + 1] - jk0;
        if (e >= 0 && e < V) {
            ev_lo = min(ev_lo, a.receiver_cell[2 * r]);
            ev_hi = max(ev_hi, a.receiver_cell[2 * r]);
        }
    }

    // The new vx on face i0, the lower face of the march's first cell (face 0 is a wall).
    real p_c[V], vx_lo[V];
    vload<real, V>(p_in + L.cell(i0), p_c);
    {
        real v[V], prev[V];
        vload<real, V>(vx_in + L.cell(i0), v);
        vload<real, V>(p_in + L.cell(i0 > 0 ? i0 - 1 : i0), prev);
        const real fa = uniform_load(m.face_a[0] + i0), fb = uniform_load(m.face_b[0] + i0);
#pragma unroll
        for (int e = 0; e < V; ++e)
            vx_lo[e] = i0 > 0 ? fa * v[e] - fb * (vel * (p_c[e] - prev[e])) : v[e];
        vstore<real, V>(vx_out + L.cell(i0), vx_lo);
    }

    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const bool up_interior = i + 1 < nx;
        // Loads, none behind a branch (missing neighbours read in place and not used).
        real vx_v[V], p_up[V], vy_lo_in[V], vy_hi_in[V], q_lo[V], q_hi[V], vz_v[V];
        vload<real, V>(vx_in + c + L.plane, vx_v);
        vload<real, V>(p_in + c + (up_interior ? L.plane : 0), p_up);
        const int64_t cy = L.face_y(i);
        vload<real, V>(vy_in + cy, vy_lo_in);
        vload<real, V>(vy_in + cy + nz, vy_hi_in);
        vload<real, V>(p_in + c - dy_lo, q_lo);
        vload<real, V>(p_in + c + dy_hi, q_hi);
        const int64_t cz = L.face_z(i);
        vload<real, V>(vz_in + cz, vz_v);
        const real vz_top = vz_in[cz + V];
        const real pz_lo = p_in[c - dz_lo];
        const real pz_hi = p_in[c + dz_hi];
        const real fax = uniform_load(m.face_a[0] + i + 1), fbx = uniform_load(m.face_b[0] + i + 1);
        const real cax = uniform_load(m.cell_a[0] + i), cbx = uniform_load(m.cell_b[0] + i);
        const bool layer_x = cax != one;

        // Faces: interior faces updated, wall faces kept (rigid walls behind the layer).
        real vx_hi[V], vy_lo[V], vy_hi[V], vz_f[V + 1];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            vx_hi[e] = up_interior ? fax * vx_v[e] - fbx * (vel * (p_up[e] - p_c[e])) : vx_v[e];
            vy_lo[e] = has_ylo ? fay_lo * vy_lo_in[e] - fby_lo * (vel * (p_c[e] - q_lo[e]))
                               : vy_lo_in[e];
            vy_hi[e] = has_yhi ? fay_hi * vy_hi_in[e] - fby_hi * (vel * (q_hi[e] - p_c[e]))
                               : vy_hi_in[e];
            if (e > 0)
                vz_f[e] = faz[e] * vz_v[e] - fbz[e] * (vel * (p_c[e] - p_c[e - 1]));
            else
                vz_f[0] = has_zlo ? faz[0] * vz_v[0] - fbz[0] * (vel * (p_c[0] - pz_lo)) : vz_v[0];
        }
        vz_f[V] = has_zhi ? faz[V] * vz_top - fbz[V] * (vel * (pz_hi - p_c[V - 1])) : vz_top;
        if (i + 1 < i1 || i + 1 == nx) vstore<real, V>(vx_out + c + L.plane, vx_hi);
        vstore<real, V>(vy_out + cy, vy_lo);
        if (!has_yhi) vstore<real, V>(vy_out + cy + nz, vy_hi);
        {
            real lower[V];
#pragma unroll
            for (int e = 0; e < V; ++e) lower[e] = vz_f[e];
            vstore<real, V>(vz_out + cz, lower);
        }
        if (!has_zhi) vz_out[cz + V] = vz_f[V];

        // Pressure: the split update in the layer, the plain one elsewhere.
        const bool split = layer_x || any_yz;
        real py[V], pz[V], p_new[V];
        if (split) {
            vload<real, V>(p_y + c, py);
            vload<real, V>(p_z + c, pz);
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real dx = vx_hi[e] - vx_lo[e];
            const real dy = vy_hi[e] - vy_lo[e];
            const real dz = vz_f[e + 1] - vz_f[e];
            if (split && (layer_x || layer_yz[e])) {
                const real px = (p_c[e] - py[e]) - pz[e];
                const real nx_ = cax * px - cbx * (pres * dx);
                const real ny_ = cay * py[e] - cby * (pres * dy);
                const real nz_ = caz[e] * pz[e] - cbz[e] * (pres * dz);
                py[e] = ny_;
                pz[e] = nz_;
                p_new[e] = (nx_ + ny_) + nz_;
            } else {
                p_new[e] = p_c[e] - ((dx + dy) + dz) * pres;
            }
        }
        if (split) {
            vstore<real, V>(p_y + c, py);
            vstore<real, V>(p_z + c, pz);
        }
        if (i >= ev_lo && i <= ev_hi) {
            // Sources in insertion order, then receivers (pml_pressure).
            for (int s = 0; s < a.n_sources; ++s) {
                const int e = a.source_cell[2 * s + 1] - jk0;
                if (a.source_cell[2 * s] == i && e >= 0 && e < V && step < a.source_length[s])
                    lane_add<real, V>(p_new, e, a.src * a.source_signal[s * a.source_stride + step]);
            }
        }
        vstore<real, V>(p_out + c, p_new);
        if (i >= ev_lo && i <= ev_hi) {
            for (int r = 0; r < a.n_receivers; ++r) {
                const int e = a.receiver_cell[2 * r + 1] - jk0;
                if (a.receiver_cell[2 * r] == i && e >= 0 && e < V)
                    a.recording[r * a.recording_stride + column] = lane_get<real, V>(p_new, e);
            }
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            vx_lo[e] = vx_hi[e];
            p_c[e] = p_up[e];
        }
    }
}

template <typename real, int V>
__global__ __launch_bounds__(512) void pml_fused_step(const SweepArgs<real> a, const PMLArgs<real> m,
                                                      const FusedIO<real> io, const int64_t step,
                                                      const int64_t column) {
    int kb, jb, xc, j, k0;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    if (!flat_lane(a.tiles, a.ny, a.zp, jb, j, k0)) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    pml_march<real, V>(a, m, io, i0, i1, j, k0, step, column);
}

template <typename real>
using PMLFusedKernel = void (*)(const SweepArgs<real>, const PMLArgs<real>, const FusedIO<real>,
                                const int64_t, const int64_t);

template <typename real>
PMLFusedKernel<real> pick_pml_fused(int vec) {
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) return pml_fused_step<real, 4>;
    }
    if (vec == 2) return pml_fused_step<real, 2>;
    return pml_fused_step<real, 1>;
}

template <typename real>
int pml_run(const acfdtd_hip_desc &d, const acfdtd_hip_pml_desc &pml, int64_t step_offset,
            int64_t n_steps, int64_t column, hipStream_t s) {
    SweepArgs<real> a;
    fill_args(d, a);
    PMLArgs<real> m{};
    for (int q = 0; q < 3; ++q) {
        m.cell_a[q] = static_cast<const real *>(pml.cell_a[q]);
        m.cell_b[q] = static_cast<const real *>(pml.cell_b[q]);
        m.face_a[q] = static_cast<const real *>(pml.face_a[q]);
        m.face_b[q] = static_cast<const real *>(pml.face_b[q]);
    }
    m.p_y = static_cast<real *>(pml.p_y);
    m.p_z = static_cast<real *>(pml.p_z);
    int64_t n = 0;
    if (has_swap(d)) {
        // Fused steps in pairs, from the first set to the swap set and back.
        SweepArgs<real> f = a;
        f.tiles = fused_tiling(d);
        const PMLFusedKernel<real> k = pick_pml_fused<real>(f.tiles.vec);
        const FusedIO<real> forward = fused_io<real>(d, false), backward = fused_io<real>(d, true);
        const dim3 fgrid(f.tiles.launched), fblock(64, f.tiles.tile_j);
        for (; n + 1 < n_steps; n += 2) {
            hipLaunchKernelGGL(k, fgrid, fblock, 0, s, f, m, forward, step_offset + n, column + n);
            hipLaunchKernelGGL(k, fgrid, fblock, 0, s, f, m, backward, step_offset + n + 1,
                               column + n + 1);
        }
    }
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    for (; n < n_steps; ++n) {
        hipLaunchKernelGGL(pml_velocity<real>, grid, block, 0, s, a, m);
        hipLaunchKernelGGL(pml_pressure<real>, grid, block, 0, s, a, m, step_offset + n, column + n);
    }
    return check_hip(hipGetLastError(), "PML launch");
}

}  // namespace

}  // namespace acfdtd

using namespace acfdtd;

extern "C" int acfdtd_hip_pml_run(const acfdtd_hip_desc *desc, const acfdtd_hip_pml_desc *pml,
                                  int64_t step_offset, int64_t n_steps, int64_t recording_column,
                                  void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (pml == nullptr) return fail(ACFDTD_HIP_EINVAL, "PML descriptor is NULL");
    if (pml->abi_version != ACFDTD_HIP_ABI_VERSION || pml->struct_size != sizeof(acfdtd_hip_pml_desc))
        return fail(ACFDTD_HIP_EABI, "PML descriptor version or size mismatch");
    for (int q = 0; q < 3; ++q)
        if (!pml->cell_a[q] || !pml->cell_b[q] || !pml->face_a[q] || !pml->face_b[q])
            return fail(ACFDTD_HIP_EINVAL, "PML profile is NULL");
    if (!pml->p_y || !pml->p_z) return fail(ACFDTD_HIP_EINVAL, "PML split fields are NULL");
    if (desc->wall_mask != 0 || desc->n_relax != 0 || desc->has_layer || desc->flags != 0 ||
        desc->wall_tape)
        return fail(ACFDTD_HIP_EINVAL,
                    "the PML runs with rigid walls, no air, no sponge, no flags and no tape");
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + n_steps > desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording is too short for this many steps");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return pml_run<double>(*desc, *pml, step_offset, n_steps, recording_column, s);
    return pml_run<float>(*desc, *pml, step_offset, n_steps, recording_column, s);
}
