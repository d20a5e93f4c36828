// The split-field perfectly matched layer on gfx950: a velocity sweep and a pressure sweep per
// step (include/acfdtd_hip.h, acfdtd_hip_pml_desc, states the arithmetic; the CPU oracle
// restates it and these kernels reproduce it bit for bit, built with -ffp-contract=off).
//
// Mapping: the first-order two sweeps' (acfdtd.c:110-178 structure): 64 lanes along z per
// wavefront, tile_j rows along y, a chunk of planes marched along x with the x neighbour in a
// register. Outside the layer a cell does the reference's update and touches no split part;
// in the layer it reads and writes p_y and p_z too. So the split parts cost HBM traffic only
// in the layer (two extra reads and writes per layer cell), and the layer's per-axis
// coefficients are 1D arrays read through the scalar and L1 caches.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
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
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    for (int64_t n = 0; n < n_steps; ++n) {
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
