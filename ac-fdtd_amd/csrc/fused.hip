// The fused one-sweep step (scalar and V-wide lanes): kernels, dispatch, launch.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {



template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED,
          bool PADDED = false>
__global__ __launch_bounds__(512) void fused_step_vec(const SweepArgs<real> a,
                                                      const FusedIO<real> io, const int64_t step,
                                                      const int64_t column) {
    int kb, jb, xc, j, k0;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    if (!flat_lane(a.tiles, a.ny, a.zp, jb, j, k0)) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    fused_march_vec<real, V, NR, WALLS, LAYER, COMPENSATED, PADDED>(
        a, i0, i1, j, k0, step, column, io.p_in, io.vx_in, io.vy_in, io.vz_in, io.p_out,
        io.vx_out, io.vy_out, io.vz_out, a.psi, a.wall_table, a.cell[0], a.face[0], a.cell[1],
        a.face[1], a.cell[2], a.face[2], a.recording);
}


template <typename real>
using FusedKernel = void (*)(const SweepArgs<real>, const FusedIO<real>, const int64_t,
                             const int64_t);

template <typename real, int V, int NR, bool W, bool L, bool C, bool P>
FusedKernel<real> fused_kernel() {
    return fused_step_vec<real, V, NR, W, L, C, P>;
}

// Compensated kernels exist for fp32 only, and 4-wide lanes only for fp32 (16 bytes).
template <typename real, int V, int NR, bool P>
FusedKernel<real> pick_fused_nr(bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (comp) {
            if (walls) return layer ? fused_kernel<real, V, NR, true, true, true, P>()
                                    : fused_kernel<real, V, NR, true, false, true, P>();
            return layer ? fused_kernel<real, V, NR, false, true, true, P>()
                         : fused_kernel<real, V, NR, false, false, true, P>();
        }
    }
    (void)comp;
    if (walls) return layer ? fused_kernel<real, V, NR, true, true, false, P>()
                            : fused_kernel<real, V, NR, true, false, false, P>();
    return layer ? fused_kernel<real, V, NR, false, true, false, P>()
                 : fused_kernel<real, V, NR, false, false, false, P>();
}

template <typename real, int V, bool P = false>
FusedKernel<real> pick_fused_v(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_fused_nr<real, V, 0, P>(walls, layer, comp);
        case 1: return pick_fused_nr<real, V, 1, P>(walls, layer, comp);
        case 2: return pick_fused_nr<real, V, 2, P>(walls, layer, comp);
        case 3: return pick_fused_nr<real, V, 3, P>(walls, layer, comp);
        default: return pick_fused_nr<real, V, 4, P>(walls, layer, comp);
    }
}

// Padded rows (z_pitch) take the kernels of the widest lanes (4 cells fp32, 2 fp64) only; null
// when the pitch allows narrower lanes alone (the caller then runs the in-place sweeps).
template <typename real>
FusedKernel<real> pick_fused(int vec, int nr, bool walls, bool layer, bool comp,
                             bool padded = false) {
    constexpr int widest = sizeof(real) == 4 ? 4 : 2;
    if (padded) return vec == widest ? pick_fused_v<real, widest, true>(nr, walls, layer, comp)
                                     : nullptr;
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) return pick_fused_v<real, 4>(nr, walls, layer, comp);
    }
    if (vec == 2) return pick_fused_v<real, 2>(nr, walls, layer, comp);
    return pick_fused_v<real, 1>(nr, walls, layer, comp);
}

template <typename real>
int64_t launch_fused_pairs(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps,
                           int64_t column, hipStream_t stream) {
    SweepArgs<real> f;
    fill_args(d, f);
    f.tiles = fused_tiling(d);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const FusedKernel<real> fk = pick_fused<real>(f.tiles.vec, d.n_relax, d.wall_mask != 0,
                                                  d.has_layer != 0, comp, d.z_pitch > 0);
    if (fk == nullptr) return 0;
    const FusedIO<real> forward = fused_io<real>(d, false), backward = fused_io<real>(d, true);
    const dim3 grid(f.tiles.launched), block(64, f.tiles.tile_j);
    int64_t n = 0;
    for (; n + 1 < n_steps; n += 2) {
        hipLaunchKernelGGL(fk, grid, block, 0, stream, f, forward, step_offset + n, column + n);
        hipLaunchKernelGGL(fk, grid, block, 0, stream, f, backward, step_offset + n + 1,
                           column + n + 1);
    }
    return n;
}

template int64_t launch_fused_pairs<float>(const acfdtd_hip_desc &, int64_t, int64_t, int64_t,
                                           hipStream_t);
template int64_t launch_fused_pairs<double>(const acfdtd_hip_desc &, int64_t, int64_t, int64_t,
                                            hipStream_t);

template <typename real>
int fused_once(const acfdtd_hip_desc &d, int64_t step, int64_t column, bool reverse,
               hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    a.tiles = fused_tiling(d);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const FusedKernel<real> fk = pick_fused<real>(a.tiles.vec, d.n_relax, d.wall_mask != 0,
                                                  d.has_layer != 0, comp);
    hipLaunchKernelGGL(fk, dim3(a.tiles.launched), dim3(64, a.tiles.tile_j), 0, stream, a,
                       fused_io<real>(d, reverse), step, column);
    return check_hip(hipGetLastError(), "fused launch");
}

}  // namespace acfdtd

using namespace acfdtd;

extern "C" int acfdtd_hip_fused_step(const acfdtd_hip_desc *desc, int64_t step,
                                     int64_t recording_column, int reverse, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (!has_swap(*desc)) return fail(ACFDTD_HIP_EINVAL, "fused step needs the four swap buffers");
    if (step < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "step and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column >= desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording column out of range");
    if (desc->wall_tape &&
        (step < desc->wall_tape_step0 || step - desc->wall_tape_step0 >= desc->wall_tape_rows))
        return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover this step");
    if (!div_tape_covers(*desc, step, step))
        return fail(ACFDTD_HIP_EINVAL, "divergence tape does not cover this step");
    if (desc->layer_tape) return fail(ACFDTD_HIP_EINVAL, "the layer tape is written by acfdtd_hip_run only");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return fused_once<double>(*desc, step, recording_column, reverse != 0, s);
    return fused_once<float>(*desc, step, recording_column, reverse != 0, s);
}

extern "C" int acfdtd_hip_run_states(const acfdtd_hip_desc *desc, int64_t step_offset,
                                     int64_t n_steps, void *p_states, void *vx_states,
                                     void *vy_states, void *vz_states, void *stream) {
    if (desc == nullptr) return fail(ACFDTD_HIP_EINVAL, "descriptor is NULL");
    if (!p_states || !vx_states || !vy_states || !vz_states)
        return fail(ACFDTD_HIP_EINVAL, "state ring pointer is NULL");
    if (step_offset < 0 || n_steps < 0)
        return fail(ACFDTD_HIP_EINVAL, "step_offset and n_steps must be >= 0");
    // The ring replaces the fields; nothing is recorded and no tape is written.
    acfdtd_hip_desc d = *desc;
    d.p = p_states;
    d.vx = vx_states;
    d.vy = vy_states;
    d.vz = vz_states;
    d.n_receivers = 0;
    d.receiver_cell = nullptr;
    d.recording = nullptr;
    d.recording_stride = 0;
    d.wall_tape = nullptr;
    d.wall_tape_stride = 0;
    d.wall_tape_step0 = 0;
    d.wall_tape_rows = 0;
    d.div_tape = nullptr;
    d.div_tape_step0 = 0;
    d.div_tape_rows = 0;
    d.layer_tape = nullptr;
    d.layer_tape_step0 = 0;
    d.layer_tape_rows = 0;
    d.layer_tape_stride = 0;
    const int status = validate(&d);
    if (status != ACFDTD_HIP_OK) return status;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    const int64_t sizes[4] = {n, (int64_t)(d.nx + 1) * d.ny * d.nz,
                              (int64_t)d.nx * (d.ny + 1) * d.nz, (int64_t)d.nx * d.ny * (d.nz + 1)};
    char *bases[4] = {static_cast<char *>(p_states), static_cast<char *>(vx_states),
                      static_cast<char *>(vy_states), static_cast<char *>(vz_states)};
    void **in[4] = {&d.p, &d.vx, &d.vy, &d.vz};
    void **out[4] = {&d.p_swap, &d.vx_swap, &d.vy_swap, &d.vz_swap};
    for (int64_t q = 0; q < n_steps; ++q) {
        for (int f = 0; f < 4; ++f) {
            *in[f] = bases[f] + q * sizes[f] * d.real_bytes;
            *out[f] = bases[f] + (q + 1) * sizes[f] * d.real_bytes;
        }
        const int rc = d.real_bytes == 8 ? fused_once<double>(d, step_offset + q, 0, false, s)
                                         : fused_once<float>(d, step_offset + q, 0, false, s);
        if (rc != ACFDTD_HIP_OK) return rc;
    }
    return ACFDTD_HIP_OK;
}
