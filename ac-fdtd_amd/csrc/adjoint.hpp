// Adjoint pieces shared by the two-sweep adjoint (adjoint.hip) and the fused adjoint step
// (adjoint_fused.hip): argument block, gradient slots, per-workgroup partial sums.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "acfdtd_hip.h"
#include "sweeps.hpp"

namespace acfdtd {

enum : int { kGradWall = 0, kGradSponge = 6, kGradScale = 7, kGradGain = 8, kGradSlots = 12 };

template <typename real>
struct AdjointArgs {
    real *lp, *lvx, *lvy, *lvz, *lpsi, *ldiv;
    const real *grad_recording;
    int64_t grad_recording_stride;
    int64_t grad_recording_step0;
    real *grad_signal;
    const real *fwd_p, *fwd_vx, *fwd_vy, *fwd_vz;
    const real *w_cell[3];
    const real *w_face[3];
    double *partials;
    uint32_t grads;
    real dt;
};

template <typename real>
void fill_adjoint(const acfdtd_hip_adjoint_desc &g, AdjointArgs<real> &a) {
    a.lp = static_cast<real *>(g.lp);
    a.lvx = static_cast<real *>(g.lvx);
    a.lvy = static_cast<real *>(g.lvy);
    a.lvz = static_cast<real *>(g.lvz);
    a.lpsi = static_cast<real *>(g.lpsi);
    a.ldiv = static_cast<real *>(g.ldiv);
    a.grad_recording = static_cast<const real *>(g.grad_recording);
    a.grad_recording_stride = g.grad_recording_stride;
    a.grad_recording_step0 = g.grad_recording_step0;
    a.grad_signal = static_cast<real *>(g.grad_signal);
    a.fwd_p = static_cast<const real *>(g.fwd_p);
    a.fwd_vx = static_cast<const real *>(g.fwd_vx);
    a.fwd_vy = static_cast<const real *>(g.fwd_vy);
    a.fwd_vz = static_cast<const real *>(g.fwd_vz);
    for (int q = 0; q < 3; ++q) {
        a.w_cell[q] = static_cast<const real *>(g.weight_cell[q]);
        a.w_face[q] = static_cast<const real *>(g.weight_face[q]);
    }
    a.partials = g.partials;
    a.grads = g.grads;
    a.dt = (real)g.dt;
}

// Like decode_tile, also returning the tile's own index (its gradient slot).
__device__ inline bool decode_tile_slot(const Tiling &t, int &kb, int &jb, int &xc, int &slot) {
    const unsigned n = (unsigned)t.launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);
    if (b >= (unsigned)t.tiles) return false;
    slot = (int)b;
    kb = (int)(b % (unsigned)t.nblk_k);
    b /= (unsigned)t.nblk_k;
    jb = (int)(b % (unsigned)t.nblk_j);
    xc = (int)(b / (unsigned)t.nblk_j);
    return true;
}

// Add each wave's sums of acc[FIRST, FIRST + COUNT) into its own row of `partials`,
// partials[(slot * blockDim.y + wave) * kGradSlots + q]. A row belongs to one wave of one
// launch, and launches on a stream run in order, so the totals are deterministic. Waves whose
// lanes all hold zeros skip the reduction (away from the walls, every wave when only the wall
// gradients are asked for). `tile` must be uniform over the workgroup.
template <int FIRST = 0, int COUNT = kGradSlots>
__device__ inline void flush_partials(const double (&acc)[kGradSlots], bool tile, int slot,
                                      double *partials) {
    if (!tile) return;
    bool nonzero = false;
#pragma unroll
    for (int q = FIRST; q < FIRST + COUNT; ++q) nonzero = nonzero || acc[q] != 0;
    if (!__any(nonzero)) return;
    double sum[COUNT];
#pragma unroll
    for (int q = 0; q < COUNT; ++q) sum[q] = wave_sum(acc[FIRST + q]);
    if (threadIdx.x == 0) {
        double *row = partials + ((int64_t)slot * blockDim.y + threadIdx.y) * kGradSlots + FIRST;
        double old[COUNT];
#pragma unroll
        for (int q = 0; q < COUNT; ++q) old[q] = row[q];
#pragma unroll
        for (int q = 0; q < COUNT; ++q) row[q] = old[q] + sum[q];
    }
}

// Forward state fields by slot of the ring acfdtd_hip_adjoint_desc::fwd_* describes, or the
// single state the face and cell sweeps take (slot 0).
template <typename real>
struct ForwardState {
    const real *p, *vx, *vy, *vz;
};

template <typename real>
ForwardState<real> forward_slot(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &g,
                                int64_t slot) {
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    const int64_t nvx = (int64_t)(d.nx + 1) * d.ny * d.nz;
    const int64_t nvy = (int64_t)d.nx * (d.ny + 1) * d.nz;
    const int64_t nvz = (int64_t)d.nx * d.ny * (d.nz + 1);
    ForwardState<real> s{};
    if (g.fwd_p) s.p = static_cast<const real *>(g.fwd_p) + slot * n;
    if (g.fwd_vx) s.vx = static_cast<const real *>(g.fwd_vx) + slot * nvx;
    if (g.fwd_vy) s.vy = static_cast<const real *>(g.fwd_vy) + slot * nvy;
    if (g.fwd_vz) s.vz = static_cast<const real *>(g.fwd_vz) + slot * nvz;
    return s;
}

// Waves (tiles x rows per tile) of the fused adjoint step, for sizing the partials; defined in
// adjoint_fused.hip.
int64_t adjoint_fused_waves(const acfdtd_hip_desc &d);

// Launch fused adjoint steps m = step_hi, step_hi - 1, ... for the leading even number of
// n_steps, in pairs from the first set to the swap set and back; returns the number launched.
// Defined in adjoint_fused.hip.
template <typename real>
int64_t launch_adjoint_pairs(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &g,
                             int64_t step_hi, int64_t n_steps, hipStream_t stream);

// Launch two-step adjoint sweeps for the leading multiple of four of n_steps (steps step_hi,
// step_hi - 1, ..., each with its local part), in pairs from the first set to the swap set and
// back; returns the number of steps done (0 when the descriptors do not allow it). Defined in
// adjoint2.hip and adjoint2_64.hip.
template <typename real>
int64_t launch_adjoint2_quads(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &g,
                              int64_t step_hi, int64_t n_steps, hipStream_t stream);

inline bool has_adjoint_swap(const acfdtd_hip_adjoint_desc &g) {
    return g.ldiv_swap && g.lvx_swap && g.lvy_swap && g.lvz_swap;
}

}  // namespace acfdtd
