// The two-step sweep: kernel, dispatch, launch of the leading multiple of four steps. Templates
// only; fused2.hip and fused2_64.hip instantiate the launchers and the kernels of 16-byte
// lanes (4 cells fp32, 2 fp64), fused2_v2.hip, fused2_v1.hip and fused2_64_v1.hip those of
// narrower lanes (nz not a multiple of 4 or 2), fused2s.hip and fused2s_64.hip the staged ones.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "acfdtd_hip.h"
#include "marches.hpp"
#include "marches2.hpp"
#include "sweeps.hpp"

namespace acfdtd {

// Waves per SIMD the compiler must leave room for (caps the registers per lane: 2 -> 256,
// 3 -> 168, 4 -> 128).
#ifndef ACFDTD_FUSED2_MIN_WAVES
#define ACFDTD_FUSED2_MIN_WAVES 2
#endif

// One workgroup of the two-step sweep on tile b of a room (b < t2.tiles): the body of
// fused2_step, also run by batched rooms. The LDS it uses, declared by the kernel: lds, the
// exchanged p^(n+1) planes (3 x kFused2Threads x V values); the source and receiver tables,
// copied once (the lanes that hold a source or a receiver scan them every plane, and with a
// barrier per plane the whole workgroup waits for those lanes; in LDS their reads wait for
// nothing else, while a read from global memory waits for every store the lane issued before
// it); and, when STAGED, each lane's slots for the next plane's first-step inputs.
template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED, bool STAGED,
          bool PADDED = false>
__device__ __forceinline__ void fused2_block(const SweepArgs<real> &a, const Fused2IO<real> &io,
                                             const Tiling2 &t2, const unsigned b,
                                             const int64_t step, const int64_t column,
                                             real *lds, int32_t *src_lds, int64_t *len_lds,
                                             real *now_lds, int32_t *rx_lds,
                                             unsigned char *stage) {
    const int blk = (int)(b % (unsigned)t2.nblk);
    const int xc = (int)(b / (unsigned)t2.nblk);
    const int bg = blk % t2.nblk_g, bj = blk / t2.nblk_g;
    const int t = (int)threadIdx.x;
    const int r = t / t2.pitch - 1;             // block row, -1 and rows: halo
    const int sl = t % t2.pitch - t2.zhalo;     // segment position, -1 and seg: halo
    const int jj = bj * t2.rows + r;
    const int gg = bg * t2.seg + sl;
    const bool valid = t < (t2.rows + 2) * t2.pitch && jj >= 0 && jj < a.ny && gg >= 0 &&
                       gg < t2.groups;
    const bool owner = valid && r >= 0 && r < t2.rows && sl >= 0 && sl < t2.seg;
    const int j = valid ? jj : 0;
    const int k0 = valid ? gg * V : 0;
    int i0, i1;
    chunk_planes(t2, blk, xc, a.nx, i0, i1);
    // fused2_applies admits at most kFused2Sources sources and kFused2Receivers receivers.
    for (int q = t; q < a.n_sources; q += (int)blockDim.x) {
        src_lds[2 * q] = a.source_cell[2 * q];
        src_lds[2 * q + 1] = a.source_cell[2 * q + 1];
        const int64_t len = a.source_length[q];
        len_lds[q] = len;
        // Indexed by step parity; a sample past the signal's end is never read.
        for (int h = 0; h < 2; ++h) {
            const int64_t n = step + h;
            now_lds[2 * q + (int)(n & 1)] = n < len ? a.source_signal[q * a.source_stride + n] : real(0);
        }
    }
    for (int q = t; q < 2 * a.n_receivers; q += (int)blockDim.x) rx_lds[q] = a.receiver_cell[q];
    lds_barrier();
    fused2_march<real, V, NR, WALLS, LAYER, COMPENSATED, STAGED, PADDED>(
        a, t2, i0, i1, j, k0, valid, owner, t, step, column, io.p_in, io.vx_in, io.vy_in,
        io.vz_in, io.psi_in, io.p_out, io.vx_out, io.vy_out, io.vz_out, io.psi_out, lds, src_lds,
        len_lds, now_lds, rx_lds, stage);
}

// The LDS declarations of a two-step kernel (see fused2_block).
#define ACFDTD_FUSED2_SHARED(real, V, NR, STAGED)                                          \
    __shared__ real lds[3 * kFused2Threads * V];                                           \
    __shared__ __align__(16) unsigned char stage[STAGED ? stage_bytes<real, V, NR>() : 16]; \
    __shared__ int32_t src_lds[2 * kFused2Sources];                                        \
    __shared__ int64_t len_lds[kFused2Sources];                                            \
    __shared__ real now_lds[2 * kFused2Sources];                                           \
    __shared__ int32_t rx_lds[2 * kFused2Receivers]

// The single-room kernel keeps its own copy of the workgroup body (fused2_block, which the
// batched kernel runs, is the same code): through the shared function the compiler spills
// more scalar registers here (about 800 more lane reads of spilled scalars per kernel), which
// measured 6 % slower at 256^3.
template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED, bool STAGED,
          bool DIVT = false, bool PADDED = false>
__global__ __launch_bounds__(kFused2Threads, ACFDTD_FUSED2_MIN_WAVES) void fused2_step(const SweepArgs<real> a,
                                                             const Fused2IO<real> io,
                                                             const Tiling2 t2, const int64_t step,
                                                             const int64_t column) {
    __shared__ real lds[3 * kFused2Threads * V];
    // STAGED: each lane's slots for the next plane's first-step inputs (fused2_march).
    __shared__ __align__(16) unsigned char stage[STAGED ? stage_bytes<real, V, NR>() : 16];
    // Source and receiver tables, copied once: the lanes that hold a source or a receiver scan
    // them every plane, and with a barrier per plane the whole workgroup waits for those lanes.
    // In LDS their reads wait for nothing else; a read from global (or generic) memory waits
    // for every store the lane issued before it.
    __shared__ int32_t src_lds[2 * kFused2Sources];
    __shared__ int64_t len_lds[kFused2Sources];
    __shared__ real now_lds[2 * kFused2Sources];  // samples of steps `step` and `step` + 1
    __shared__ int32_t rx_lds[2 * kFused2Receivers];
    const unsigned n = (unsigned)t2.launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);  // contiguous tiles per XCD, as decode_tile
    if (b >= (unsigned)t2.tiles) return;  // whole workgroup, before any barrier
    const int blk = (int)(b % (unsigned)t2.nblk);
    const int xc = (int)(b / (unsigned)t2.nblk);
    const int bg = blk % t2.nblk_g, bj = blk / t2.nblk_g;
    const int t = (int)threadIdx.x;
    const int r = t / t2.pitch - 1;             // block row, -1 and rows: halo
    const int sl = t % t2.pitch - t2.zhalo;     // segment position, -1 and seg: halo
    const int jj = bj * t2.rows + r;
    const int gg = bg * t2.seg + sl;
    const bool valid = t < (t2.rows + 2) * t2.pitch && jj >= 0 && jj < a.ny && gg >= 0 &&
                       gg < t2.groups;
    const bool owner = valid && r >= 0 && r < t2.rows && sl >= 0 && sl < t2.seg;
    const int j = valid ? jj : 0;
    const int k0 = valid ? gg * V : 0;
    int i0, i1;
    chunk_planes(t2, blk, xc, a.nx, i0, i1);
    // fused2_applies admits at most kFused2Sources sources and kFused2Receivers receivers.
    for (int q = t; q < a.n_sources; q += (int)blockDim.x) {
        src_lds[2 * q] = a.source_cell[2 * q];
        src_lds[2 * q + 1] = a.source_cell[2 * q + 1];
        const int64_t len = a.source_length[q];
        len_lds[q] = len;
        // Indexed by step parity; a sample past the signal's end is never read.
        for (int h = 0; h < 2; ++h) {
            const int64_t n = step + h;
            now_lds[2 * q + (int)(n & 1)] = n < len ? a.source_signal[q * a.source_stride + n] : real(0);
        }
    }
    for (int q = t; q < 2 * a.n_receivers; q += (int)blockDim.x) rx_lds[q] = a.receiver_cell[q];
    lds_barrier();
    fused2_march<real, V, NR, WALLS, LAYER, COMPENSATED, STAGED, PADDED, DIVT>(
        a, t2, i0, i1, j, k0, valid, owner, t, step, column, io.p_in, io.vx_in, io.vy_in,
        io.vz_in, io.psi_in, io.p_out, io.vx_out, io.vy_out, io.vz_out, io.psi_out, lds, src_lds,
        len_lds, now_lds, rx_lds, stage);
}

template <typename real>
using Fused2Kernel = void (*)(const SweepArgs<real>, const Fused2IO<real>, const Tiling2,
                              const int64_t, const int64_t);

template <typename real, int V, int NR, bool STAGED, bool DIVT = false, bool PADDED = false>
Fused2Kernel<real> pick_fused2_nr(bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (comp) {
            if (walls) return layer ? fused2_step<real, V, NR, true, true, true, STAGED, DIVT, PADDED>
                                    : fused2_step<real, V, NR, true, false, true, STAGED, DIVT, PADDED>;
            return layer ? fused2_step<real, V, NR, false, true, true, STAGED, DIVT, PADDED>
                         : fused2_step<real, V, NR, false, false, true, STAGED, DIVT, PADDED>;
        }
    }
    (void)comp;
    if (walls) return layer ? fused2_step<real, V, NR, true, true, false, STAGED, DIVT, PADDED>
                            : fused2_step<real, V, NR, true, false, false, STAGED, DIVT, PADDED>;
    return layer ? fused2_step<real, V, NR, false, true, false, STAGED, DIVT, PADDED>
                 : fused2_step<real, V, NR, false, false, false, STAGED, DIVT, PADDED>;
}

// The kernel for a lane width V (cells per lane), plain or staged.
template <typename real, int V, bool STAGED>
Fused2Kernel<real> pick_fused2_kind(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_fused2_nr<real, V, 0, STAGED>(walls, layer, comp);
        case 1: return pick_fused2_nr<real, V, 1, STAGED>(walls, layer, comp);
        case 2: return pick_fused2_nr<real, V, 2, STAGED>(walls, layer, comp);
        case 3: return pick_fused2_nr<real, V, 3, STAGED>(walls, layer, comp);
        default: return pick_fused2_nr<real, V, 4, STAGED>(walls, layer, comp);
    }
}

// The plain sweep on padded rows (z_pitch), for the widest lanes (V = 4 fp32, 2 fp64).
template <typename real, int V>
Fused2Kernel<real> pick_fused2_padded(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_fused2_nr<real, V, 0, false, false, true>(walls, layer, comp);
        case 1: return pick_fused2_nr<real, V, 1, false, false, true>(walls, layer, comp);
        case 2: return pick_fused2_nr<real, V, 2, false, false, true>(walls, layer, comp);
        case 3: return pick_fused2_nr<real, V, 3, false, false, true>(walls, layer, comp);
        default: return pick_fused2_nr<real, V, 4, false, false, true>(walls, layer, comp);
    }
}

// The plain sweep that also writes the gradient tapes (divergence, layer), for the widest
// lanes (V = 4 fp32, 2 fp64); without relaxation states only the layer tape applies.
template <typename real, int V>
Fused2Kernel<real> pick_fused2_divt(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_fused2_nr<real, V, 0, false, true>(walls, layer, comp);
        case 1: return pick_fused2_nr<real, V, 1, false, true>(walls, layer, comp);
        case 2: return pick_fused2_nr<real, V, 2, false, true>(walls, layer, comp);
        case 3: return pick_fused2_nr<real, V, 3, false, true>(walls, layer, comp);
        default: return pick_fused2_nr<real, V, 4, false, true>(walls, layer, comp);
    }
}

// Each (precision, lane width, staged) family of kernels is instantiated in a translation unit
// of its own (fused2*.hip), so that they compile in parallel; elsewhere they are only declared.
extern template Fused2Kernel<float> pick_fused2_kind<float, 4, false>(int, bool, bool, bool);
extern template Fused2Kernel<float> pick_fused2_kind<float, 4, true>(int, bool, bool, bool);
extern template Fused2Kernel<float> pick_fused2_kind<float, 2, false>(int, bool, bool, bool);
extern template Fused2Kernel<float> pick_fused2_kind<float, 1, false>(int, bool, bool, bool);
extern template Fused2Kernel<double> pick_fused2_kind<double, 2, false>(int, bool, bool, bool);
extern template Fused2Kernel<double> pick_fused2_kind<double, 2, true>(int, bool, bool, bool);
extern template Fused2Kernel<double> pick_fused2_kind<double, 1, false>(int, bool, bool, bool);
extern template Fused2Kernel<float> pick_fused2_divt<float, 4>(int, bool, bool, bool);
extern template Fused2Kernel<double> pick_fused2_divt<double, 2>(int, bool, bool, bool);
extern template Fused2Kernel<float> pick_fused2_padded<float, 4>(int, bool, bool, bool);
extern template Fused2Kernel<double> pick_fused2_padded<double, 2>(int, bool, bool, bool);

// Whether the staged variant may run: asked for, 16-byte lanes (4 cells fp32, 2 fp64), and
// every field it stages 16-byte aligned in both sets (otherwise the plain two-step sweep runs,
// with the same results).
inline bool staged_applies(const acfdtd_hip_desc &d, const int vec) {
    if (!(d.flags & ACFDTD_HIP_STAGED) || vec * d.real_bytes != 16 || d.div_tape || d.layer_tape ||
        d.z_pitch)
        return false;
    const void *fields[] = {d.p, d.vx, d.vy, d.p_swap, d.vx_swap, d.vy_swap};
    for (const void *f : fields)
        if (reinterpret_cast<uintptr_t>(f) % 16 != 0) return false;
    if (d.n_relax > 0 && (reinterpret_cast<uintptr_t>(d.psi) % 16 != 0 ||
                          reinterpret_cast<uintptr_t>(d.psi_swap) % 16 != 0))
        return false;
    return true;
}

// The two-step sweep's kernel for d: lanes as wide as nz allows (as the fused step's).
template <typename real>
Fused2Kernel<real> pick_fused2(const acfdtd_hip_desc &d) {
    const int vec = fused_vec(d);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const bool walls = d.wall_mask != 0, layer = d.has_layer != 0;
    const bool staged = staged_applies(d, vec);
    if (d.z_pitch > 0) {
        // fused2_rows_apply: the widest lanes only.
        if constexpr (sizeof(real) == 4) return pick_fused2_padded<real, 4>(d.n_relax, walls, layer, comp);
        else return pick_fused2_padded<real, 2>(d.n_relax, walls, layer, comp);
    }
    if (d.div_tape || d.layer_tape) {
        // fused2_divt_applies: the widest lanes only.
        if constexpr (sizeof(real) == 4) return pick_fused2_divt<real, 4>(d.n_relax, walls, layer, comp);
        else return pick_fused2_divt<real, 2>(d.n_relax, walls, layer, comp);
    }
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) return staged ? pick_fused2_kind<real, 4, true>(d.n_relax, walls, layer, comp)
                                    : pick_fused2_kind<real, 4, false>(d.n_relax, walls, layer, comp);
        if (vec == 2) return pick_fused2_kind<real, 2, false>(d.n_relax, walls, layer, comp);
        return pick_fused2_kind<real, 1, false>(d.n_relax, walls, layer, comp);
    } else {
        if (vec == 2) return staged ? pick_fused2_kind<real, 2, true>(d.n_relax, walls, layer, comp)
                                    : pick_fused2_kind<real, 2, false>(d.n_relax, walls, layer, comp);
        return pick_fused2_kind<real, 1, false>(d.n_relax, walls, layer, comp);
    }
}

// The two directions of a two-step sweep: forward reads the descriptor's fields and writes the
// swap set, backward the other way round.
template <typename real>
void fused2_io(const acfdtd_hip_desc &d, Fused2IO<real> &forward, Fused2IO<real> &backward) {
    forward = Fused2IO<real>{};
    backward = Fused2IO<real>{};
    forward.p_in = backward.p_out = static_cast<real *>(d.p);
    forward.vx_in = backward.vx_out = static_cast<real *>(d.vx);
    forward.vy_in = backward.vy_out = static_cast<real *>(d.vy);
    forward.vz_in = backward.vz_out = static_cast<real *>(d.vz);
    forward.psi_in = backward.psi_out = static_cast<real *>(d.psi);
    forward.p_out = static_cast<real *>(d.p_swap);
    forward.vx_out = static_cast<real *>(d.vx_swap);
    forward.vy_out = static_cast<real *>(d.vy_swap);
    forward.vz_out = static_cast<real *>(d.vz_swap);
    forward.psi_out = static_cast<real *>(d.psi_swap);
    backward.p_in = forward.p_out;
    backward.vx_in = forward.vx_out;
    backward.vy_in = forward.vy_out;
    backward.vz_in = forward.vz_out;
    backward.psi_in = forward.psi_out;
}

// With a divergence or layer tape, and with padded rows, the two-step sweep has kernels for the
// widest lanes only.
inline bool fused2_divt_applies(const acfdtd_hip_desc &d) {
    return (d.div_tape == nullptr && d.layer_tape == nullptr && d.z_pitch == 0) ||
           fused_vec(d) * d.real_bytes == 16;
}

template <typename real>
int64_t launch_fused2_quads(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps,
                            int64_t column, hipStream_t stream) {
    if (!fused2_applies(d) || !fused2_divt_applies(d)) return 0;
    const Fused2Kernel<real> k = pick_fused2<real>(d);
    Tiling2 t2;
    if (!make_tiling2(d, fused_vec(d), t2, reinterpret_cast<const void *>(k))) return 0;
    balance_tiling2(t2, d.nx);
    SweepArgs<real> a;
    fill_args(d, a);
    Fused2IO<real> forward, backward;
    fused2_io<real>(d, forward, backward);
    const dim3 grid(t2.launched), block(t2.threads);
    int64_t n = 0;
    for (; n + 3 < n_steps; n += 4) {
        hipLaunchKernelGGL(k, grid, block, 0, stream, a, forward, t2, step_offset + n, column + n);
        hipLaunchKernelGGL(k, grid, block, 0, stream, a, backward, t2, step_offset + n + 2,
                           column + n + 2);
    }
    return n;
}

// One two-step sweep for steps `step` and `step` + 1 (reverse: from the swap set back);
// ACFDTD_HIP_EINVAL when the descriptor does not allow it.
template <typename real>
int launch_fused2_once(const acfdtd_hip_desc &d, int64_t step, int64_t column, bool reverse,
                       hipStream_t stream) {
    Tiling2 t2;
    const Fused2Kernel<real> k =
        fused2_applies(d) && fused2_divt_applies(d) ? pick_fused2<real>(d) : nullptr;
    if (k == nullptr || !make_tiling2(d, fused_vec(d), t2, reinterpret_cast<const void *>(k)))
        return fail(ACFDTD_HIP_EINVAL, "the two-step sweep does not apply to this descriptor");
    balance_tiling2(t2, d.nx);
    SweepArgs<real> a;
    fill_args(d, a);
    Fused2IO<real> forward, backward;
    fused2_io<real>(d, forward, backward);
    hipLaunchKernelGGL(k,
                       dim3(t2.launched), dim3(t2.threads), 0, stream, a,
                       reverse ? backward : forward, t2, step, column);
    return check_hip(hipGetLastError(), "two-step sweep launch");
}

}  // namespace acfdtd
