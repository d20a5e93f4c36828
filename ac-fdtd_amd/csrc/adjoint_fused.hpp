// The adjoint step as one sweep, V z-cells per lane: the face sweep and stencil of adjoint
// step m and the cell-local part of step m - 1 (the two kernels of adjoint.hip, in one pass).
// Templates only; adjoint_fused.hip and adjoint_fused64.hip instantiate them for fp32 and fp64.
//
// Adjoint step m reads ldiv (the local part of step m left it) on the six neighbours of each
// cell, as the forward step reads p, and the old adjoint velocities on the cell's faces; it
// writes the new adjoint velocities, and ldiv of step m - 1. Both are therefore ping-ponged
// between two sets (`in` -> `out`), while lp and lpsi, which only their own cell reads, are
// updated in place. Each cell computes the new adjoint velocity on all six of its faces from
// the same inputs with the same operations as the face sweep (so a shared face gets the same
// bits in both cells and in the face sweep), and only the face's owner (the cell above it;
// the last cell for the far wall) writes it and adds its sponge gradient. The arithmetic per
// cell is that of adjoint_face_sweep / adjoint_cell_sweep term by term, so the adjoint fields
// are bit-identical to the two-kernel path; only the fp64 gradient partials are summed in
// another order.
//
// Traffic per cell and adjoint step: ldiv, lvx, lvy, lvz read and written, lp and each lpsi
// read and written: 10 + 2 n_relax elements, against 13 + 2 n_relax for the two sweeps.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "acfdtd_hip.h"
#include "adjoint.hpp"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {

template <typename real>
struct AdjointIO {
    const real *ldiv_in, *lvx_in, *lvy_in, *lvz_in;
    real *ldiv_out, *lvx_out, *lvy_out, *lvz_out;
    // Forward state after step m (x^(m+1)), for the sponge gradient of the faces, and before
    // it (x^m), for the sponge and air gradients of the local part of step m - 1.
    const real *after_vx, *after_vy, *after_vz;
    const real *before_p, *before_vx, *before_vy, *before_vz;
};

template <typename real, int V, int NR, bool LAYER, bool FWDGRAD>
__device__ __forceinline__ void adjoint_march_vec(
    const SweepArgs<real> &a, const AdjointArgs<real> &g, const AdjointIO<real> &io,
    const int i0, const int i1, const int j, const int k0, const int64_t m,
    const bool want_walls, const bool sponge_face, const bool sponge_cell, const bool air,
    const real *__restrict__ ld_in, const real *__restrict__ lvx_in,
    const real *__restrict__ lvy_in, const real *__restrict__ lvz_in,
    real *__restrict__ ld_out, real *__restrict__ lvx_out, real *__restrict__ lvy_out,
    real *__restrict__ lvz_out, real *__restrict__ lp, real *__restrict__ lpsi,
    const real *__restrict__ wall_table, double (&acc)[kGradSlots]) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const Layout L(nx, ny, nz, j, k0);
    const int64_t n_cells = L.plane * nx;
    const int jk0 = j * nz + k0;
    // Air gradients: the divergences of step m - 1 from the divergence tape when there is one,
    // else from the forward velocities of that step.
    const real *div_row =
        FWDGRAD && air && a.div_tape ? a.div_tape + (m - 1 - a.div_tape_step0) * n_cells : nullptr;
    const real vel = a.vel;
    const bool local = m >= 1;
    const int64_t step = m - 1;  // the step whose local part this sweep transposes
    const double mdt = -(double)g.dt;

    const bool has_ylo = j > 0, has_yhi = j + 1 < ny;
    const bool has_zlo = k0 > 0, has_zhi = k0 + V < nz;
    const int64_t dy_lo = has_ylo ? nz : 0, dy_hi = has_yhi ? nz : 0;
    const int64_t dz_lo = has_zlo ? 1 : 0, dz_hi = has_zhi ? V : V - 1;

    // Sponge factors and, for its gradient, grading weights of the lane's rows.
    real fyc = 1, fyf_lo = 1, fyf_hi = 1, fzc[V], fzf[V + 1];
    real wyc = 0, wyf_lo = 0, wyf_hi = 0, wzc[V], wzf[V + 1];
#pragma unroll
    for (int e = 0; e < V; ++e) fzc[e] = 1, wzc[e] = 0;
#pragma unroll
    for (int e = 0; e <= V; ++e) fzf[e] = 1, wzf[e] = 0;
    if (LAYER) {
        fyc = a.cell[1][j];
        fyf_lo = a.face[1][j];
        fyf_hi = a.face[1][j + 1];
#pragma unroll
        for (int e = 0; e < V; ++e) fzc[e] = a.cell[2][k0 + e];
#pragma unroll
        for (int e = 0; e <= V; ++e) fzf[e] = a.face[2][k0 + e];
        if (sponge_face || sponge_cell) {
            wyc = g.w_cell[1][j];
            wyf_lo = g.w_face[1][j];
            wyf_hi = g.w_face[1][j + 1];
#pragma unroll
            for (int e = 0; e < V; ++e) wzc[e] = g.w_cell[2][k0 + e];
#pragma unroll
            for (int e = 0; e <= V; ++e) wzf[e] = g.w_face[2][k0 + e];
        }
    }

    // Wall classes (y and z part) and the factors of interior planes; a wall cell here is one
    // with a nonzero code when the table exists, as in adjoint_cell_sweep.
    int code_yz[V];
    bool wall_in[V];
    real fu_in[V], fp_in[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        code_yz[e] = (wall_class(j, ny) << 2) | (wall_class(k0 + e, nz) << 4);
        wall_in[e] = wall_table != nullptr && code_yz[e] != 0;
        fu_in[e] = 1;
        fp_in[e] = 0;
        if (wall_in[e]) {
            fu_in[e] = wall_table[code_yz[e]];
            fp_in[e] = wall_table[ACFDTD_HIP_WALL_CODES + code_yz[e]];
        }
    }
    const bool yz_boundary = !has_ylo || !has_yhi || !has_zlo || !has_zhi;

    real mg = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) mg += a.mean[r] * a.gain[r];
    const real cdiv = (a.relax_scale - a.pres) - mg;

    bool special = false;
    for (int s = 0; s < a.n_sources; ++s) {
        const int e = a.source_cell[2 * s + 1] - jk0;
        special = special || (e >= 0 && e < V);
    }
    for (int r = 0; r < a.n_receivers; ++r) {
        const int e = a.receiver_cell[2 * r + 1] - jk0;
        special = special || (e >= 0 && e < V);
    }
    const TapePlanes planes(nx, ny, nz);
    const real *tape_row =
        want_walls ? a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride : nullptr;

    // New adjoint vx on face i0, the lower face of the first cell of the march.
    real ld_c[V], lvx_lo[V];
    vload<real, V>(ld_in + L.cell(i0), ld_c);
    {
        real v[V], prev[V];
        vload<real, V>(lvx_in + L.cell(i0), v);
        vload<real, V>(ld_in + L.cell(i0 > 0 ? i0 - 1 : i0), prev);
        const real fxf = LAYER ? a.face[0][i0] : real(1);
        real wxf = 0;
        if (sponge_face) wxf = g.w_face[0][i0];
        real after[V];
        if (sponge_face) vload<real, V>(io.after_vx + L.cell(i0), after);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real t = v[e] + ((i0 > 0 ? prev[e] : real(0)) - ld_c[e]);
            if (sponge_face) acc[kGradSponge] += mdt * (double)(wxf + wyc + wzc[e]) * t * after[e];
            lvx_lo[e] = damp3<real, LAYER>(t, fxf, fyc, fzc[e]);
        }
        vstore<real, V>(lvx_out + L.cell(i0), lvx_lo);
    }

    auto plane = [&](const int i, auto xwall_tag) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        const int64_t c = L.cell(i);
        const bool up_interior = !XWALL || i + 1 < nx;
        const bool lo_interior = !XWALL || i > 0;
        const real fxc = LAYER ? uniform_load(a.cell[0] + i) : real(1);
        const real fxf_hi = LAYER ? uniform_load(a.face[0] + i + 1) : real(1);

        // Loads, none behind a per-lane branch.
        real lvx_v[V], ld_up[V], lvy_lo_in[V], lvy_hi_in[V], ld_ylo[V], ld_yhi[V], lvz_v[V];
        real lpc[V];
        vload<real, V>(lvx_in + c + L.plane, lvx_v);
        vload<real, V>(ld_in + c + (up_interior ? L.plane : 0), ld_up);
        const int64_t cy = L.face_y(i);
        vload<real, V>(lvy_in + cy, lvy_lo_in);
        vload<real, V>(lvy_in + cy + nz, lvy_hi_in);
        vload<real, V>(ld_in + c - dy_lo, ld_ylo);
        vload<real, V>(ld_in + c + dy_hi, ld_yhi);
        const int64_t cz = L.face_z(i);
        vload<real, V>(lvz_in + cz, lvz_v);
        const real lvz_top = lvz_in[cz + V];
        const real ldz_lo = ld_in[c - dz_lo];
        const real ldz_hi = ld_in[c + dz_hi];
        vload<real, V>(lp + c, lpc);
        real states[NR > 0 ? NR : 1][V];
        if (local) {
#pragma unroll
            for (int r = 0; r < NR; ++r) vload<real, V>(lpsi + (int64_t)r * n_cells + c, states[r]);
        }
        // Boundary tape of step m - 1 for the lane's wall cells, read with the other loads so
        // that waiting for it does not also wait for this plane's stores. Every face plane a
        // cell lies on holds the same value: the x row on x-wall planes, else the y row for
        // the first and last rows, else the z rows for the first and last cells of a row.
        real tv[V];
#pragma unroll
        for (int e = 0; e < V; ++e) tv[e] = 0;
        if (want_walls && (XWALL || yz_boundary)) {
            {
                // Offsets, not a pointer that may be null: such a pointer is no longer known to
                // point to global memory, and the loads through it become flat ones (which the
                // LDS waits also wait for).
                int64_t row = -1;
                if (XWALL) row = (i == 0 ? planes.x_lo : planes.x_hi) + (int64_t)j * nz + k0;
                else if (!has_ylo || !has_yhi)
                    row = (!has_ylo ? planes.y_lo : planes.y_hi) + (int64_t)i * nz + k0;
                if (row >= 0) {
                    const real *src = tape_row + row;
                    if (a.wall_tape_rows16) {
                        vload<real, V>(src, tv);
                    } else {
#pragma unroll
                        for (int e = 0; e < V; ++e) tv[e] = src[e];
                    }
                } else {
                    const int64_t xy = (int64_t)i * ny + j;
                    if (!has_zhi) tv[V - 1] = tape_row[planes.z_hi + xy];
                    if (!has_zlo) tv[0] = tape_row[planes.z_lo + xy];
                }
            }
        }

        // New adjoint velocities on the six faces of each cell (adjoint_face_sweep).
        real tx[V], ty_lo[V], ty_hi[V], tz[V + 1];
        real lvx_hi[V], lvy_lo[V], lvy_hi[V], lvz_f[V + 1];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            tx[e] = lvx_v[e] + (ld_c[e] - (up_interior ? ld_up[e] : real(0)));
            lvx_hi[e] = damp3<real, LAYER>(tx[e], fxf_hi, fyc, fzc[e]);
            ty_lo[e] = lvy_lo_in[e] + ((has_ylo ? ld_ylo[e] : real(0)) - ld_c[e]);
            lvy_lo[e] = damp3<real, LAYER>(ty_lo[e], fxc, fyf_lo, fzc[e]);
            ty_hi[e] = lvy_hi_in[e] + (ld_c[e] - (has_yhi ? ld_yhi[e] : real(0)));
            lvy_hi[e] = damp3<real, LAYER>(ty_hi[e], fxc, fyf_hi, fzc[e]);
            const real below = e > 0 ? ld_c[e - 1] : (has_zlo ? ldz_lo : real(0));
            tz[e] = lvz_v[e] + (below - ld_c[e]);
            lvz_f[e] = damp3<real, LAYER>(tz[e], fxc, fyc, fzf[e]);
        }
        tz[V] = lvz_top + (ld_c[V - 1] - (has_zhi ? ldz_hi : real(0)));
        lvz_f[V] = damp3<real, LAYER>(tz[V], fxc, fyc, fzf[V]);

        // Owned faces: stores and their sponge gradient.
        const bool store_x = i + 1 < i1 || i + 1 == nx;
        if (store_x) vstore<real, V>(lvx_out + c + L.plane, lvx_hi);
        vstore<real, V>(lvy_out + cy, lvy_lo);
        if (!has_yhi) vstore<real, V>(lvy_out + cy + nz, lvy_hi);
        {
            real lower[V];
#pragma unroll
            for (int e = 0; e < V; ++e) lower[e] = lvz_f[e];
            vstore<real, V>(lvz_out + cz, lower);
        }
        if (!has_zhi) lvz_out[cz + V] = lvz_f[V];
        if (sponge_face) {
            const real wxc = uniform_load(g.w_cell[0] + i), wxf_hi = uniform_load(g.w_face[0] + i + 1);
            real ax[V], ay_lo[V], ay_hi[V], az[V];
            vload<real, V>(io.after_vx + c + L.plane, ax);
            vload<real, V>(io.after_vy + cy, ay_lo);
            vload<real, V>(io.after_vy + cy + nz, ay_hi);
            vload<real, V>(io.after_vz + cz, az);
            const real az_top = io.after_vz[cz + V];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                double s = (double)(wxc + wyf_lo + wzc[e]) * ty_lo[e] * ay_lo[e];
                s += (double)(wxc + wyc + wzf[e]) * tz[e] * az[e];
                if (store_x) s += (double)(wxf_hi + wyc + wzc[e]) * tx[e] * ax[e];
                if (!has_yhi) s += (double)(wxc + wyf_hi + wzc[e]) * ty_hi[e] * ay_hi[e];
                acc[kGradSponge] += mdt * s;
            }
            if (!has_zhi)
                acc[kGradSponge] +=
                    mdt * (double)(wxc + wyc + wzf[V]) * tz[V] * az_top;
        }

        // Stencil part of step m (adjoint_cell_sweep, stencil = 1).
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real lo_x = lo_interior ? lvx_lo[e] : real(0);
            const real up_x = up_interior ? lvx_hi[e] : real(0);
            const real lo_y = has_ylo ? lvy_lo[e] : real(0);
            const real up_y = has_yhi ? lvy_hi[e] : real(0);
            const real lo_z = (e > 0 || has_zlo) ? lvz_f[e] : real(0);
            const real up_z = (e + 1 < V || has_zhi) ? lvz_f[e + 1] : real(0);
            lpc[e] = lpc[e] + vel * (((up_x - lo_x) + (up_y - lo_y)) + (up_z - lo_z));
        }

        if (!local) {
            vstore<real, V>(lp + c, lpc);
        } else {
            // Local part of step m - 1 (adjoint_cell_sweep, local = 1).
            real injected[V];
#pragma unroll
            for (int e = 0; e < V; ++e) injected[e] = 0;
            if (special) {
                for (int r = 0; r < a.n_receivers; ++r) {
                    const int e = a.receiver_cell[2 * r + 1] - jk0;
                    if (a.receiver_cell[2 * r] == i && e >= 0 && e < V) {
                        lane_add<real, V>(lpc, e,
                                          g.grad_recording[r * g.grad_recording_stride +
                                                           (step - g.grad_recording_step0)]);
                    }
                }
                for (int s = 0; s < a.n_sources; ++s) {
                    const int e = a.source_cell[2 * s + 1] - jk0;
                    if (a.source_cell[2 * s] == i && e >= 0 && e < V &&
                        step < a.source_length[s]) {
                        if (g.grad_signal)
                            g.grad_signal[s * a.source_stride + step] =
                                a.src * lane_get<real, V>(lpc, e);
                        if (sponge_cell)
                            lane_add<real, V>(injected, e,
                                              a.src * a.source_signal[s * a.source_stride + step]);
                    }
                }
            }
            real bp[V], bvx_lo[V], bvx_hi[V], bvy_lo[V], bvy_hi[V], bvz[V], dvt[V];
            real bvz_top = 0;
            if (sponge_cell) vload<real, V>(io.before_p + c, bp);
            if (air && div_row) {
                vload<real, V>(div_row + c, dvt);
            } else if (air) {
                vload<real, V>(io.before_vx + c, bvx_lo);
                vload<real, V>(io.before_vx + c + L.plane, bvx_hi);
                vload<real, V>(io.before_vy + cy, bvy_lo);
                vload<real, V>(io.before_vy + cy + nz, bvy_hi);
                vload<real, V>(io.before_vz + cz, bvz);
                bvz_top = io.before_vz[cz + V];
            }
            const int cx = XWALL ? wall_class(i, nx) : 0;
            real wxc = 0;
            if (sponge_cell) wxc = uniform_load(g.w_cell[0] + i);
            real lp_out[V], ld_new[V];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const real lpe = lpc[e];
                const real lhat = LAYER ? ((lpe * fxc) * fyc) * fzc[e] : lpe;
                if (sponge_cell)
                    acc[kGradSponge] += mdt * (double)(wxc + (wyc + wzc[e])) * (double)lpe *
                                        (double)(bp[e] - injected[e]);
                real ltil = lhat, lp_new = 0;
                bool is_wall = wall_in[e];
                real fu = fu_in[e], fp = fp_in[e];
                const int code = code_yz[e] | cx;
                if (XWALL) {
                    is_wall = wall_table != nullptr && code != 0;
                    if (is_wall) {
                        fu = wall_table[code];
                        fp = wall_table[ACFDTD_HIP_WALL_CODES + code];
                    }
                }
                if (is_wall) {
                    ltil = fu * lhat;
                    lp_new = -fp * lhat;
                    if (want_walls && (XWALL || yz_boundary)) {
                        // Added to each face the cell lies on, in face order.
                        const double ga = -(double)lhat * (double)fu * (double)fu * (double)tv[e];
                        if (XWALL && i == 0) acc[kGradWall + 0] += ga;
                        if (XWALL && i == nx - 1) acc[kGradWall + 1] += ga;
                        if (!has_ylo) acc[kGradWall + 2] += ga;
                        if (!has_yhi) acc[kGradWall + 3] += ga;
                        if (e == 0 && !has_zlo) acc[kGradWall + 4] += ga;
                        if (e == V - 1 && !has_zhi) acc[kGradWall + 5] += ga;
                    }
                }
                lp_new = lp_new + ltil;
                real ldv = ltil * cdiv;
                double div = 0;
                if (air) {
                    if (div_row) {
                        div = (double)dvt[e];
                    } else {
                        const real top = e + 1 < V ? bvz[e + 1] : bvz_top;
                        div = (double)(((bvx_hi[e] - bvx_lo[e]) + (bvy_hi[e] - bvy_lo[e])) +
                                       (top - bvz[e]));
                    }
                    acc[kGradScale] += div * (double)ltil;
                }
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const real old = states[r][e];
                    states[r][e] = a.mean[r] * ltil + a.decay[r] * old;
                    ldv = ldv + old * (a.gain[r] * (real(1) - a.decay[r]));
                    if (air)
                        acc[kGradGain + r] += div * (-(double)a.mean[r] * (double)ltil +
                                                     (1.0 - (double)a.decay[r]) * (double)old);
                }
                lp_out[e] = lp_new;
                ld_new[e] = ldv;
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) vstore<real, V>(lpsi + (int64_t)r * n_cells + c, states[r]);
            vstore<real, V>(lp + c, lp_out);
            vstore<real, V>(ld_out + c, ld_new);
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            lvx_lo[e] = lvx_hi[e];
            ld_c[e] = ld_up[e];
        }
    };

    using Wall = std::integral_constant<bool, true>;
    using Interior = std::integral_constant<bool, false>;
    int i = i0;
    if (i == 0) {
        plane(0, Wall{});
        ++i;
    }
    const int stop = i1 == nx ? nx - 1 : i1;
#pragma unroll kXUnroll
    for (; i < stop; ++i) plane(i, Interior{});
    if (i1 == nx && i == nx - 1) plane(nx - 1, Wall{});
}

template <typename real, int V, int NR, bool LAYER, bool FWDGRAD>
__global__ __launch_bounds__(512) void adjoint_fused_step(const SweepArgs<real> a,
                                                          const AdjointArgs<real> g,
                                                          const AdjointIO<real> io,
                                                          const int64_t m) {
    double acc[kGradSlots] = {};
    int kb = 0, jb = 0, xc = 0, slot = 0, j = 0, k0 = 0;
    const bool tile = decode_tile_slot(a.tiles, kb, jb, xc, slot);
    const bool active = tile && flat_lane(a.tiles, a.ny, a.nz, jb, j, k0);
    const bool local = m >= 1;
    const bool want_walls = local && (g.grads & ACFDTD_HIP_GRAD_WALLS) && a.wall_tape;
    const bool sponge_face = FWDGRAD && LAYER && (g.grads & ACFDTD_HIP_GRAD_SPONGE) && io.after_vx;
    const bool sponge_cell =
        FWDGRAD && local && LAYER && (g.grads & ACFDTD_HIP_GRAD_SPONGE) && io.before_p;
    const bool air = FWDGRAD && local && NR > 0 && (g.grads & ACFDTD_HIP_GRAD_AIR) &&
                     (io.before_vx || a.div_tape);
    if (active) {
        const int i0 = xc * a.tiles.chunk;
        const int i1 = min(i0 + a.tiles.chunk, a.nx);
        adjoint_march_vec<real, V, NR, LAYER, FWDGRAD>(
            a, g, io, i0, i1, j, k0, m, want_walls, sponge_face, sponge_cell, air, io.ldiv_in,
            io.lvx_in, io.lvy_in, io.lvz_in, io.ldiv_out, io.lvx_out, io.lvy_out, io.lvz_out,
            g.lp, g.lpsi, a.wall_table, acc);
    }
    if (FWDGRAD) {
        if (want_walls || sponge_face || sponge_cell || air)
            flush_partials(acc, tile, slot, g.partials);
    } else if (want_walls) {
        flush_partials<kGradWall, 6>(acc, tile, slot, g.partials);
    }
}

template <typename real>
using AdjointFusedKernel = void (*)(const SweepArgs<real>, const AdjointArgs<real>,
                                    const AdjointIO<real>, const int64_t);

template <typename real, int V, bool F>
AdjointFusedKernel<real> pick_adjoint_vf(int nr, bool layer) {
    switch (nr) {
        case 0: return layer ? adjoint_fused_step<real, V, 0, true, F> : adjoint_fused_step<real, V, 0, false, F>;
        case 1: return layer ? adjoint_fused_step<real, V, 1, true, F> : adjoint_fused_step<real, V, 1, false, F>;
        case 2: return layer ? adjoint_fused_step<real, V, 2, true, F> : adjoint_fused_step<real, V, 2, false, F>;
        case 3: return layer ? adjoint_fused_step<real, V, 3, true, F> : adjoint_fused_step<real, V, 3, false, F>;
        default: return layer ? adjoint_fused_step<real, V, 4, true, F> : adjoint_fused_step<real, V, 4, false, F>;
    }
}

// FWDGRAD kernels carry the sponge and air gradients (forward states read every step); the
// others only the wall and source gradients, with fewer registers.
template <typename real, int V>
AdjointFusedKernel<real> pick_adjoint_v(int nr, bool layer, bool fwdgrad) {
    return fwdgrad ? pick_adjoint_vf<real, V, true>(nr, layer)
                   : pick_adjoint_vf<real, V, false>(nr, layer);
}

template <typename real>
AdjointFusedKernel<real> pick_adjoint(int vec, int nr, bool layer, bool fwdgrad) {
    if constexpr (sizeof(real) == 4) {
        if (vec == 4) return pick_adjoint_v<real, 4>(nr, layer, fwdgrad);
    }
    if (vec == 2) return pick_adjoint_v<real, 2>(nr, layer, fwdgrad);
    return pick_adjoint_v<real, 1>(nr, layer, fwdgrad);
}

template <typename real>
int64_t launch_adjoint_pairs(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &gd,
                             int64_t step_hi, int64_t n_steps, hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    a.tiles = fused_tiling(d);
    AdjointArgs<real> g;
    fill_adjoint(gd, g);
    const bool air = (gd.grads & ACFDTD_HIP_GRAD_AIR) && d.n_relax > 0;
    const bool fwdgrad = (gd.fwd_count > 0 &&
                          (((gd.grads & ACFDTD_HIP_GRAD_SPONGE) && d.has_layer) || air)) ||
                         (air && d.div_tape);
    const AdjointFusedKernel<real> k =
        pick_adjoint<real>(a.tiles.vec, d.n_relax, d.has_layer != 0, fwdgrad);
    real *first[4] = {static_cast<real *>(gd.ldiv), static_cast<real *>(gd.lvx),
                      static_cast<real *>(gd.lvy), static_cast<real *>(gd.lvz)};
    real *second[4] = {static_cast<real *>(gd.ldiv_swap), static_cast<real *>(gd.lvx_swap),
                       static_cast<real *>(gd.lvy_swap), static_cast<real *>(gd.lvz_swap)};
    const bool ring = gd.fwd_count > 0;
    const dim3 grid(a.tiles.launched), block(64, a.tiles.tile_j);
    int64_t n = 0;
    for (; n + 1 < n_steps; n += 2) {
        for (int half = 0; half < 2; ++half) {
            const int64_t m = step_hi - n - half;
            real **in = half ? second : first;
            real **out = half ? first : second;
            AdjointIO<real> io{};
            io.ldiv_in = in[0];
            io.lvx_in = in[1];
            io.lvy_in = in[2];
            io.lvz_in = in[3];
            io.ldiv_out = out[0];
            io.lvx_out = out[1];
            io.lvy_out = out[2];
            io.lvz_out = out[3];
            if (ring) {
                const ForwardState<real> after = forward_slot<real>(d, gd, m + 1 - gd.fwd_step0);
                const ForwardState<real> before = forward_slot<real>(d, gd, m - gd.fwd_step0);
                io.after_vx = after.vx;
                io.after_vy = after.vy;
                io.after_vz = after.vz;
                io.before_p = before.p;
                io.before_vx = before.vx;
                io.before_vy = before.vy;
                io.before_vz = before.vz;
            }
            hipLaunchKernelGGL(k, grid, block, 0, stream, a, g, io, m);
        }
    }
    return n;
}

}  // namespace acfdtd
