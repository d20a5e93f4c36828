// Two adjoint steps per sweep: adjoint steps m and m - 1 (the faces and stencil of each, and
// the local parts of steps m - 1 and m - 2) in one pass over memory. Templates only;
// adjoint2.hip and adjoint2_64.hip instantiate them for fp32 (4 cells per lane) and fp64 (2).
//
// The adjoint step has the forward step's data flow: it reads ldiv on the six neighbours of a
// cell (as the forward step reads p) and the adjoint velocities on the cell's own faces, and
// lp and lpsi only at the cell. So the sweep is the two-step forward sweep (marches2.hpp) with
// the adjoint body: along x a lane runs step A (adjoint step m) one plane ahead of step B
// (adjoint step m - 1); step B of plane i needs step A's ldiv of planes i - 1, i, i + 1
// (registers of the same lane) and of its four y and z neighbours in plane i (other lanes,
// through LDS); the workgroup runs step A, storing nothing and adding no gradient, on one halo
// row on each side, on a halo lane group at each end of a row segment, and on one halo plane
// at each end of its chunk. A halo cell's step A includes its receiver's injection, since the
// ldiv it hands to step B depends on it. Halo lanes read lp and lpsi of cells other
// workgroups own, so every field ping-pongs between two sets (the single-step adjoint updates
// lp and lpsi in place).
//
// Every cell's arithmetic is adjoint_march_vec's term by term, so the adjoint fields and the
// source-signal gradient are bit-identical to single adjoint steps; the fp64 gradient partials
// are summed in another order. The wall and signal gradients are carried, and with AIR the
// air-strength gradients, whose forward input (every cell's divergence of every step) comes
// from the forward pass's divergence tape; with SPONGE the sponge-strength gradient, whose
// forward input (the damped fields of the layer's cells at every step) comes from the forward
// pass's layer tape.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "acfdtd_hip.h"
#include "adjoint.hpp"
#include "marches.hpp"
#include "marches2.hpp"
#include "sweeps.hpp"

namespace acfdtd {

template <typename real>
struct Adjoint2IO {
    const real *ld_in, *lvx_in, *lvy_in, *lvz_in, *lp_in, *lpsi_in;
    real *ld_out, *lvx_out, *lvy_out, *lvz_out, *lp_out, *lpsi_out;
};

// LDS layout of the block's grading weights (SPONGE): z cell weights, z face weights, y cell
// weights, y face weights. They are read where the sponge terms need them rather than held in
// registers through the march (the two-step adjoint has none to spare, and an LDS read does
// not wait for the vector memory operations in flight).
template <typename real, int V>
struct WeightLds {
    static constexpr int kZ = kFused2Threads * V + 8;
    static constexpr int kY = 2 * 8 + 4;  // rows of a block (at most 16) and its halo rows
    static constexpr int zc = 0, zf = kZ, yc = 2 * kZ, yf = 2 * kZ + kY, size = 2 * kZ + 2 * kY;
};

// Per-lane constants, as adjoint_march_vec computes them before its march.
template <typename real, int V>
struct AdjLane {
    const int32_t *src_cell, *rx_cell;  // LDS copies of the tables
    const int64_t *src_len;
    const real *rx_now;  // dL/d(recording) of the sweep's two local steps, two per receiver
    const real *wtab;    // LDS copy of the wall table, or null without walls
    int ev_lo, ev_hi;    // planes of the lane's sources and receivers (empty: lo > hi)
    int j, k0, jk0;
    bool has_ylo, has_yhi, has_zlo, has_zhi, yz_boundary, special;
    int64_t dy_lo, dy_hi, dz_lo, dz_hi;
    real fyc, fyf_lo, fyf_hi, fzc[V], fzf[V + 1];
    int code_yz[V];
    bool wall_in[V];
    real fu_in[V], fp_in[V];
    real cdiv;
    // SPONGE: the grading weights of the lane's rows and cells (in the block's LDS copy, see
    // WeightLds), and its place in the frame planes of the layer tape (-1 outside the frame).
    const real *wl;  // LDS weights of the block
    int wy, wz;      // the lane's row and first cell in them
    int64_t lt_pos;
    __device__ real w_yc() const { return wl[WeightLds<real, V>::yc + wy]; }
    __device__ real w_yf(const int h) const { return wl[WeightLds<real, V>::yf + wy + h]; }
    __device__ real w_zc(const int e) const { return wl[WeightLds<real, V>::zc + wz + e]; }
    __device__ real w_zf(const int e) const { return wl[WeightLds<real, V>::zf + wz + e]; }
};


template <typename real, int V>
__device__ __forceinline__ void adj_lane_sponge(const SweepArgs<real> &a, const real *wl,
                                                const int wy, const int wz, const int j,
                                                const int k0, AdjLane<real, V> &c) {
    c.wl = wl;
    c.wy = wy;
    c.wz = wz;
    c.lt_pos = a.layer.in_frame(j, k0);
}

template <typename real, int V, int NR, bool LAYER>
__device__ __forceinline__ void adj_lane(const SweepArgs<real> &a, const int j, const int k0,
                                         const int32_t *src_cell, const int64_t *src_len,
                                         const int32_t *rx_cell, const real *rx_now,
                                         const real *wtab, AdjLane<real, V> &c) {
    const int ny = a.ny, nz = a.nz;
    c.src_cell = src_cell;
    c.src_len = src_len;
    c.rx_cell = rx_cell;
    c.rx_now = rx_now;
    c.wtab = wtab;
    c.j = j;
    c.k0 = k0;
    c.jk0 = j * nz + k0;
    c.has_ylo = j > 0;
    c.has_yhi = j + 1 < ny;
    c.has_zlo = k0 > 0;
    c.has_zhi = k0 + V < nz;
    c.dy_lo = c.has_ylo ? nz : 0;
    c.dy_hi = c.has_yhi ? nz : 0;
    c.dz_lo = c.has_zlo ? 1 : 0;
    c.dz_hi = c.has_zhi ? V : V - 1;
    c.fyc = 1;
    c.fyf_lo = 1;
    c.fyf_hi = 1;
#pragma unroll
    for (int e = 0; e < V; ++e) c.fzc[e] = 1;
#pragma unroll
    for (int e = 0; e <= V; ++e) c.fzf[e] = 1;
    if (LAYER) {
        c.fyc = a.cell[1][j];
        c.fyf_lo = a.face[1][j];
        c.fyf_hi = a.face[1][j + 1];
#pragma unroll
        for (int e = 0; e < V; ++e) c.fzc[e] = a.cell[2][k0 + e];
#pragma unroll
        for (int e = 0; e <= V; ++e) c.fzf[e] = a.face[2][k0 + e];
    }
    // A wall cell is one with a nonzero class when the table exists (adjoint_cell_sweep).
#pragma unroll
    for (int e = 0; e < V; ++e) {
        c.code_yz[e] = (wall_class(j, ny) << 2) | (wall_class(k0 + e, nz) << 4);
        c.wall_in[e] = wtab != nullptr && c.code_yz[e] != 0;
        c.fu_in[e] = 1;
        c.fp_in[e] = 0;
        if (c.wall_in[e]) {
            c.fu_in[e] = wtab[c.code_yz[e]];
            c.fp_in[e] = wtab[ACFDTD_HIP_WALL_CODES + c.code_yz[e]];
        }
    }
    c.yz_boundary = !c.has_ylo || !c.has_yhi || !c.has_zlo || !c.has_zhi;
    real mg = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) mg += a.mean[r] * a.gain[r];
    c.cdiv = (a.relax_scale - a.pres) - mg;
    int lo = a.nx, hi = -1;
    for (int s = 0; s < a.n_sources; ++s) {
        const int e = src_cell[2 * s + 1] - c.jk0;
        if (e >= 0 && e < V) {
            lo = min(lo, src_cell[2 * s]);
            hi = max(hi, src_cell[2 * s]);
        }
    }
    for (int r = 0; r < a.n_receivers; ++r) {
        const int e = rx_cell[2 * r + 1] - c.jk0;
        if (e >= 0 && e < V) {
            lo = min(lo, rx_cell[2 * r]);
            hi = max(hi, rx_cell[2 * r]);
        }
    }
    c.ev_lo = lo;
    c.ev_hi = hi;
    c.special = hi >= 0;
}

// What one adjoint step leaves in a lane for one plane: the new adjoint velocities on the faces
// the next step reads again (upper x, both y, all z), lp, ldiv and lpsi after the local part;
// with the air gradients, the divergences step B of the plane needs (loaded with step A's
// inputs, so that step B, which otherwise reads only registers and LDS, waits for no load).
template <typename real, int V, int NR>
struct AdjPlane {
    real lvx_hi[V], lvy_lo[V], lvy_hi[V], lvz[V + 1], lp[V], ld[V];
    real psi[NR > 0 ? NR : 1][V];
    real dv_b[V];
};

// One adjoint step of the lane's V cells in plane i, adjoint_march_vec's plane body without the
// sponge and air gradients: the faces from ldiv of the cell (ld_c) and its neighbours (ld_up:
// plane i + 1, ld_ylo / ld_yhi: rows j -/+ 1, ldz_lo / ldz_hi: cells k0 - 1, k0 + V) and the
// old adjoint velocities, with the new lower x face lvx_lo; the stencil into lp; the local part
// of `step` (receiver injection rx_now[2 r + h], the source-signal gradient, sponge, walls,
// relaxation). `own` marks the cells whose gradients this lane adds and writes; tv holds the
// boundary tape of `step` for the lane's wall cells, dv (AIR) the forward divergences of `step`,
// whose air-gradient terms go to air_acc (relax_scale, then one per relaxation state).
// SPONGE: the sponge-strength gradient of `own` cells into sp_acc, from the layer tape rows of
// step + 1 (lt_v: the velocities the faces' terms multiply) and of `step` (lt_p: the pressures
// of the local part); own_xhi: this lane also adds the term of its cells' upper x face.
template <typename real, int V, int NR, bool LAYER, bool XWALL, bool AIR = false, bool SPONGE = false>
__device__ __forceinline__ void adjoint_cells(
    const SweepArgs<real> &a, const AdjLane<real, V> &c, const int i, const int64_t step,
    const int h, const bool own, const bool want_walls, real *grad_signal, const real (&tv)[V],
    const real (&ld_c)[V], const real (&ld_up)[V], const real (&ld_ylo)[V],
    const real (&ld_yhi)[V], const real ldz_lo, const real ldz_hi, const real (&lvx_lo)[V],
    const real (&lvx_v)[V], const real (&lvy_lo_in)[V], const real (&lvy_hi_in)[V],
    const real (&lvz_v)[V], const real lvz_top, const real (&lp_in)[V],
    const real (&states_in)[NR > 0 ? NR : 1][V], AdjPlane<real, V, NR> &o, double (&acc)[6],
    const real (&dv)[V], double (&air_acc)[1 + (NR > 0 ? NR : 1)],
    const AdjointArgs<real> &g, const real *lt_v, const real *lt_p, const bool own_xhi,
    double &sp_acc) {
    const int nx = a.nx;
    const real vel = a.vel;
    const bool up_interior = !XWALL || i + 1 < nx;
    const bool lo_interior = !XWALL || i > 0;
    const real fxc = LAYER ? uniform_load(a.cell[0] + i) : real(1);
    const real fxf_hi = LAYER ? uniform_load(a.face[0] + i + 1) : real(1);

    // Faces (adjoint_face_sweep); SPONGE keeps the undamped values for the sponge terms.
    real s_tx[V], s_ty_lo[V], s_ty_hi[V], s_tz[V + 1];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const real tx = lvx_v[e] + (ld_c[e] - (up_interior ? ld_up[e] : real(0)));
        o.lvx_hi[e] = damp3<real, LAYER>(tx, fxf_hi, c.fyc, c.fzc[e]);
        const real ty_lo = lvy_lo_in[e] + ((c.has_ylo ? ld_ylo[e] : real(0)) - ld_c[e]);
        o.lvy_lo[e] = damp3<real, LAYER>(ty_lo, fxc, c.fyf_lo, c.fzc[e]);
        const real ty_hi = lvy_hi_in[e] + (ld_c[e] - (c.has_yhi ? ld_yhi[e] : real(0)));
        o.lvy_hi[e] = damp3<real, LAYER>(ty_hi, fxc, c.fyf_hi, c.fzc[e]);
        const real below = e > 0 ? ld_c[e - 1] : (c.has_zlo ? ldz_lo : real(0));
        const real tz = lvz_v[e] + (below - ld_c[e]);
        o.lvz[e] = damp3<real, LAYER>(tz, fxc, c.fyc, c.fzf[e]);
        s_tx[e] = tx;
        s_ty_lo[e] = ty_lo;
        s_ty_hi[e] = ty_hi;
        s_tz[e] = tz;
    }
    {
        const real tz = lvz_top + (ld_c[V - 1] - (c.has_zhi ? ldz_hi : real(0)));
        o.lvz[V] = damp3<real, LAYER>(tz, fxc, c.fyc, c.fzf[V]);
        s_tz[V] = tz;
    }

    // Sponge terms of the faces, as adjoint_march_vec adds them: the undamped adjoint velocity
    // times the forward velocity after the step (tape row lt_v) times -dt times the face's
    // summed weights.
    real wxc = 0;
    int64_t lt_q = -1;
    const double mdt = -(double)g.dt;
    if constexpr (SPONGE) {
        const LayerFrame &f = a.layer;
        wxc = uniform_load(g.w_cell[0] + i);
        lt_q = f.offset(i, c.j, c.k0, c.lt_pos);
        if (own) {
            double s = 0;
            if (lt_q >= 0) {
                real fy[V], fz[V];
                vload<real, V>(lt_v + 2 * f.cells + lt_q, fy);
                vload<real, V>(lt_v + 3 * f.cells + lt_q, fz);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    s += (double)(wxc + c.w_yf(0) + c.w_zc(e)) * s_ty_lo[e] * fy[e];
                    s += (double)(wxc + c.w_yc() + c.w_zf(e)) * s_tz[e] * fz[e];
                }
            }
            if (!c.has_yhi) {
                real fy[V];
                vload<real, V>(lt_v + f.far_y() + (int64_t)i * a.nz + c.k0, fy);
#pragma unroll
                for (int e = 0; e < V; ++e) s += (double)(wxc + c.w_yf(1) + c.w_zc(e)) * s_ty_hi[e] * fy[e];
            }
            if (!c.has_zhi)
                s += (double)(wxc + c.w_yc() + c.w_zf(V)) * s_tz[V] *
                     lt_v[f.far_z() + (int64_t)i * a.ny + c.j];
            sp_acc += mdt * s;
        }
        if (own_xhi) {
            // The upper x face is the lower face of plane i + 1 (the far face of the last plane).
            const int64_t q1 = i + 1 == nx ? f.far_x() + (int64_t)c.j * a.nz + c.k0
                                           : f.cells + f.offset(i + 1, c.j, c.k0, c.lt_pos);
            if (i + 1 == nx || q1 >= f.cells) {
                const real wxf = uniform_load(g.w_face[0] + i + 1);
                real fx[V];
                vload<real, V>(lt_v + q1, fx);
                double s = 0;
#pragma unroll
                for (int e = 0; e < V; ++e) s += (double)(wxf + c.w_yc() + c.w_zc(e)) * s_tx[e] * fx[e];
                sp_acc += mdt * s;
            }
        }
    }

    // Stencil (adjoint_cell_sweep, stencil = 1).
    real lpc[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const real lo_x = lo_interior ? lvx_lo[e] : real(0);
        const real up_x = up_interior ? o.lvx_hi[e] : real(0);
        const real lo_y = c.has_ylo ? o.lvy_lo[e] : real(0);
        const real up_y = c.has_yhi ? o.lvy_hi[e] : real(0);
        const real lo_z = (e > 0 || c.has_zlo) ? o.lvz[e] : real(0);
        const real up_z = (e + 1 < V || c.has_zhi) ? o.lvz[e + 1] : real(0);
        lpc[e] = lp_in[e] + vel * (((up_x - lo_x) + (up_y - lo_y)) + (up_z - lo_z));
    }

    // Local part of `step` (adjoint_cell_sweep, local = 1): receivers in, signal gradient out;
    // SPONGE: the sources' injection of `step`, which the tape's pressures include and the
    // sponge term excludes.
    real injected[V];
#pragma unroll
    for (int e = 0; e < V; ++e) injected[e] = 0;
    if (c.special && i >= c.ev_lo && i <= c.ev_hi) {
        if (SPONGE && own) {
            for (int s = 0; s < a.n_sources; ++s) {
                const int e = c.src_cell[2 * s + 1] - c.jk0;
                if (c.src_cell[2 * s] == i && e >= 0 && e < V && step < c.src_len[s])
                    lane_add<real, V>(injected, e, a.src * a.source_signal[s * a.source_stride + step]);
            }
        }
        for (int r = 0; r < a.n_receivers; ++r) {
            const int e = c.rx_cell[2 * r + 1] - c.jk0;
            if (c.rx_cell[2 * r] == i && e >= 0 && e < V) lane_add<real, V>(lpc, e, c.rx_now[2 * r + h]);
        }
        if (own && grad_signal) {
            for (int s = 0; s < a.n_sources; ++s) {
                const int e = c.src_cell[2 * s + 1] - c.jk0;
                if (c.src_cell[2 * s] == i && e >= 0 && e < V && step < c.src_len[s])
                    grad_signal[s * a.source_stride + step] = a.src * lane_get<real, V>(lpc, e);
            }
        }
    }
    const int cx = XWALL ? wall_class(i, nx) : 0;
    // Sponge term of each cell: lp times the forward pressure of `step` (tape row lt_p) less
    // its injection, times -dt times the cell's summed weights.
    real bp[V];
#pragma unroll
    for (int e = 0; e < V; ++e) bp[e] = 0;
    if (SPONGE && own && lt_q >= 0) vload<real, V>(lt_p + lt_q, bp);
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const real lpe = lpc[e];
        if (SPONGE && own && lt_q >= 0)
            sp_acc += mdt * (double)(wxc + (c.w_yc() + c.w_zc(e))) * (double)lpe *
                      (double)(bp[e] - injected[e]);
        const real lhat = LAYER ? ((lpe * fxc) * c.fyc) * c.fzc[e] : lpe;
        real ltil = lhat, lp_new = 0;
        bool is_wall = c.wall_in[e];
        real fu = c.fu_in[e], fp = c.fp_in[e];
        if (XWALL) {
            const int code = c.code_yz[e] | cx;
            is_wall = c.wtab != nullptr;
            if (is_wall) {
                fu = c.wtab[code];
                fp = c.wtab[ACFDTD_HIP_WALL_CODES + code];
            }
        }
        if (is_wall) {
            ltil = fu * lhat;
            lp_new = -fp * lhat;
            if (own && want_walls && (XWALL || c.yz_boundary)) {
                // Added to each face the cell lies on, in face order.
                const double ga = -(double)lhat * (double)fu * (double)fu * (double)tv[e];
                if (XWALL && i == 0) acc[0] += ga;
                if (XWALL && i == nx - 1) acc[1] += ga;
                if (!c.has_ylo) acc[2] += ga;
                if (!c.has_yhi) acc[3] += ga;
                if (e == 0 && !c.has_zlo) acc[4] += ga;
                if (e == V - 1 && !c.has_zhi) acc[5] += ga;
            }
        }
        lp_new = lp_new + ltil;
        real ldv = ltil * c.cdiv;
        const double div = AIR ? (double)dv[e] : 0.0;
        if (AIR && own) air_acc[0] += div * (double)ltil;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const real old = states_in[r][e];
            o.psi[r][e] = a.mean[r] * ltil + a.decay[r] * old;
            ldv = ldv + old * (a.gain[r] * (real(1) - a.decay[r]));
            if (AIR && own)
                air_acc[1 + r] += div * (-(double)a.mean[r] * (double)ltil +
                                         (1.0 - (double)a.decay[r]) * (double)old);
        }
        o.lp[e] = lp_new;
        o.ld[e] = ldv;
    }
}

// Boundary tape of `step` for the lane's wall cells in plane i: every face plane a cell lies on
// holds the same value, so the x row on x-wall planes, else the y row for the first and last
// rows, else the z rows for the first and last cells of a row.
template <typename real, int V, bool XWALL>
__device__ __forceinline__ void load_tape(const SweepArgs<real> &a, const AdjLane<real, V> &c,
                                          const int i, const int64_t step, real (&tv)[V]) {
#pragma unroll
    for (int e = 0; e < V; ++e) tv[e] = 0;
    if (!(XWALL || c.yz_boundary)) return;
    const TapePlanes planes(a.nx, a.ny, a.nz);
    const real *row = a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride;
    const real *src = nullptr;
    if (XWALL) src = row + (i == 0 ? planes.x_lo : planes.x_hi) + (int64_t)c.j * a.nz + c.k0;
    else if (!c.has_ylo || !c.has_yhi)
        src = row + (!c.has_ylo ? planes.y_lo : planes.y_hi) + (int64_t)i * a.nz + c.k0;
    if (src != nullptr) {
        if (a.wall_tape_rows16) {
            vload<real, V>(src, tv);
        } else {
#pragma unroll
            for (int e = 0; e < V; ++e) tv[e] = src[e];
        }
    } else {
        const int64_t xy = (int64_t)i * a.ny + c.j;
        if (!c.has_zhi) tv[V - 1] = row[planes.z_hi + xy];
        if (!c.has_zlo) tv[0] = row[planes.z_lo + xy];
    }
}

template <typename real, int V, int NR, bool LAYER, bool AIR = false, bool SPONGE = false>
__device__ __forceinline__ void adjoint2_march(
    const SweepArgs<real> &a, const Tiling2 &t2, const int i0, const int i1, const int j,
    const int k0, const bool valid, const bool owner, const int t, const int64_t m,
    const bool want_walls, real *grad_signal, const Adjoint2IO<real> &io, real *__restrict__ lds,
    const int32_t *src_cell, const int64_t *src_len, const int32_t *rx_cell, const real *rx_now,
    const real *wtab, double (&acc)[6], double (&air_acc)[1 + (NR > 0 ? NR : 1)],
    const AdjointArgs<real> &g, double &sp_acc, const real *wl, const int wy, const int wz) {
    static_assert(!SPONGE || LAYER, "the sponge gradient needs the layer");
    using Wall = std::integral_constant<bool, true>;
    using Interior = std::integral_constant<bool, false>;
    const int nx = a.nx, nz = a.nz;
    const Layout L(nx, a.ny, nz, j, k0);
    const int64_t n_cells = L.plane * nx;
    const int G = t2.pitch, T = t2.threads;
    const real *__restrict__ ld_in = io.ld_in;
    const real *__restrict__ lvx_in = io.lvx_in;
    const real *__restrict__ lvy_in = io.lvy_in;
    const real *__restrict__ lvz_in = io.lvz_in;
    const real *__restrict__ lp_in = io.lp_in;
    const real *__restrict__ lpsi_in = io.lpsi_in;
    AdjLane<real, V> c;
    if (valid) adj_lane<real, V, NR, LAYER>(a, j, k0, src_cell, src_len, rx_cell, rx_now, wtab, c);
    if (SPONGE && valid) adj_lane_sponge<real, V>(a, wl, wy, wz, j, k0, c);
    // The lane constants' loads complete on every path into the march (see fused2_march).
    wait_memory();
    // SPONGE: layer tape rows of forward steps m, m - 1, m - 2 (step A's faces read row m and
    // its local part row m - 1; step B's faces row m - 1 and its local part row m - 2).
    const real *lt_row[3] = {nullptr, nullptr, nullptr};
    if (SPONGE) {
#pragma unroll
        for (int h = 0; h < 3; ++h)
            lt_row[h] = a.layer_tape + (m - h - a.layer_tape_step0) * a.layer_tape_stride;
    }
    // The sponge term of a lower x face from its undamped adjoint velocity w (first faces).
    auto sponge_xface = [&](const int s, const real *row, const real (&w)[V]) __attribute__((always_inline)) {
        if constexpr (SPONGE) {
            const LayerFrame &f = a.layer;
            const int64_t q = f.offset(s, j, k0, c.lt_pos);
            if (q >= 0) {
                const real wxf = uniform_load(g.w_face[0] + s);
                real fx[V];
                vload<real, V>(row + f.cells + q, fx);
                double sum = 0;
#pragma unroll
                for (int e = 0; e < V; ++e) sum += (double)(wxf + c.w_yc() + c.w_zc(e)) * w[e] * fx[e];
                sp_acc += -(double)g.dt * sum;
            }
        }
    };

    auto slot = [&](const int plane, const int lane) __attribute__((always_inline)) {
        return lds + ((plane % 3) * T + lane) * V;
    };

    // Step A's ldiv of the plane being advanced, and its new lower x face.
    real ld_a[V], s1_lvx_lo[V];
    auto first_face = [&](const int s) __attribute__((always_inline)) {
        real v[V], prev[V];
        vload<real, V>(ld_in + L.cell(s), ld_a);
        vload<real, V>(lvx_in + L.cell(s), v);
        vload<real, V>(ld_in + L.cell(s > 0 ? s - 1 : s), prev);
        const real fxf = LAYER ? uniform_load(a.face[0] + s) : real(1);
        real ws[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real w = v[e] + ((s > 0 ? prev[e] : real(0)) - ld_a[e]);
            s1_lvx_lo[e] = damp3<real, LAYER>(w, fxf, c.fyc, c.fzc[e]);
            ws[e] = w;
        }
        // Face 0 belongs to no halo plane: its first step's sponge term is added here.
        if (SPONGE && owner && s == 0 && i0 == 0) sponge_xface(0, lt_row[0], ws);
    };
    // Step A (adjoint step m, local part of m - 1) of plane s from global memory into o.
    auto stepA = [&](const int s, AdjPlane<real, V, NR> &o, auto xwall_tag) __attribute__((always_inline)) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        const int64_t cc = L.cell(s);
        const bool up_interior = !XWALL || s + 1 < nx;
        real lvx_v[V], ld_up[V], lvy_lo_in[V], lvy_hi_in[V], ld_ylo[V], ld_yhi[V], lvz_v[V], lpc[V];
        vload<real, V>(lvx_in + cc + L.plane, lvx_v);
        vload<real, V>(ld_in + cc + (up_interior ? L.plane : 0), ld_up);
        const int64_t cy = L.face_y(s);
        vload<real, V>(lvy_in + cy, lvy_lo_in);
        vload<real, V>(lvy_in + cy + nz, lvy_hi_in);
        vload<real, V>(ld_in + cc - c.dy_lo, ld_ylo);
        vload<real, V>(ld_in + cc + c.dy_hi, ld_yhi);
        const int64_t cz = L.face_z(s);
        vload<real, V>(lvz_in + cz, lvz_v);
        const real lvz_top = lvz_in[cz + V];
        const real ldz_lo = ld_in[cc - c.dz_lo];
        const real ldz_hi = ld_in[cc + c.dz_hi];
        vload<real, V>(lp_in + cc, lpc);
        real states[NR > 0 ? NR : 1][V];
#pragma unroll
        for (int r = 0; r < NR; ++r) vload<real, V>(lpsi_in + (int64_t)r * n_cells + cc, states[r]);
        const bool own = owner && s >= i0 && s < i1;
        real tv[V];
        if (want_walls && own) load_tape<real, V, XWALL>(a, c, s, m - 1, tv);
        else
#pragma unroll
            for (int e = 0; e < V; ++e) tv[e] = 0;
        real dv[V];
        if (AIR && own) {
            const real *row = a.div_tape + (m - 1 - a.div_tape_step0) * n_cells + cc;
            vload<real, V>(row, dv);
            vload<real, V>(row - n_cells, o.dv_b);  // step m - 2, for step B of this plane
        } else {
#pragma unroll
            for (int e = 0; e < V; ++e) dv[e] = 0;
        }
        // Upper x face s + 1: this chunk's when its owner cell s + 1 is, or it is the far face.
        const bool own_xhi = SPONGE && owner && s + 1 >= i0 && s < i1 && (s + 1 < i1 || s + 1 == nx);
        adjoint_cells<real, V, NR, LAYER, XWALL, AIR, SPONGE>(
            a, c, s, m - 1, 0, own, want_walls, grad_signal, tv, ld_a, ld_up, ld_ylo, ld_yhi,
            ldz_lo, ldz_hi, s1_lvx_lo, lvx_v, lvy_lo_in, lvy_hi_in, lvz_v, lvz_top, lpc, states, o,
            acc, dv, air_acc, g, lt_row[0], lt_row[1], own_xhi, sp_acc);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            ld_a[e] = ld_up[e];
            s1_lvx_lo[e] = o.lvx_hi[e];
        }
    };
    auto stepA_at = [&](const int s, AdjPlane<real, V, NR> &o) __attribute__((always_inline)) {
        if (s == 0 || s == nx - 1) stepA(s, o, Wall{});
        else stepA(s, o, Interior{});
    };

    // Step B (adjoint step m - 1, local part of m - 2) of plane i from step A's results: A
    // (plane i), ld_next (step A's ldiv of plane i + 1), the new lower face s2_lvx_lo and the
    // neighbours in LDS; stores the plane.
    real s2_lvx_lo[V];
    auto stepB = [&](const int i, const AdjPlane<real, V, NR> &A, const real (&ld_next)[V],
                     auto xwall_tag) __attribute__((always_inline)) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        const int64_t cc = L.cell(i);
        real ld_ylo[V], ld_yhi[V];
        vload<real, V>(slot(i, t - G), ld_ylo);
        vload<real, V>(slot(i, t + G), ld_yhi);
        const real ldz_lo = slot(i, t - 1)[V - 1];
        const real ldz_hi = slot(i, t + 1)[0];
        real tv[V];
        if (want_walls) load_tape<real, V, XWALL>(a, c, i, m - 2, tv);
        real lvz_v[V];
#pragma unroll
        for (int e = 0; e < V; ++e) lvz_v[e] = A.lvz[e];
        AdjPlane<real, V, NR> o;
        adjoint_cells<real, V, NR, LAYER, XWALL, AIR, SPONGE>(
            a, c, i, m - 2, 1, true, want_walls, grad_signal, tv, A.ld, ld_next, ld_ylo, ld_yhi,
            ldz_lo, ldz_hi, s2_lvx_lo, A.lvx_hi, A.lvy_lo, A.lvy_hi, lvz_v, A.lvz[V], A.lp, A.psi,
            o, acc, A.dv_b, air_acc, g, lt_row[1], lt_row[2],
            SPONGE && (i + 1 < i1 || i + 1 == nx), sp_acc);
        const int64_t cy = L.face_y(i), cz = L.face_z(i);
        if (i + 1 < i1 || i + 1 == nx) vstore<real, V>(io.lvx_out + cc + L.plane, o.lvx_hi);
        vstore<real, V>(io.lvy_out + cy, o.lvy_lo);
        if (!c.has_yhi) vstore<real, V>(io.lvy_out + cy + nz, o.lvy_hi);
        {
            real lower[V];
#pragma unroll
            for (int e = 0; e < V; ++e) lower[e] = o.lvz[e];
            vstore<real, V>(io.lvz_out + cz, lower);
        }
        if (!c.has_zhi) io.lvz_out[cz + V] = o.lvz[V];
#pragma unroll
        for (int r = 0; r < NR; ++r) vstore<real, V>(io.lpsi_out + (int64_t)r * n_cells + cc, o.psi[r]);
        vstore<real, V>(io.lp_out + cc, o.lp);
        vstore<real, V>(io.ld_out + cc, o.ld);
#pragma unroll
        for (int e = 0; e < V; ++e) s2_lvx_lo[e] = o.lvx_hi[e];
    };

    // Prologue: step A of the halo plane i0 - 1 (its ldiv only) and of plane i0.
    AdjPlane<real, V, NR> A, B;
    real ld_prev[V];
    real v_half_lo[V];  // step A's adjoint vx on face i0
    if (valid) {
        const int s0 = i0 > 0 ? i0 - 1 : i0;
        first_face(s0);
        if (i0 > 0) {
            stepA_at(i0 - 1, B);
#pragma unroll
            for (int e = 0; e < V; ++e) ld_prev[e] = B.ld[e];
        }
#pragma unroll
        for (int e = 0; e < V; ++e) v_half_lo[e] = s1_lvx_lo[e];
        stepA_at(i0, A);
        vstore<real, V>(slot(i0, t), A.ld);
    }
    lds_barrier();
    if (owner) {
        // Step B's adjoint vx on face i0, the lower face of the first cell (the prologue of
        // adjoint_march_vec with step A's results as inputs).
        const real fxf = LAYER ? uniform_load(a.face[0] + i0) : real(1);
        real ws[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real w = v_half_lo[e] + ((i0 > 0 ? ld_prev[e] : real(0)) - A.ld[e]);
            s2_lvx_lo[e] = damp3<real, LAYER>(w, fxf, c.fyc, c.fzc[e]);
            ws[e] = w;
        }
        vstore<real, V>(io.lvx_out + L.cell(i0), s2_lvx_lo);
        sponge_xface(i0, lt_row[1], ws);
    }

    // One plane: step A of plane i + 1 into `next`, step B of plane i from `here`. The march
    // alternates the two states' roles, unrolled by two, instead of copying `next` into `here`
    // after each plane (a copy into registers a load may still target waits for that load,
    // and with it for every store issued since: see fused2_march).
    auto plane = [&](const int i, const AdjPlane<real, V, NR> &here,
                     AdjPlane<real, V, NR> &next) __attribute__((always_inline)) {
        const bool more = i + 1 < nx;
        if (valid && more) {
            stepA_at(i + 1, next);
            vstore<real, V>(slot(i + 1, t), next.ld);
        }
        lds_barrier();
        if (owner) {
            real ld_next[V];
#pragma unroll
            for (int e = 0; e < V; ++e) ld_next[e] = more ? next.ld[e] : here.ld[e];
            if (i == 0 || i == nx - 1) stepB(i, here, ld_next, Wall{});
            else stepB(i, here, ld_next, Interior{});
        }
    };
    for (int i = i0; i < i1; i += 2) {
        plane(i, A, B);
        if (i + 1 < i1) plane(i + 1, B, A);
    }
}

// Each wave adds its sums of the six wall partials into its own row of `partials`.
__device__ inline void flush_wall_rows(const double (&acc)[6], const int64_t row, double *partials) {
    bool nonzero = false;
#pragma unroll
    for (int q = 0; q < 6; ++q) nonzero = nonzero || acc[q] != 0;
    if (!__any(nonzero)) return;
    double sum[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) sum[q] = wave_sum(acc[q]);
    if ((threadIdx.x & 63u) == 0) {
        double *p = partials + row * kGradSlots + kGradWall;
        double old[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) old[q] = p[q];
#pragma unroll
        for (int q = 0; q < 6; ++q) p[q] = old[q] + sum[q];
    }
}

// Each wave adds its sums of the air partials (relax_scale, then one per relaxation state) into
// its own row of `partials`.
template <int NA>
__device__ inline void flush_air_rows(const double (&acc)[NA], const int count, const int64_t row,
                                      double *partials) {
    bool nonzero = false;
#pragma unroll
    for (int q = 0; q < NA; ++q) nonzero = nonzero || acc[q] != 0;
    if (!__any(nonzero)) return;
    double sum[NA];
#pragma unroll
    for (int q = 0; q < NA; ++q) sum[q] = wave_sum(acc[q]);
    if ((threadIdx.x & 63u) == 0) {
        double *p = partials + row * kGradSlots + kGradScale;
        for (int q = 0; q < count; ++q) p[q] = p[q] + sum[q];
    }
}

// Each wave adds its sum of the sponge partial into its own row of `partials`.
__device__ inline void flush_sponge_rows(const double acc, const int64_t row, double *partials) {
    if (!__any(acc != 0)) return;
    const double sum = wave_sum(acc);
    if ((threadIdx.x & 63u) == 0) {
        double *p = partials + row * kGradSlots + kGradSponge;
        *p = *p + sum;
    }
}

template <typename real, int V, int NR, bool LAYER, bool AIR = false, bool SPONGE = false>
__global__ __launch_bounds__(kFused2Threads, 2) void adjoint2_step(
    const SweepArgs<real> a, const AdjointArgs<real> g, const Adjoint2IO<real> io,
    const Tiling2 t2, const int64_t m) {
    __shared__ real lds[3 * kFused2Threads * V];
    // Tables read inside the march, copied once (see fused2_step): sources, receivers, the
    // receivers' dL/d(recording) of the two local steps, the wall table.
    __shared__ int32_t src_lds[2 * kFused2Sources];
    __shared__ int64_t len_lds[kFused2Sources];
    __shared__ int32_t rx_lds[2 * kFused2Receivers];
    __shared__ real rxg_lds[2 * kFused2Receivers];
    __shared__ real wt_lds[2 * ACFDTD_HIP_WALL_CODES];
    __shared__ real w_lds[SPONGE ? WeightLds<real, V>::size : 1];
    const unsigned n = (unsigned)t2.launched;
    unsigned b = blockIdx.x;
    b = (b & 7u) * (n >> 3) + (b >> 3);  // contiguous tiles per XCD, as decode_tile
    if (b >= (unsigned)t2.tiles) return;  // whole workgroup, before any barrier
    const int blk = (int)(b % (unsigned)t2.nblk);
    const int xc = (int)(b / (unsigned)t2.nblk);
    const int bg = blk % t2.nblk_g, bj = blk / t2.nblk_g;
    // SPONGE: the grading weights of the block's rows (one halo row each side) and cells.
    const int kbase = max(0, (bg * t2.seg - t2.zhalo) * V);
    const int jbase = bj * t2.rows - 1;
    if (SPONGE) {
        using W = WeightLds<real, V>;
        const int nzc = min(W::kZ - 1, (t2.seg + 2 * t2.zhalo) * V);
        for (int q = (int)threadIdx.x; q <= nzc; q += (int)blockDim.x) {
            const int k = kbase + q;
            w_lds[W::zc + q] = k < a.nz ? g.w_cell[2][k] : real(0);
            w_lds[W::zf + q] = k <= a.nz ? g.w_face[2][k] : real(0);
        }
        for (int q = (int)threadIdx.x; q < t2.rows + 3 && q < W::kY; q += (int)blockDim.x) {
            const int jr = jbase + q;
            w_lds[W::yc + q] = jr >= 0 && jr < a.ny ? g.w_cell[1][jr] : real(0);
            w_lds[W::yf + q] = jr >= 0 && jr <= a.ny ? g.w_face[1][jr] : real(0);
        }
    }
    const int t = (int)threadIdx.x;
    const int r = t / t2.pitch - 1;
    const int sl = t % t2.pitch - t2.zhalo;
    const int jj = bj * t2.rows + r;
    const int gg = bg * t2.seg + sl;
    const bool valid = t < (t2.rows + 2) * t2.pitch && jj >= 0 && jj < a.ny && gg >= 0 &&
                       gg < t2.groups;
    const bool owner = valid && r >= 0 && r < t2.rows && sl >= 0 && sl < t2.seg;
    const int j = valid ? jj : 0;
    const int k0 = valid ? gg * V : 0;
    int i0, i1;
    chunk_planes(t2, blk, xc, a.nx, i0, i1);
    // adjoint2_applies admits at most kFused2Sources sources and kFused2Receivers receivers.
    for (int q = t; q < a.n_sources; q += (int)blockDim.x) {
        src_lds[2 * q] = a.source_cell[2 * q];
        src_lds[2 * q + 1] = a.source_cell[2 * q + 1];
        len_lds[q] = a.source_length[q];
    }
    for (int q = t; q < a.n_receivers; q += (int)blockDim.x) {
        rx_lds[2 * q] = a.receiver_cell[2 * q];
        rx_lds[2 * q + 1] = a.receiver_cell[2 * q + 1];
        for (int h = 0; h < 2; ++h)
            rxg_lds[2 * q + h] = g.grad_recording[q * g.grad_recording_stride +
                                                  (m - 1 - h - g.grad_recording_step0)];
    }
    const bool walls = a.wall_table != nullptr;
    if (walls)
        for (int q = t; q < 2 * ACFDTD_HIP_WALL_CODES; q += (int)blockDim.x) wt_lds[q] = a.wall_table[q];
    lds_barrier();
    const bool want_walls = (g.grads & ACFDTD_HIP_GRAD_WALLS) && a.wall_tape;
    double acc[6] = {};
    double air_acc[1 + (NR > 0 ? NR : 1)] = {};
    double sp_acc = 0;
    adjoint2_march<real, V, NR, LAYER, AIR, SPONGE>(
        a, t2, i0, i1, j, k0, valid, owner, t, m, want_walls, g.grad_signal, io, lds, src_lds,
        len_lds, rx_lds, rxg_lds, walls ? wt_lds : nullptr, acc, air_acc, g, sp_acc, w_lds,
        jj - jbase, k0 - kbase);
    const int64_t row = (int64_t)b * (t2.threads / 64) + t / 64;
    if (want_walls) flush_wall_rows(acc, row, g.partials);
    if (AIR) flush_air_rows(air_acc, 1 + NR, row, g.partials);
    if (SPONGE) flush_sponge_rows(sp_acc, row, g.partials);
}

template <typename real>
using Adjoint2Kernel = void (*)(const SweepArgs<real>, const AdjointArgs<real>,
                                const Adjoint2IO<real>, const Tiling2, const int64_t);

template <typename real, int NR>
Adjoint2Kernel<real> pick_adjoint2_nr(bool layer, bool air) {
    constexpr int V = sizeof(real) == 4 ? 4 : 2;
    if constexpr (NR > 0) {
        if (air) return layer ? adjoint2_step<real, V, NR, true, true>
                              : adjoint2_step<real, V, NR, false, true>;
    }
    (void)air;
    return layer ? adjoint2_step<real, V, NR, true> : adjoint2_step<real, V, NR, false>;
}

// SPONGE kernels (the sponge gradient from the layer tape; the layer is present), with or
// without the air gradients. Instantiated in adjoint2s.hip and adjoint2s_64.hip.
template <typename real>
Adjoint2Kernel<real> pick_adjoint2_sponge(int nr, bool air) {
    constexpr int V = sizeof(real) == 4 ? 4 : 2;
    switch (nr) {
        case 0: return adjoint2_step<real, V, 0, true, false, true>;
        case 1: return air ? adjoint2_step<real, V, 1, true, true, true> : adjoint2_step<real, V, 1, true, false, true>;
        case 2: return air ? adjoint2_step<real, V, 2, true, true, true> : adjoint2_step<real, V, 2, true, false, true>;
        case 3: return air ? adjoint2_step<real, V, 3, true, true, true> : adjoint2_step<real, V, 3, true, false, true>;
        default: return air ? adjoint2_step<real, V, 4, true, true, true> : adjoint2_step<real, V, 4, true, false, true>;
    }
}
extern template Adjoint2Kernel<float> pick_adjoint2_sponge<float>(int, bool);
extern template Adjoint2Kernel<double> pick_adjoint2_sponge<double>(int, bool);

// AIR kernels (air gradients from the divergence tape) when the air gradients are asked for.
template <typename real>
Adjoint2Kernel<real> pick_adjoint2(int nr, bool layer, bool air) {
    switch (nr) {
        case 0: return pick_adjoint2_nr<real, 0>(layer, false);
        case 1: return pick_adjoint2_nr<real, 1>(layer, air);
        case 2: return pick_adjoint2_nr<real, 2>(layer, air);
        case 3: return pick_adjoint2_nr<real, 3>(layer, air);
        default: return pick_adjoint2_nr<real, 4>(layer, air);
    }
}

// Whether acfdtd_hip_adjoint_run may advance two adjoint steps per sweep: asked for, both
// sets complete, no forward states given (the steps that read them run one per sweep), the
// sponge gradient only from a layer tape and the air gradients only from a divergence tape,
// the source and receiver tables within the LDS copies.
inline bool adjoint2_applies(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &g) {
    const bool sponge = (g.grads & ACFDTD_HIP_GRAD_SPONGE) && d.has_layer;
    const bool air = (g.grads & ACFDTD_HIP_GRAD_AIR) && d.n_relax > 0;
    const bool states = g.fwd_count > 0 && (sponge || (air && d.div_tape == nullptr));
    return g.time_block == 2 && has_adjoint_swap(g) && g.lp_swap != nullptr &&
           (d.n_relax == 0 || g.lpsi_swap != nullptr) && !states &&
           !(sponge && d.layer_tape == nullptr) && !(air && d.div_tape == nullptr) &&
           d.n_sources <= kFused2Sources && d.n_receivers <= kFused2Receivers;
}

template <typename real>
int64_t launch_adjoint2_quads(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &gd,
                              int64_t step_hi, int64_t n_steps, hipStream_t stream) {
    Tiling2 t2;
    if (!adjoint2_applies(d, gd) || !make_tiling2(d, sizeof(real) == 4 ? 4 : 2, t2)) return 0;
    balance_tiling2(t2, d.nx);
    SweepArgs<real> a;
    fill_args(d, a);
    AdjointArgs<real> g;
    fill_adjoint(gd, g);
    const bool air = (gd.grads & ACFDTD_HIP_GRAD_AIR) && d.n_relax > 0 && d.div_tape != nullptr;
    const bool sponge = (gd.grads & ACFDTD_HIP_GRAD_SPONGE) && d.has_layer && d.layer_tape != nullptr;
    const Adjoint2Kernel<real> k = sponge ? pick_adjoint2_sponge<real>(d.n_relax, air)
                                          : pick_adjoint2<real>(d.n_relax, d.has_layer != 0, air);
    real *first[6] = {static_cast<real *>(gd.ldiv), static_cast<real *>(gd.lvx),
                      static_cast<real *>(gd.lvy), static_cast<real *>(gd.lvz),
                      static_cast<real *>(gd.lp), static_cast<real *>(gd.lpsi)};
    real *second[6] = {static_cast<real *>(gd.ldiv_swap), static_cast<real *>(gd.lvx_swap),
                       static_cast<real *>(gd.lvy_swap), static_cast<real *>(gd.lvz_swap),
                       static_cast<real *>(gd.lp_swap), static_cast<real *>(gd.lpsi_swap)};
    auto io_of = [](real *const *in, real *const *out) {
        Adjoint2IO<real> io{};
        io.ld_in = in[0];
        io.lvx_in = in[1];
        io.lvy_in = in[2];
        io.lvz_in = in[3];
        io.lp_in = in[4];
        io.lpsi_in = in[5];
        io.ld_out = out[0];
        io.lvx_out = out[1];
        io.lvy_out = out[2];
        io.lvz_out = out[3];
        io.lp_out = out[4];
        io.lpsi_out = out[5];
        return io;
    };
    const Adjoint2IO<real> forward = io_of(first, second), backward = io_of(second, first);
    // Every step of a quad needs its local part: the lowest, step_hi - n - 3, must be >= 1.
    const int64_t usable = n_steps < step_hi ? n_steps : step_hi;
    const dim3 grid(t2.launched), block(t2.threads);
    int64_t n = 0;
    for (; n + 3 < usable; n += 4) {
        hipLaunchKernelGGL(k, grid, block, 0, stream, a, g, forward, t2, step_hi - n);
        hipLaunchKernelGGL(k, grid, block, 0, stream, a, g, backward, t2, step_hi - n - 2);
    }
    return n;
}

}  // namespace acfdtd
