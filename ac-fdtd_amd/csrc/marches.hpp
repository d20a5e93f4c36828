// Device bodies of the time step, shared by the single-room, fused and batched kernels.
//
// Each body marches one lane's column of cells along x over [i0, i1): velocity_march and
// pressure_march are the two in-place sweeps, fused_march_vec the one-sweep step from an input
// set of fields to an output set with V z-cells per lane. Pointers are
// passed as restrict-qualified arguments of inlined functions, so the compiler may issue the
// loads of the next planes before the stores of the current one retire.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "acfdtd_hip.h"
#include "sweeps.hpp"

namespace acfdtd {

// Planes marched per loop trip; with the restrict-qualified bodies below, the loads of the
// next planes can be issued before the stores of the current one retire.
#ifndef ACFDTD_X_UNROLL
#define ACFDTD_X_UNROLL 2
#endif
constexpr int kXUnroll = ACFDTD_X_UNROLL;

template <typename real, bool LAYER, bool FUSED_FMA>
__device__ __forceinline__ void velocity_march(
    const SweepArgs<real> &a, const int i0, const int i1, const int j, const int k,
    const real *__restrict__ p, real *__restrict__ vx, real *__restrict__ vy,
    real *__restrict__ vz, const real *__restrict__ cell_x, const real *__restrict__ face_x,
    const real *__restrict__ cell_y, const real *__restrict__ face_y,
    const real *__restrict__ cell_z, const real *__restrict__ face_z) {
    const Layout L = room_layout(a, j, k);
    const real vel = a.vel;
    real fyc = 1, fzc = 1, fyf = 1, fzf = 1;
    if (LAYER) {
        fyc = cell_y[j];
        fzc = cell_z[k];
        fyf = face_y[j];
        fzf = face_z[k];
    }
    real p_prev = i0 > 0 ? p[L.cell(i0 - 1)] : real(0);
#pragma unroll kXUnroll
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real pc = p[c];
        real fxc = 1, fxf = 1;
        if (LAYER) {
            fxc = cell_x[i];
            fxf = face_x[i];
        }
        // vx on face i (lower face of cell i); face 0 is a wall and is only ever damped.
        if (LAYER || i > 0) {
            real v = vx[c];
            if (i > 0) v = face_update<real, FUSED_FMA>(v, vel, pc - p_prev);
            if (LAYER) v = ((v * fxf) * fyc) * fzc;
            vx[c] = v;
        }
        const int64_t cy = L.face_y(i);
        if (LAYER || j > 0) {
            real v = vy[cy];
            if (j > 0) v = face_update<real, FUSED_FMA>(v, vel, pc - p[c - a.zp]);
            if (LAYER) v = ((v * fxc) * fyf) * fzc;
            vy[cy] = v;
        }
        const int64_t cz = L.face_z(i);
        if (LAYER || k > 0) {
            real v = vz[cz];
            if (k > 0) v = face_update<real, FUSED_FMA>(v, vel, pc - p[c - 1]);
            if (LAYER) v = ((v * fxc) * fyc) * fzf;
            vz[cz] = v;
        }
        if (LAYER) {
            // The far wall faces belong to the last cell on each axis; like face 0 they are
            // never updated, only damped (the reference damps whole slabs, walls included).
            if (i == a.nx - 1) {
                const int64_t t = c + L.plane;
                vx[t] = ((vx[t] * face_x[a.nx]) * fyc) * fzc;
            }
            if (j == a.ny - 1) {
                const int64_t t = cy + a.zp;
                vy[t] = ((vy[t] * fxc) * face_y[a.ny]) * fzc;
            }
            if (k == a.nz - 1) {
                const int64_t t = cz + 1;
                vz[t] = ((vz[t] * fxc) * fyc) * face_z[a.nz];
            }
        }
        p_prev = pc;
    }
}


template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__device__ __forceinline__ void pressure_march(
    const SweepArgs<real> &a, const int i0, const int i1, const int j, const int k,
    const int64_t step, const int64_t column, real *__restrict__ p, real *__restrict__ psi,
    const real *__restrict__ vx, const real *__restrict__ vy, const real *__restrict__ vz,
    const real *__restrict__ wall_table, const real *__restrict__ cell_x,
    const real *__restrict__ cell_y, const real *__restrict__ cell_z,
    real *__restrict__ recording) {
    const Layout L = room_layout(a, j, k);
    const int64_t n_cells = L.plane * a.nx;
    const int jk = j * a.zp + k;
    const int code_yz = WALLS ? (wall_class(j, a.ny) << 2) | (wall_class(k, a.nz) << 4) : 0;
    real fyc = 1, fzc = 1;
    if (LAYER) {
        fyc = cell_y[j];
        fzc = cell_z[k];
    }

    real vx_lo_next = vx[L.cell(i0)];
#pragma unroll kXUnroll
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real vx_lo = vx_lo_next;
        const real vx_hi = vx[c + L.plane];
        vx_lo_next = vx_hi;
        const int64_t cy = L.face_y(i);
        const int64_t cz = L.face_z(i);
        const real div = ((vx_hi - vx_lo) + (vy[cy + a.zp] - vy[cy])) + (vz[cz + 1] - vz[cz]);
        if (a.div_tape) a.div_tape[(step - a.div_tape_step0) * n_cells + c] = div;

        const real p_old = p[c];
        real pn;
        if (COMPENSATED) {
            // p_old + sum_r off_r mean_r + div relax_scale - div pres, summed with TwoSum so the
            // increment costs one rounding instead of NR + 2, and the dominant product exact.
            real s = p_old, e = 0;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                real *state = psi + (int64_t)r * n_cells + c;
                const real target = div * a.gain[r];
                const real offset = *state - target;
                *state = offset * a.decay[r] + target;
                two_sum_acc(s, offset * a.mean[r], e);
            }
            if (NR > 0) two_sum_acc(s, div * a.relax_scale, e);
            const real prod = div * a.pres;
            const real prod_err = fma_real(div, a.pres, -prod);
            two_sum_acc(s, -prod, e);
            pn = s + (e - prod_err);
        } else {
            pn = p_old;
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
        }

        if (a.wall_tape) {
            const int code = (wall_class(j, a.ny) << 2) | (wall_class(k, a.nz) << 4) |
                             wall_class(i, a.nx);
            if (code)
                tape_store(a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride,
                           TapePlanes(a.nx, a.ny, a.nz), a.nx, a.ny, a.nz, i, j, k, code,
                           pn + p_old);
        }
        if (WALLS) {
            const int code = code_yz | wall_class(i, a.nx);
            if ((a.wall_mask >> code) & 1ull)
                pn = pn * wall_table[code] - p_old * wall_table[ACFDTD_HIP_WALL_CODES + code];
        }
        if (LAYER) pn = ((pn * cell_x[i]) * fyc) * fzc;

        // Sources, in insertion order; the plane test is wave-uniform (scalar).
        for (int s = 0; s < a.n_sources; ++s) {
            if (a.source_cell[2 * s] == i && a.source_cell[2 * s + 1] == jk &&
                step < a.source_length[s])
                pn = pn + a.src * a.source_signal[s * a.source_stride + step];
        }
        p[c] = pn;
        for (int r = 0; r < a.n_receivers; ++r) {
            if (a.receiver_cell[2 * r] == i && a.receiver_cell[2 * r + 1] == jk)
                recording[r * a.recording_stride + column] = pn;
        }
    }
}


// ------------------------------------------------------------------ fused single sweep --
//
// One step in one pass over memory: read p^n and v^{n-1/2} from the `in` set, write
// v^{n+1/2} and p^{n+1} to the `out` set, advance psi in place. Each cell computes the new
// velocity on all six of its faces; a face shared with a neighbouring cell is computed by both
// from the same unmodified inputs with the same operations, so both get the same bits, and
// only the face's owner (the cell above it; the last cell for the far wall) writes it. The
// x-neighbour face comes from a register carried along the march. Wall faces are copied
// (damped when there is a sponge), as the two-sweep path leaves them in place.

template <typename real>
struct FusedIO {
    const real *p_in, *vx_in, *vy_in, *vz_in;
    real *p_out, *vx_out, *vy_out, *vz_out;
};

template <typename real, bool LAYER>
__device__ __forceinline__ real damp3(real v, real f1, real f2, real f3) {
    if (LAYER) return ((v * f1) * f2) * f3;
    return v;
}

// ----------------------------------------------------------- fused sweep, V cells per lane --
//
// Each lane owns V consecutive z cells (V = 1, 2 or 4), so every field access is a V-wide
// vector access (dwordx2 / dwordx4): fp32 with one cell per lane issues too few bytes per
// memory instruction to keep HBM busy. Used when nz is a multiple of V:
// then each lane's cells lie in one row and every vector is V-element aligned in all fields
// except vz (rows of nz + 1), whose accesses only need the 4-byte alignment gfx950 requires.
//
// The body is written so that no load depends on a branch: neighbours that do not exist at a
// wall are loaded from a clamped (valid) address and discarded with a select, wall factors of
// the interior planes are looked up once per lane, the two x-wall planes are peeled off the
// march, and sources, receivers and the boundary tape sit behind per-lane flags computed once.
// The compiler can then issue all of a plane's loads before the first wait.

template <typename real, int V>
struct VecOf {
    typedef real type __attribute__((ext_vector_type(V), aligned(sizeof(real))));
};

template <typename real, int V>
__device__ __forceinline__ void vload(const real *__restrict__ p, real (&out)[V]) {
    if constexpr (V == 1) {
        out[0] = *p;
    } else {
        const typename VecOf<real, V>::type v =
            *reinterpret_cast<const typename VecOf<real, V>::type *>(p);
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] = v[e];
    }
}

template <typename real, int V>
__device__ __forceinline__ void vstore(real *__restrict__ p, const real (&in)[V]) {
    if constexpr (V == 1) {
        *p = in[0];
    } else {
        typename VecOf<real, V>::type v;
#pragma unroll
        for (int e = 0; e < V; ++e) v[e] = in[e];
        *reinterpret_cast<typename VecOf<real, V>::type *>(p) = v;
    }
}

// x[e] and x[e] += add for a run-time e < V, through a vector value: indexing the array itself
// with a run-time index would place it (and what the optimiser merges with it) in scratch
// memory.
template <typename real, int V>
__device__ __forceinline__ real lane_get(const real (&x)[V], const int e) {
    if constexpr (V == 1) {
        (void)e;
        return x[0];
    } else {
        typename VecOf<real, V>::type v;
#pragma unroll
        for (int q = 0; q < V; ++q) v[q] = x[q];
        return v[e];
    }
}

template <typename real, int V>
__device__ __forceinline__ void lane_add(real (&x)[V], const int e, const real add) {
    if constexpr (V == 1) {
        (void)e;
        x[0] = x[0] + add;
    } else {
        typename VecOf<real, V>::type v;
#pragma unroll
        for (int q = 0; q < V; ++q) v[q] = x[q];
        v[e] = v[e] + add;
#pragma unroll
        for (int q = 0; q < V; ++q) x[q] = v[q];
    }
}

// Lane position in the flattened (y, z / V) plane used by the V-wide fused kernels: consecutive
// lanes take consecutive V-cell groups of a row and run on into the next row, so no lane idles
// whatever nz is. Returns false for lanes past the plane.
__device__ inline bool flat_lane(const Tiling &t, int ny, int zp, int jb, int &j, int &k0) {
    const int groups = zp / t.vec;
    const int64_t flat = (int64_t)jb * (64 * t.tile_j) + threadIdx.y * 64 + threadIdx.x;
    if (flat >= (int64_t)ny * groups) return false;
    j = (int)(flat / groups);
    k0 = (int)(flat - (int64_t)j * groups) * t.vec;
    return true;
}

// PADDED: the room's rows may be padded (z_pitch), so the last lane of a row may hold fewer
// than V of the room's cells: its cells from ztop = nz - k0 on are padding, and its face ztop is
// the room's upper z wall. That face keeps its input value (damped, as every wall face), the
// sponge factors of padding cells are read clamped to the room's profiles, and nothing is
// stored past the lane's own V elements. Padding cells compute values that reach no cell of the
// room: the only face between them and the room is the wall face, which reads no pressure.
template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED,
          bool PADDED = false>
__device__ __forceinline__ void fused_march_vec(
    const SweepArgs<real> &a, const int i0, const int i1, const int j, const int k0,
    const int64_t step, const int64_t column, const real *__restrict__ p_in,
    const real *__restrict__ vx_in, const real *__restrict__ vy_in,
    const real *__restrict__ vz_in, real *__restrict__ p_out, real *__restrict__ vx_out,
    real *__restrict__ vy_out, real *__restrict__ vz_out, real *__restrict__ psi,
    const real *__restrict__ wall_table, const real *__restrict__ cell_x,
    const real *__restrict__ face_x, const real *__restrict__ cell_y,
    const real *__restrict__ face_y, const real *__restrict__ cell_z,
    const real *__restrict__ face_z, real *__restrict__ recording) {
    const int nx = a.nx, ny = a.ny, nz = a.nz, zp = a.zp;
    const Layout L = room_layout(a, j, k0);
    const int64_t n_cells = L.plane * nx;
    const int jk0 = j * zp + k0;
    const real vel = a.vel;

    // Neighbours that exist, and clamped offsets to read in their place when they do not.
    const bool has_ylo = j > 0, has_yhi = j + 1 < ny;
    const bool has_zlo = k0 > 0, has_zhi = k0 + V < nz;
    const int64_t dy_lo = has_ylo ? zp : 0, dy_hi = has_yhi ? zp : 0;
    const int64_t dz_lo = has_zlo ? 1 : 0, dz_hi = has_zhi ? V : V - 1;
    // PADDED: the lane's first padding cell (its upper wall face), V or more when it has none;
    // the upper wall face past the lane (vz_top) is then read only in lanes without padding.
    const int ztop = nz - k0;
    const bool partial = PADDED && ztop < V;
    const int64_t dz_top = partial ? V - 1 : V;

    real fyc = 1, fyf_lo = 1, fyf_hi = 1, fzc[V], fzf[V + 1];
#pragma unroll
    for (int e = 0; e < V; ++e) fzc[e] = 1;
#pragma unroll
    for (int e = 0; e <= V; ++e) fzf[e] = 1;
    if (LAYER) {
        fyc = cell_y[j];
        fyf_lo = face_y[j];
        fyf_hi = face_y[j + 1];
#pragma unroll
        for (int e = 0; e < V; ++e) fzc[e] = cell_z[PADDED ? min(k0 + e, nz - 1) : k0 + e];
#pragma unroll
        for (int e = 0; e <= V; ++e) fzf[e] = face_z[PADDED ? min(k0 + e, nz) : k0 + e];
    }

    // Wall classes: y and z part per cell; the factors of interior planes (no x wall).
    int code_yz[V];
    bool wall_in[V];
    real fu_in[V], fp_in[V];
    const bool yz_boundary = !has_ylo || !has_yhi || !has_zlo || !has_zhi;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        code_yz[e] = (wall_class(j, ny) << 2) | (wall_class(k0 + e, nz) << 4);
        wall_in[e] = WALLS && ((a.wall_mask >> code_yz[e]) & 1ull);
        fu_in[e] = 1;
        fp_in[e] = 0;
    }
    if (WALLS && yz_boundary) {
#pragma unroll
        for (int e = 0; e < V; ++e)
            if (wall_in[e]) {
                fu_in[e] = wall_table[code_yz[e]];
                fp_in[e] = wall_table[ACFDTD_HIP_WALL_CODES + code_yz[e]];
            }
    }

    // The planes [ev_lo, ev_hi] of the sources and receivers in this lane's cells (none when
    // ev_lo > ev_hi): the per-plane scans of the tables run only there.
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

    // New vx on face i0, the lower face of the first cell of the march.
    real p_c[V], vx_lo[V];
    vload<real, V>(p_in + L.cell(i0), p_c);
    {
        real v[V], prev[V];
        vload<real, V>(vx_in + L.cell(i0), v);
        vload<real, V>(p_in + L.cell(i0 > 0 ? i0 - 1 : i0), prev);
        const real fxf = LAYER ? face_x[i0] : real(1);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real w = i0 > 0 ? v[e] - vel * (p_c[e] - prev[e]) : v[e];
            vx_lo[e] = damp3<real, LAYER>(w, fxf, fyc, fzc[e]);
        }
        vstore<real, V>(vx_out + L.cell(i0), vx_lo);
    }

    // One plane; XWALL marks the planes i = 0 and i = nx - 1, where x walls and the missing
    // upper neighbour are handled.
    auto plane = [&](const int i, auto xwall_tag) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        const int64_t c = L.cell(i);
        const bool up_interior = !XWALL || i + 1 < nx;
        const real fxc = LAYER ? uniform_load(cell_x + i) : real(1);
        const real fxf_hi = LAYER ? uniform_load(face_x + i + 1) : real(1);

        // Loads, none behind a branch.
        real vx_v[V], p_up[V], vy_lo_in[V], vy_hi_in[V], q_lo[V], q_hi[V], vz_v[V];
        vload<real, V>(vx_in + c + L.plane, vx_v);
        vload<real, V>(p_in + c + (up_interior ? L.plane : 0), p_up);
        const int64_t cy = L.face_y(i);
        vload<real, V>(vy_in + cy, vy_lo_in);
        vload<real, V>(vy_in + cy + zp, vy_hi_in);
        vload<real, V>(p_in + c - dy_lo, q_lo);
        vload<real, V>(p_in + c + dy_hi, q_hi);
        const int64_t cz = L.face_z(i);
        vload<real, V>(vz_in + cz, vz_v);
        const real vz_top = vz_in[cz + dz_top];
        const real pz_lo = p_in[c - dz_lo];
        const real pz_hi = p_in[c + dz_hi];
        real states[NR > 0 ? NR : 1][V];
#pragma unroll
        for (int r = 0; r < NR; ++r) vload<real, V>(psi + (int64_t)r * n_cells + c, states[r]);

        // New velocities on the six faces of each cell.
        real vx_hi[V], vy_lo[V], vy_hi[V], vz_f[V + 1];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real wx = up_interior ? vx_v[e] - vel * (p_up[e] - p_c[e]) : vx_v[e];
            vx_hi[e] = damp3<real, LAYER>(wx, fxf_hi, fyc, fzc[e]);
            const real wl = has_ylo ? vy_lo_in[e] - vel * (p_c[e] - q_lo[e]) : vy_lo_in[e];
            vy_lo[e] = damp3<real, LAYER>(wl, fxc, fyf_lo, fzc[e]);
            const real wh = has_yhi ? vy_hi_in[e] - vel * (q_hi[e] - p_c[e]) : vy_hi_in[e];
            vy_hi[e] = damp3<real, LAYER>(wh, fxc, fyf_hi, fzc[e]);
            real wz;
            if (e > 0) {
                wz = vz_v[e] - vel * (p_c[e] - p_c[e - 1]);
                if (PADDED && e == ztop) wz = vz_v[e];  // the upper wall face
            } else {
                wz = has_zlo ? vz_v[0] - vel * (p_c[0] - pz_lo) : vz_v[0];
            }
            vz_f[e] = damp3<real, LAYER>(wz, fxc, fyc, fzf[e]);
        }
        {
            const real wz = has_zhi ? vz_top - vel * (pz_hi - p_c[V - 1]) : vz_top;
            vz_f[V] = damp3<real, LAYER>(wz, fxc, fyc, fzf[V]);
        }
        if (i + 1 < i1 || i + 1 == nx) vstore<real, V>(vx_out + c + L.plane, vx_hi);
        vstore<real, V>(vy_out + cy, vy_lo);
        if (!has_yhi) vstore<real, V>(vy_out + cy + zp, vy_hi);
        {
            real lower[V];
#pragma unroll
            for (int e = 0; e < V; ++e) lower[e] = vz_f[e];
            vstore<real, V>(vz_out + cz, lower);
        }
        if (!has_zhi && !partial) vz_out[cz + V] = vz_f[V];

        // Pressure with the relaxation states, walls and sponge.
        const int cx = XWALL ? wall_class(i, nx) : 0;
        real p_new[V], tape_value[V], div_value[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real div =
                ((vx_hi[e] - vx_lo[e]) + (vy_hi[e] - vy_lo[e])) + (vz_f[e + 1] - vz_f[e]);
            div_value[e] = div;
            const real p_old = p_c[e];
            real pn;
            if (COMPENSATED) {
                real s = p_old, err = 0;
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const real target = div * a.gain[r];
                    const real offset = states[r][e] - target;
                    states[r][e] = offset * a.decay[r] + target;
                    two_sum_acc(s, offset * a.mean[r], err);
                }
                if (NR > 0) two_sum_acc(s, div * a.relax_scale, err);
                const real prod = div * a.pres;
                const real prod_err = fma_real(div, a.pres, -prod);
                two_sum_acc(s, -prod, err);
                pn = s + (err - prod_err);
            } else {
                pn = p_old;
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const real target = div * a.gain[r];
                    const real offset = states[r][e] - target;
                    states[r][e] = offset * a.decay[r] + target;
                    pn = pn + offset * a.mean[r];
                }
                if (NR > 0) pn = pn + div * a.relax_scale;
                pn = pn - div * a.pres;
            }
            tape_value[e] = pn + p_old;
            if (WALLS) {
                bool is_wall = wall_in[e];
                real fu = fu_in[e], fp = fp_in[e];
                if (XWALL) {
                    const int code = code_yz[e] | cx;
                    is_wall = (a.wall_mask >> code) & 1ull;
                    if (is_wall) {
                        fu = wall_table[code];
                        fp = wall_table[ACFDTD_HIP_WALL_CODES + code];
                    }
                }
                if (is_wall) pn = pn * fu - p_old * fp;
            }
            if (LAYER) pn = ((pn * fxc) * fyc) * fzc[e];
            p_new[e] = pn;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) vstore<real, V>(psi + (int64_t)r * n_cells + c, states[r]);
        if (a.div_tape) vstore<real, V>(a.div_tape + (step - a.div_tape_step0) * n_cells + c, div_value);
        if (a.wall_tape && (XWALL || yz_boundary)) {
            real *row = a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const int code = code_yz[e] | cx;
                if (code)
                    tape_store(row, TapePlanes(nx, ny, nz), nx, ny, nz, i, j, k0 + e, code,
                               tape_value[e]);
            }
        }
        if (i >= ev_lo && i <= ev_hi) {
            // Sources in insertion order, after the sponge (scheme.py:360-363).
            for (int s = 0; s < a.n_sources; ++s) {
                const int e = a.source_cell[2 * s + 1] - jk0;
                if (a.source_cell[2 * s] == i && e >= 0 && e < V && step < a.source_length[s]) {
                    const real add = a.src * a.source_signal[s * a.source_stride + step];
#pragma unroll
                    for (int q = 0; q < V; ++q)
                        if (q == e) p_new[q] = p_new[q] + add;
                }
            }
        }
        vstore<real, V>(p_out + c, p_new);
        if (i >= ev_lo && i <= ev_hi) {
            for (int r = 0; r < a.n_receivers; ++r) {
                const int e = a.receiver_cell[2 * r + 1] - jk0;
                if (a.receiver_cell[2 * r] == i && e >= 0 && e < V) {
#pragma unroll
                    for (int q = 0; q < V; ++q)
                        if (q == e) recording[r * a.recording_stride + column] = p_new[q];
                }
            }
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            vx_lo[e] = vx_hi[e];
            p_c[e] = p_up[e];
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

template <typename real>
FusedIO<real> fused_io(const acfdtd_hip_desc &d, bool reverse) {
    FusedIO<real> io;
    real *a[4] = {static_cast<real *>(d.p), static_cast<real *>(d.vx), static_cast<real *>(d.vy),
                  static_cast<real *>(d.vz)};
    real *b[4] = {static_cast<real *>(d.p_swap), static_cast<real *>(d.vx_swap),
                  static_cast<real *>(d.vy_swap), static_cast<real *>(d.vz_swap)};
    real **in = reverse ? b : a;
    real **out = reverse ? a : b;
    io.p_in = in[0];
    io.vx_in = in[1];
    io.vy_in = in[2];
    io.vz_in = in[3];
    io.p_out = out[0];
    io.vx_out = out[1];
    io.vy_out = out[2];
    io.vz_out = out[3];
    return io;
}

inline bool has_swap(const acfdtd_hip_desc &d) {
    return d.p_swap && d.vx_swap && d.vy_swap && d.vz_swap;
}

// Launch n_steps steps as the two in-place sweeps each (velocity, then pressure); the state
// stays in the descriptor's fields. Takes padded rows. Defined in sweeps.hip.
template <typename real>
int inplace_steps(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps, int64_t column,
                  hipStream_t stream);

// Launch the fused step for the leading even number of n_steps (pairs A -> B -> A); returns the
// number of steps launched. Defined in fused.hip.
template <typename real>
int64_t launch_fused_pairs(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps,
                           int64_t column, hipStream_t stream);

// Launch two-step sweeps for the leading multiple of four of n_steps (pairs A -> B -> A of two
// steps each) when the descriptor asks for them and allows them; returns the number of steps
// launched (0 otherwise). Defined in fused2.hpp, instantiated in fused2.hip and fused2_64.hip.
template <typename real>
int64_t launch_fused2_quads(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps,
                            int64_t column, hipStream_t stream);

// One two-step sweep (steps `step`, `step` + 1), for per-launch timing; defined beside
// launch_fused2_quads.
template <typename real>
int launch_fused2_once(const acfdtd_hip_desc &d, int64_t step, int64_t column, bool reverse,
                       hipStream_t stream);

}  // namespace acfdtd
