// The two fused sweeps of one time step of the staggered pressure-velocity scheme, for gfx950.
//
// One step is (reference src/ac_fdtd/scheme.py:285-296):
//   velocity sweep : v_a -= vel * (p[hi] - p[lo]) on interior faces, then the sponge
//                    v_a = ((v_a * f_x) * f_y) * f_z                      (scheme.py:307-311, :298)
//   pressure sweep : div = ((dvx + dvy) + dvz); relaxation states; p -= pres * div; wall closed
//                    form; sponge; sources; receiver samples               (scheme.py:313-363)
// The sweeps keep the NumPy reference's association term by term, and the library is built
// with -ffp-contract=off, so the fp64 kernels reproduce the CPU oracle bit for bit. The fp32
// kernels run the same code in float; with ACFDTD_HIP_COMPENSATED they instead form the
// pressure increment with error-free transformations and update velocity with a single fma.
//
// Work decomposition (both sweeps): a workgroup is 64 lanes along z (one wavefront per row,
// unit-stride, coalesced) times tile_j rows along y, and marches a chunk of chunk_x planes
// along x. The x-neighbour each cell needs (p[i-1] for the velocity sweep, vx[i+1] for the
// pressure sweep) is carried in a register from the previous plane, so it is read from memory
// once; the y and z neighbours are re-read from L1/L2. Workgroup ids are remapped so that each
// of the 8 XCDs, which receive consecutive ids round-robin, owns a contiguous range of tiles
// and the halo rows/planes shared between neighbouring tiles stay in one L2.
//
// The scheme is a bandwidth-bound stencil (about 0.4 flop per byte in fp32): no MFMA.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "acfdtd_hip.h"
#include "sweeps.hpp"

namespace acfdtd {

thread_local std::string g_last_error;

int fail(int code, const std::string &message) {
    g_last_error = message;
    return code;
}

int check_hip(hipError_t status, const char *what) {
    if (status == hipSuccess) return ACFDTD_HIP_OK;
    return fail(ACFDTD_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(status));
}

// ------------------------------------------------------------------------------- kernels --

// Planes marched per loop trip; with the restrict-qualified bodies below, the loads of the
// next planes can be issued before the stores of the current one retire.
#ifndef ACFDTD_X_UNROLL
#define ACFDTD_X_UNROLL 2
#endif

template <typename real, bool LAYER, bool FUSED_FMA>
__device__ __forceinline__ void velocity_march(
    const SweepArgs<real> &a, const int i0, const int i1, const int j, const int k,
    const real *__restrict__ p, real *__restrict__ vx, real *__restrict__ vy,
    real *__restrict__ vz, const real *__restrict__ cell_x, const real *__restrict__ face_x,
    const real *__restrict__ cell_y, const real *__restrict__ face_y,
    const real *__restrict__ cell_z, const real *__restrict__ face_z) {
    const Layout L(a.nx, a.ny, a.nz, j, k);
    const real vel = a.vel;
    real fyc = 1, fzc = 1, fyf = 1, fzf = 1;
    if (LAYER) {
        fyc = cell_y[j];
        fzc = cell_z[k];
        fyf = face_y[j];
        fzf = face_z[k];
    }
    real p_prev = i0 > 0 ? p[L.cell(i0 - 1)] : real(0);
#pragma unroll ACFDTD_X_UNROLL
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
            if (j > 0) v = face_update<real, FUSED_FMA>(v, vel, pc - p[c - a.nz]);
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
                const int64_t t = cy + a.nz;
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

template <typename real, bool LAYER, bool FUSED_FMA>
__global__ __launch_bounds__(512) void velocity_sweep(const SweepArgs<real> a) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    velocity_march<real, LAYER, FUSED_FMA>(a, i0, i1, j, k, a.p, a.vx, a.vy, a.vz, a.cell[0],
                                           a.face[0], a.cell[1], a.face[1], a.cell[2],
                                           a.face[2]);
}

template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__device__ __forceinline__ void pressure_march(
    const SweepArgs<real> &a, const int i0, const int i1, const int j, const int k,
    const int64_t step, const int64_t column, real *__restrict__ p, real *__restrict__ psi,
    const real *__restrict__ vx, const real *__restrict__ vy, const real *__restrict__ vz,
    const real *__restrict__ wall_table, const real *__restrict__ cell_x,
    const real *__restrict__ cell_y, const real *__restrict__ cell_z,
    real *__restrict__ recording) {
    const Layout L(a.nx, a.ny, a.nz, j, k);
    const int64_t n_cells = L.plane * a.nx;
    const int jk = j * a.nz + k;
    const int code_yz = WALLS ? (wall_class(j, a.ny) << 2) | (wall_class(k, a.nz) << 4) : 0;
    real fyc = 1, fzc = 1;
    if (LAYER) {
        fyc = cell_y[j];
        fzc = cell_z[k];
    }

    real vx_lo_next = vx[L.cell(i0)];
#pragma unroll ACFDTD_X_UNROLL
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real vx_lo = vx_lo_next;
        const real vx_hi = vx[c + L.plane];
        vx_lo_next = vx_hi;
        const int64_t cy = L.face_y(i);
        const int64_t cz = L.face_z(i);
        const real div = ((vx_hi - vx_lo) + (vy[cy + a.nz] - vy[cy])) + (vz[cz + 1] - vz[cz]);

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

template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(512) void pressure_sweep(const SweepArgs<real> a, const int64_t step,
                                                      const int64_t column) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    pressure_march<real, NR, WALLS, LAYER, COMPENSATED>(
        a, i0, i1, j, k, step, column, a.p, a.psi, a.vx, a.vy, a.vz, a.wall_table, a.cell[0],
        a.cell[1], a.cell[2], a.recording);
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

template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__device__ __forceinline__ void fused_march(
    const SweepArgs<real> &a, const int i0, const int i1, const int j, const int k,
    const int64_t step, const int64_t column, const real *__restrict__ p_in,
    const real *__restrict__ vx_in, const real *__restrict__ vy_in,
    const real *__restrict__ vz_in, real *__restrict__ p_out, real *__restrict__ vx_out,
    real *__restrict__ vy_out, real *__restrict__ vz_out, real *__restrict__ psi,
    const real *__restrict__ wall_table, const real *__restrict__ cell_x,
    const real *__restrict__ face_x, const real *__restrict__ cell_y,
    const real *__restrict__ face_y, const real *__restrict__ cell_z,
    const real *__restrict__ face_z, real *__restrict__ recording) {
    const Layout L(a.nx, a.ny, a.nz, j, k);
    const int64_t n_cells = L.plane * a.nx;
    const int jk = j * a.nz + k;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const real vel = a.vel;
    const int code_yz = WALLS ? (wall_class(j, ny) << 2) | (wall_class(k, nz) << 4) : 0;
    real fyc = 1, fzc = 1, fyf_lo = 1, fyf_hi = 1, fzf_lo = 1, fzf_hi = 1;
    if (LAYER) {
        fyc = cell_y[j];
        fzc = cell_z[k];
        fyf_lo = face_y[j];
        fyf_hi = face_y[j + 1];
        fzf_lo = face_z[k];
        fzf_hi = face_z[k + 1];
    }

    // New vx on face i0, the lower face of the first cell of the march.
    real p_c = p_in[L.cell(i0)];
    real vx_lo;
    {
        const real fxf = LAYER ? face_x[i0] : real(1);
        real v = vx_in[L.cell(i0)];
        if (i0 > 0) v = v - vel * (p_c - p_in[L.cell(i0 - 1)]);
        vx_lo = damp3<real, LAYER>(v, fxf, fyc, fzc);
        vx_out[L.cell(i0)] = vx_lo;
    }

#pragma unroll ACFDTD_X_UNROLL
    for (int i = i0; i < i1; ++i) {
        const int64_t c = L.cell(i);
        const real fxc = LAYER ? cell_x[i] : real(1);

        // x: upper face i + 1 (interior, or the far wall when i is the last cell).
        real vx_hi;
        real p_up = 0;
        {
            const real fxf = LAYER ? face_x[i + 1] : real(1);
            real v = vx_in[c + L.plane];
            if (i + 1 < nx) {
                p_up = p_in[c + L.plane];
                v = v - vel * (p_up - p_c);
            }
            vx_hi = damp3<real, LAYER>(v, fxf, fyc, fzc);
            if (i + 1 < i1 || i + 1 == nx) vx_out[c + L.plane] = vx_hi;
        }
        // y: lower face j (owned), upper face j + 1 (written only for the far wall).
        const int64_t cy = L.face_y(i);
        real vy_lo, vy_hi;
        {
            real v = vy_in[cy];
            if (j > 0) v = v - vel * (p_c - p_in[c - nz]);
            vy_lo = damp3<real, LAYER>(v, fxc, fyf_lo, fzc);
            vy_out[cy] = vy_lo;
            real w = vy_in[cy + nz];
            if (j + 1 < ny) w = w - vel * (p_in[c + nz] - p_c);
            vy_hi = damp3<real, LAYER>(w, fxc, fyf_hi, fzc);
            if (j + 1 == ny) vy_out[cy + nz] = vy_hi;
        }
        // z: lower face k (owned), upper face k + 1 (written only for the far wall).
        const int64_t cz = L.face_z(i);
        real vz_lo, vz_hi;
        {
            real v = vz_in[cz];
            if (k > 0) v = v - vel * (p_c - p_in[c - 1]);
            vz_lo = damp3<real, LAYER>(v, fxc, fyc, fzf_lo);
            vz_out[cz] = vz_lo;
            real w = vz_in[cz + 1];
            if (k + 1 < nz) w = w - vel * (p_in[c + 1] - p_c);
            vz_hi = damp3<real, LAYER>(w, fxc, fyc, fzf_hi);
            if (k + 1 == nz) vz_out[cz + 1] = vz_hi;
        }

        const real div = ((vx_hi - vx_lo) + (vy_hi - vy_lo)) + (vz_hi - vz_lo);
        const real p_old = p_c;
        real pn;
        if (COMPENSATED) {
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
            const int code = (wall_class(j, ny) << 2) | (wall_class(k, nz) << 4) |
                             wall_class(i, nx);
            if (code)
                tape_store(a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride,
                           TapePlanes(nx, ny, nz), nx, ny, nz, i, j, k, code, pn + p_old);
        }
        if (WALLS) {
            const int code = code_yz | wall_class(i, nx);
            if ((a.wall_mask >> code) & 1ull)
                pn = pn * wall_table[code] - p_old * wall_table[ACFDTD_HIP_WALL_CODES + code];
        }
        if (LAYER) pn = ((pn * fxc) * fyc) * fzc;
        for (int s = 0; s < a.n_sources; ++s) {
            if (a.source_cell[2 * s] == i && a.source_cell[2 * s + 1] == jk &&
                step < a.source_length[s])
                pn = pn + a.src * a.source_signal[s * a.source_stride + step];
        }
        p_out[c] = pn;
        for (int r = 0; r < a.n_receivers; ++r) {
            if (a.receiver_cell[2 * r] == i && a.receiver_cell[2 * r + 1] == jk)
                recording[r * a.recording_stride + column] = pn;
        }
        vx_lo = vx_hi;
        p_c = p_up;
    }
}

template <typename real, int NR, bool WALLS, bool LAYER, bool COMPENSATED>
__global__ __launch_bounds__(512) void fused_step(const SweepArgs<real> a, const FusedIO<real> io,
                                                  const int64_t step, const int64_t column) {
    int kb, jb, xc;
    if (!decode_tile(a.tiles, kb, jb, xc)) return;
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    if (k >= a.nz || j >= a.ny) return;
    const int i0 = xc * a.tiles.chunk;
    const int i1 = min(i0 + a.tiles.chunk, a.nx);
    fused_march<real, NR, WALLS, LAYER, COMPENSATED>(
        a, i0, i1, j, k, step, column, io.p_in, io.vx_in, io.vy_in, io.vz_in, io.p_out,
        io.vx_out, io.vy_out, io.vz_out, a.psi, a.wall_table, a.cell[0], a.face[0], a.cell[1],
        a.face[1], a.cell[2], a.face[2], a.recording);
}

// Energy terms: per-workgroup fp64 partials of (sum p^2, sum v^2, sum v (p_hi - p_lo)).
template <typename real>
__global__ __launch_bounds__(512) void energy_partials(const SweepArgs<real> a, double *partials) {
    double pp = 0, vv = 0, cp = 0;
    int kb, jb, xc;
    const bool active = decode_tile(a.tiles, kb, jb, xc);
    if (active) {
        const int k = kb * 64 + (int)threadIdx.x;
        const int j = jb * (int)blockDim.y + (int)threadIdx.y;
        if (k < a.nz && j < a.ny) {
            const int i0 = xc * a.tiles.chunk;
            const int i1 = min(i0 + a.tiles.chunk, a.nx);
            const Layout L(a.nx, a.ny, a.nz, j, k);
            double p_prev = i0 > 0 ? (double)a.p[L.cell(i0 - 1)] : 0.0;
            for (int i = i0; i < i1; ++i) {
                const int64_t c = L.cell(i);
                const double pc = a.p[c];
                pp += pc * pc;
                const double vx = a.vx[c];
                vv += vx * vx;
                if (i > 0) cp += vx * (pc - p_prev);
                const int64_t cy = L.face_y(i);
                const double vy = a.vy[cy];
                vv += vy * vy;
                if (j > 0) cp += vy * (pc - (double)a.p[c - a.nz]);
                const int64_t cz = L.face_z(i);
                const double vz = a.vz[cz];
                vv += vz * vz;
                if (k > 0) cp += vz * (pc - (double)a.p[c - 1]);
                if (i == a.nx - 1) {
                    const double t = a.vx[c + L.plane];
                    vv += t * t;
                }
                if (j == a.ny - 1) {
                    const double t = a.vy[cy + a.nz];
                    vv += t * t;
                }
                if (k == a.nz - 1) {
                    const double t = a.vz[cz + 1];
                    vv += t * t;
                }
                p_prev = pc;
            }
        }
    }
    __shared__ double lds[3][8];
    block_sum3(pp, vv, cp, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        partials[3 * blockIdx.x + 0] = pp;
        partials[3 * blockIdx.x + 1] = vv;
        partials[3 * blockIdx.x + 2] = cp;
    }
}

__global__ __launch_bounds__(512) void energy_finish(const double *partials, int n, double *out) {
    double s[3] = {0, 0, 0};
    const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    const int nt = (int)(blockDim.x * blockDim.y);
    for (int b = t; b < n; b += nt)
        for (int q = 0; q < 3; ++q) s[q] += partials[3 * b + q];
    __shared__ double lds[3][8];
    block_sum3(s[0], s[1], s[2], lds);
    if (t == 0) {
        out[0] = s[0];
        out[1] = s[1];
        out[2] = s[2];
    }
}

// ------------------------------------------------------------------------------- dispatch --

template <typename real>
using PressureKernel = void (*)(const SweepArgs<real>, const int64_t, const int64_t);
template <typename real>
using VelocityKernel = void (*)(const SweepArgs<real>);

template <typename real, int NR>
PressureKernel<real> pick_pressure_nr(bool walls, bool layer, bool comp) {
    const int key = (walls ? 1 : 0) | (layer ? 2 : 0) | (comp ? 4 : 0);
    switch (key) {
        case 0: return pressure_sweep<real, NR, false, false, false>;
        case 1: return pressure_sweep<real, NR, true, false, false>;
        case 2: return pressure_sweep<real, NR, false, true, false>;
        case 3: return pressure_sweep<real, NR, true, true, false>;
        case 4: return pressure_sweep<real, NR, false, false, true>;
        case 5: return pressure_sweep<real, NR, true, false, true>;
        case 6: return pressure_sweep<real, NR, false, true, true>;
        default: return pressure_sweep<real, NR, true, true, true>;
    }
}

template <typename real>
PressureKernel<real> pick_pressure(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_pressure_nr<real, 0>(walls, layer, comp);
        case 1: return pick_pressure_nr<real, 1>(walls, layer, comp);
        case 2: return pick_pressure_nr<real, 2>(walls, layer, comp);
        case 3: return pick_pressure_nr<real, 3>(walls, layer, comp);
        default: return pick_pressure_nr<real, 4>(walls, layer, comp);
    }
}

template <typename real>
VelocityKernel<real> pick_velocity(bool layer, bool comp) {
    if (layer) return comp ? velocity_sweep<real, true, true> : velocity_sweep<real, true, false>;
    return comp ? velocity_sweep<real, false, true> : velocity_sweep<real, false, false>;
}

template <typename real>
using FusedKernel = void (*)(const SweepArgs<real>, const FusedIO<real>, const int64_t,
                             const int64_t);

template <typename real, int NR>
FusedKernel<real> pick_fused_nr(bool walls, bool layer, bool comp) {
    const int key = (walls ? 1 : 0) | (layer ? 2 : 0) | (comp ? 4 : 0);
    switch (key) {
        case 0: return fused_step<real, NR, false, false, false>;
        case 1: return fused_step<real, NR, true, false, false>;
        case 2: return fused_step<real, NR, false, true, false>;
        case 3: return fused_step<real, NR, true, true, false>;
        case 4: return fused_step<real, NR, false, false, true>;
        case 5: return fused_step<real, NR, true, false, true>;
        case 6: return fused_step<real, NR, false, true, true>;
        default: return fused_step<real, NR, true, true, true>;
    }
}

template <typename real>
FusedKernel<real> pick_fused(int nr, bool walls, bool layer, bool comp) {
    switch (nr) {
        case 0: return pick_fused_nr<real, 0>(walls, layer, comp);
        case 1: return pick_fused_nr<real, 1>(walls, layer, comp);
        case 2: return pick_fused_nr<real, 2>(walls, layer, comp);
        case 3: return pick_fused_nr<real, 3>(walls, layer, comp);
        default: return pick_fused_nr<real, 4>(walls, layer, comp);
    }
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

template <typename real>
int run_steps(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps, int64_t column,
              hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const bool walls = d.wall_mask != 0, layer = d.has_layer != 0;
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    int64_t n = 0;
    if (has_swap(d)) {
        const FusedKernel<real> fk = pick_fused<real>(d.n_relax, walls, layer, comp);
        const FusedIO<real> forward = fused_io<real>(d, false), backward = fused_io<real>(d, true);
        for (; n + 1 < n_steps; n += 2) {
            hipLaunchKernelGGL(fk, grid, block, 0, stream, a, forward, step_offset + n, column + n);
            hipLaunchKernelGGL(fk, grid, block, 0, stream, a, backward, step_offset + n + 1,
                               column + n + 1);
        }
    }
    const VelocityKernel<real> vk = pick_velocity<real>(layer, comp);
    const PressureKernel<real> pk = pick_pressure<real>(d.n_relax, walls, layer, comp);
    for (; n < n_steps; ++n) {
        hipLaunchKernelGGL(vk, grid, block, 0, stream, a);
        hipLaunchKernelGGL(pk, grid, block, 0, stream, a, step_offset + n, column + n);
    }
    return check_hip(hipGetLastError(), "kernel launch");
}

template <typename real>
int fused_once(const acfdtd_hip_desc &d, int64_t step, int64_t column, bool reverse,
               hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const FusedKernel<real> fk = pick_fused<real>(d.n_relax, d.wall_mask != 0,
                                                  d.has_layer != 0, comp);
    hipLaunchKernelGGL(fk, dim3(a.tiles.launched), dim3(64, a.tiles.tile_j), 0, stream, a,
                       fused_io<real>(d, reverse), step, column);
    return check_hip(hipGetLastError(), "fused launch");
}

template <typename real>
int single_sweep(const acfdtd_hip_desc &d, bool pressure, int64_t step, int64_t column,
                 hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    if (pressure) {
        const PressureKernel<real> pk = pick_pressure<real>(d.n_relax, d.wall_mask != 0,
                                                            d.has_layer != 0, comp);
        hipLaunchKernelGGL(pk, grid, block, 0, stream, a, step, column);
    } else {
        hipLaunchKernelGGL(pick_velocity<real>(d.has_layer != 0, comp), grid, block, 0, stream, a);
    }
    return check_hip(hipGetLastError(), "sweep launch");
}

template <typename real>
int energy_terms(const acfdtd_hip_desc &d, void *workspace, double *out, hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    double *partials = static_cast<double *>(workspace);
    hipLaunchKernelGGL(energy_partials<real>, dim3(a.tiles.launched), dim3(64, a.tiles.tile_j), 0,
                       stream, a, partials);
    hipLaunchKernelGGL(energy_finish, dim3(1), dim3(64, 8), 0, stream, partials,
                       (int)a.tiles.launched, out);
    return check_hip(hipGetLastError(), "energy launch");
}

}  // namespace acfdtd

// ---------------------------------------------------------------------------------- C-ABI --

using namespace acfdtd;

extern "C" size_t acfdtd_hip_struct_size(void) { return sizeof(acfdtd_hip_desc); }

extern "C" int acfdtd_hip_abi_version(void) { return ACFDTD_HIP_ABI_VERSION; }

extern "C" int acfdtd_hip_supports_real(int real_bytes) {
    return real_bytes == 4 || real_bytes == 8;
}

extern "C" const char *acfdtd_hip_last_error(void) { return g_last_error.c_str(); }

extern "C" int acfdtd_hip_run(const acfdtd_hip_desc *desc, int64_t step_offset, int64_t n_steps,
                              int64_t recording_column, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + n_steps > desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording is too short for this many steps");
    if (desc->wall_tape && (step_offset < desc->wall_tape_step0 ||
                            step_offset + n_steps - desc->wall_tape_step0 > desc->wall_tape_rows))
        return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover these steps");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8)
        return run_steps<double>(*desc, step_offset, n_steps, recording_column, s);
    return run_steps<float>(*desc, step_offset, n_steps, recording_column, s);
}

extern "C" int acfdtd_hip_velocity_sweep(const acfdtd_hip_desc *desc, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return single_sweep<double>(*desc, false, 0, 0, s);
    return single_sweep<float>(*desc, false, 0, 0, s);
}

extern "C" int acfdtd_hip_pressure_sweep(const acfdtd_hip_desc *desc, int64_t step,
                                         int64_t recording_column, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (step < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "step and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column >= desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording column out of range");
    if (desc->wall_tape &&
        (step < desc->wall_tape_step0 || step - desc->wall_tape_step0 >= desc->wall_tape_rows))
        return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover this step");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return single_sweep<double>(*desc, true, step, recording_column, s);
    return single_sweep<float>(*desc, true, step, recording_column, s);
}

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
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return fused_once<double>(*desc, step, recording_column, reverse != 0, s);
    return fused_once<float>(*desc, step, recording_column, reverse != 0, s);
}

extern "C" int64_t acfdtd_hip_wall_tape_stride(int32_t nx, int32_t ny, int32_t nz) {
    return TapePlanes(nx, ny, nz).stride;
}

extern "C" size_t acfdtd_hip_energy_workspace_bytes(const acfdtd_hip_desc *desc) {
    if (validate(desc) != ACFDTD_HIP_OK) return 0;
    return (size_t)tiling_for(*desc).launched * 3 * sizeof(double);
}

extern "C" int acfdtd_hip_energy(const acfdtd_hip_desc *desc, void *workspace, double *out,
                                 void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (workspace == nullptr || out == nullptr)
        return fail(ACFDTD_HIP_EINVAL, "energy needs a workspace and an output buffer");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return energy_terms<double>(*desc, workspace, out, s);
    return energy_terms<float>(*desc, workspace, out, s);
}
