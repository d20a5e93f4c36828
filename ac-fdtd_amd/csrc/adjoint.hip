// The adjoint (transposed, time-reversed) leapfrog and the parameter gradients it yields.
//
// The step of sweeps.hip is linear in the state (p, v, psi) and in the source samples, so its
// transpose needs no forward state; only the gradients with respect to the coefficients do.
// With lam_X the adjoint of X, one forward step n transposes to (derivation in DESIGN.md):
//
//   cells, local part of step n  (reverse order of the forward pressure update)
//     lam_p'' = lam_p + G[:, n] at receivers           recording r_n = p^{n+1}[receiver]
//     grad_signal[s, n] = src_coef * lam_p''           p^{n+1} = p'' + src_coef * signal[n]
//     lam_hat = P * lam_p''                            p'' = P * p_hat      (sponge, cells)
//     lam_til = fu * lam_hat, lam_p = -fp * lam_hat    p_hat = fu * p_til - fp * p^n (walls)
//     lam_p += lam_til                                 p_til = p^n + ...
//     lam_psi_r <- mean_r lam_til + decay_r lam_psi_r
//     lam_div = lam_til (relax_scale - pres - sum mean_r gain_r)
//             + sum_r gain_r (1 - decay_r) lam_psi_r(old)
//   faces   lam_v <- D * (lam_v + lam_div[lower cell] - lam_div[upper cell])     (v' = D * v~)
//   cells, stencil part of step n
//     lam_p += vel * sum_a (lam_v[upper interior face] - lam_v[lower interior face])
//
// The stencil part of step n and the local part of step n - 1 touch only the cell itself, so
// they run in one sweep: two kernels per adjoint step, the same shape and traffic as the
// forward two-sweep step. Parameter gradients, accumulated in fp64:
//   walls   dL/da_c  = -lam_hat fu^2 (p_til + p^n)      (p_til + p^n from the boundary tape)
//   sponge  dL/dsigma += -dt W lam_p'' p''  (cells),  -dt W lam_tot v'  (faces),
//           W the sum of the grading weights inside^order of the factors in P or D
//   air     dL/dgain_r += div (-mean_r lam_til + (1 - decay_r) lam_psi_r(old)),
//           dL/drelax_scale += div lam_til              (div of the forward v^{n+1/2})
// Each workgroup adds its sums to its own slot of `partials` once per launch, so the totals
// are deterministic; the host reduces the slots and applies the chain rule to the physical
// parameters.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
#include "adjoint.hpp"
#include "adjoint2.hpp"
#include "marches2.hpp"
#include "sweeps.hpp"

namespace acfdtd {

template <typename real, bool LAYER>
__global__ __launch_bounds__(512) void adjoint_face_sweep(const SweepArgs<real> a,
                                                          const AdjointArgs<real> g) {
    double acc[kGradSlots] = {};
    int kb = 0, jb = 0, xc = 0, slot = 0;
    const bool tile = decode_tile_slot(a.tiles, kb, jb, xc, slot);
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    const bool active = tile && k < a.nz && j < a.ny;
    const bool sponge = LAYER && (g.grads & ACFDTD_HIP_GRAD_SPONGE) && g.fwd_vx;
    if (active) {
        const int i0 = xc * a.tiles.chunk;
        const int i1 = min(i0 + a.tiles.chunk, a.nx);
        const Layout L(a.nx, a.ny, a.nz, j, k);
        real fyc = 1, fzc = 1, fyf = 1, fzf = 1, wyc = 0, wzc = 0, wyf = 0, wzf = 0;
        if (LAYER) {
            fyc = a.cell[1][j];
            fzc = a.cell[2][k];
            fyf = a.face[1][j];
            fzf = a.face[2][k];
        }
        if (sponge) {
            wyc = g.w_cell[1][j];
            wzc = g.w_cell[2][k];
            wyf = g.w_face[1][j];
            wzf = g.w_face[2][k];
        }
        const double mdt = -(double)g.dt;
        real ld_prev = i0 > 0 ? g.ldiv[L.cell(i0 - 1)] : real(0);
        for (int i = i0; i < i1; ++i) {
            const int64_t c = L.cell(i);
            const real ld = g.ldiv[c];
            real fxc = 1, fxf = 1, wxc = 0, wxf = 0;
            if (LAYER) {
                fxc = a.cell[0][i];
                fxf = a.face[0][i];
            }
            if (sponge) {
                wxc = g.w_cell[0][i];
                wxf = g.w_face[0][i];
            }
            {  // x face i: lower cell i - 1 (none at the wall), upper cell i
                const real t = g.lvx[c] + ((i > 0 ? ld_prev : real(0)) - ld);
                if (sponge) acc[kGradSponge] += mdt * (double)(wxf + wyc + wzc) * t * g.fwd_vx[c];
                g.lvx[c] = LAYER ? ((t * fxf) * fyc) * fzc : t;
            }
            const int64_t cy = L.face_y(i);
            {
                const real t = g.lvy[cy] + ((j > 0 ? g.ldiv[c - a.nz] : real(0)) - ld);
                if (sponge) acc[kGradSponge] += mdt * (double)(wxc + wyf + wzc) * t * g.fwd_vy[cy];
                g.lvy[cy] = LAYER ? ((t * fxc) * fyf) * fzc : t;
            }
            const int64_t cz = L.face_z(i);
            {
                const real t = g.lvz[cz] + ((k > 0 ? g.ldiv[c - 1] : real(0)) - ld);
                if (sponge) acc[kGradSponge] += mdt * (double)(wxc + wyc + wzf) * t * g.fwd_vz[cz];
                g.lvz[cz] = LAYER ? ((t * fxc) * fyc) * fzf : t;
            }
            // Far wall faces: only the lower cell exists.
            if (i == a.nx - 1) {
                const int64_t f = c + L.plane;
                const real t = g.lvx[f] + ld;
                if (sponge)
                    acc[kGradSponge] +=
                        mdt * (double)(g.w_face[0][a.nx] + wyc + wzc) * t * g.fwd_vx[f];
                g.lvx[f] = LAYER ? ((t * a.face[0][a.nx]) * fyc) * fzc : t;
            }
            if (j == a.ny - 1) {
                const int64_t f = cy + a.nz;
                const real t = g.lvy[f] + ld;
                if (sponge)
                    acc[kGradSponge] +=
                        mdt * (double)(wxc + g.w_face[1][a.ny] + wzc) * t * g.fwd_vy[f];
                g.lvy[f] = LAYER ? ((t * fxc) * a.face[1][a.ny]) * fzc : t;
            }
            if (k == a.nz - 1) {
                const int64_t f = cz + 1;
                const real t = g.lvz[f] + ld;
                if (sponge)
                    acc[kGradSponge] +=
                        mdt * (double)(wxc + wyc + g.w_face[2][a.nz]) * t * g.fwd_vz[f];
                g.lvz[f] = LAYER ? ((t * fxc) * fyc) * a.face[2][a.nz] : t;
            }
            ld_prev = ld;
        }
    }
    if (sponge) flush_partials<kGradSponge, 1>(acc, tile, slot, g.partials);
}

template <typename real, int NR, bool LAYER>
__global__ __launch_bounds__(512) void adjoint_cell_sweep(const SweepArgs<real> a,
                                                          const AdjointArgs<real> g,
                                                          const int64_t step, const int stencil,
                                                          const int local) {
    double acc[kGradSlots] = {};
    int kb = 0, jb = 0, xc = 0, slot = 0;
    const bool tile = decode_tile_slot(a.tiles, kb, jb, xc, slot);
    const int k = kb * 64 + (int)threadIdx.x;
    const int j = jb * (int)blockDim.y + (int)threadIdx.y;
    const bool active = tile && k < a.nz && j < a.ny;
    const bool walls = local && (g.grads & ACFDTD_HIP_GRAD_WALLS) && a.wall_tape;
    const bool sponge = local && LAYER && (g.grads & ACFDTD_HIP_GRAD_SPONGE) && g.fwd_p;
    // Air gradients: the divergence of step `step` from the divergence tape, else from the
    // forward velocities v^{step+1/2}.
    const bool air = local && NR > 0 && (g.grads & ACFDTD_HIP_GRAD_AIR) && (a.div_tape || g.fwd_vx);
    if (active) {
        const int i0 = xc * a.tiles.chunk;
        const int i1 = min(i0 + a.tiles.chunk, a.nx);
        const Layout L(a.nx, a.ny, a.nz, j, k);
        const int64_t n_cells = L.plane * a.nx;
        const real *div_row =
            a.div_tape ? a.div_tape + (step - a.div_tape_step0) * n_cells : nullptr;
        const int jk = j * a.nz + k;
        const int code_yz = (wall_class(j, a.ny) << 2) | (wall_class(k, a.nz) << 4);
        real fyc = 1, fzc = 1, wyz = 0;
        if (LAYER) {
            fyc = a.cell[1][j];
            fzc = a.cell[2][k];
        }
        if (sponge) wyz = g.w_cell[1][j] + g.w_cell[2][k];
        real mg = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) mg += a.mean[r] * a.gain[r];
        const real cdiv = (a.relax_scale - a.pres) - mg;
        const TapePlanes planes(a.nx, a.ny, a.nz);
        const real *tape_row =
            walls ? a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride : nullptr;

        for (int i = i0; i < i1; ++i) {
            const int64_t c = L.cell(i);
            real lpc = g.lp[c];
            if (stencil) {
                const real lo_x = i > 0 ? g.lvx[c] : real(0);
                const real up_x = i + 1 < a.nx ? g.lvx[c + L.plane] : real(0);
                const int64_t cy = L.face_y(i);
                const real lo_y = j > 0 ? g.lvy[cy] : real(0);
                const real up_y = j + 1 < a.ny ? g.lvy[cy + a.nz] : real(0);
                const int64_t cz = L.face_z(i);
                const real lo_z = k > 0 ? g.lvz[cz] : real(0);
                const real up_z = k + 1 < a.nz ? g.lvz[cz + 1] : real(0);
                lpc = lpc + a.vel * (((up_x - lo_x) + (up_y - lo_y)) + (up_z - lo_z));
            }
            if (!local) {
                g.lp[c] = lpc;
                continue;
            }
            // Local transpose of step `step`.
            for (int r = 0; r < a.n_receivers; ++r)
                if (a.receiver_cell[2 * r] == i && a.receiver_cell[2 * r + 1] == jk)
                    lpc = lpc + g.grad_recording[r * g.grad_recording_stride +
                                                 (step - g.grad_recording_step0)];
            real injected = 0;
            for (int s = 0; s < a.n_sources; ++s) {
                if (a.source_cell[2 * s] == i && a.source_cell[2 * s + 1] == jk &&
                    step < a.source_length[s]) {
                    if (g.grad_signal) g.grad_signal[s * a.source_stride + step] = a.src * lpc;
                    if (sponge) injected += a.src * a.source_signal[s * a.source_stride + step];
                }
            }
            const real lhat = LAYER ? ((lpc * a.cell[0][i]) * fyc) * fzc : lpc;
            if (sponge)
                acc[kGradSponge] += -(double)g.dt * (double)(g.w_cell[0][i] + wyz) *
                                    (double)lpc * (double)(g.fwd_p[c] - injected);
            real ltil = lhat, lp_new = 0;
            const int code = code_yz | wall_class(i, a.nx);
            if (code && a.wall_table) {
                const real fu = a.wall_table[code];
                const real fp = a.wall_table[ACFDTD_HIP_WALL_CODES + code];
                ltil = fu * lhat;
                lp_new = -fp * lhat;
                if (walls) {
                    const double ga = -(double)lhat * (double)fu * (double)fu *
                                      (double)tape_load(tape_row, planes, a.ny, a.nz, i, j, k, code);
                    if (code & 1) acc[kGradWall + 0] += ga;
                    if (code & 2) acc[kGradWall + 1] += ga;
                    if (code & 4) acc[kGradWall + 2] += ga;
                    if (code & 8) acc[kGradWall + 3] += ga;
                    if (code & 16) acc[kGradWall + 4] += ga;
                    if (code & 32) acc[kGradWall + 5] += ga;
                }
            }
            lp_new = lp_new + ltil;
            real ldv = ltil * cdiv;
            double div = 0;
            if (air) {
                if (div_row) {
                    div = (double)div_row[c];
                } else {
                    const int64_t cy = L.face_y(i), cz = L.face_z(i);
                    div = (double)(((g.fwd_vx[c + L.plane] - g.fwd_vx[c]) +
                                    (g.fwd_vy[cy + a.nz] - g.fwd_vy[cy])) +
                                   (g.fwd_vz[cz + 1] - g.fwd_vz[cz]));
                }
                acc[kGradScale] += div * (double)ltil;
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                real *state = g.lpsi + (int64_t)r * n_cells + c;
                const real old = *state;
                *state = a.mean[r] * ltil + a.decay[r] * old;
                ldv = ldv + old * (a.gain[r] * (real(1) - a.decay[r]));
                if (air)
                    acc[kGradGain + r] +=
                        div * (-(double)a.mean[r] * (double)ltil +
                               (1.0 - (double)a.decay[r]) * (double)old);
            }
            g.lp[c] = lp_new;
            g.ldiv[c] = ldv;
        }
    }
    if (walls || sponge || air) flush_partials(acc, tile, slot, g.partials);
}

template <typename real>
using FaceKernel = void (*)(const SweepArgs<real>, const AdjointArgs<real>);
template <typename real>
using CellKernel = void (*)(const SweepArgs<real>, const AdjointArgs<real>, const int64_t,
                            const int, const int);

template <typename real>
CellKernel<real> pick_cell(int nr, bool layer) {
    switch (nr) {
        case 0: return layer ? adjoint_cell_sweep<real, 0, true> : adjoint_cell_sweep<real, 0, false>;
        case 1: return layer ? adjoint_cell_sweep<real, 1, true> : adjoint_cell_sweep<real, 1, false>;
        case 2: return layer ? adjoint_cell_sweep<real, 2, true> : adjoint_cell_sweep<real, 2, false>;
        case 3: return layer ? adjoint_cell_sweep<real, 3, true> : adjoint_cell_sweep<real, 3, false>;
        default: return layer ? adjoint_cell_sweep<real, 4, true> : adjoint_cell_sweep<real, 4, false>;
    }
}

int validate_adjoint(const acfdtd_hip_desc *d, const acfdtd_hip_adjoint_desc *g) {
    const int status = validate(d);
    if (status != ACFDTD_HIP_OK) return status;
    if (g == nullptr) return fail(ACFDTD_HIP_EINVAL, "adjoint descriptor is NULL");
    if (g->abi_version != ACFDTD_HIP_ABI_VERSION || g->struct_size != sizeof(acfdtd_hip_adjoint_desc))
        return fail(ACFDTD_HIP_EABI, "adjoint descriptor ABI version or size mismatch");
    if (g->time_block < 0 || g->time_block > 2)
        return fail(ACFDTD_HIP_EINVAL, "adjoint time_block must be 0, 1 or 2");
    if (!g->lp || !g->lvx || !g->lvy || !g->lvz || !g->ldiv)
        return fail(ACFDTD_HIP_EINVAL, "adjoint field pointer is NULL");
    if (d->n_relax > 0 && !g->lpsi) return fail(ACFDTD_HIP_EINVAL, "adjoint relaxation states are NULL");
    if (d->n_receivers > 0 && !g->grad_recording)
        return fail(ACFDTD_HIP_EINVAL, "recording gradient is NULL");
    if ((g->grads & ~(ACFDTD_HIP_GRAD_WALLS | ACFDTD_HIP_GRAD_SPONGE | ACFDTD_HIP_GRAD_AIR)) != 0)
        return fail(ACFDTD_HIP_EINVAL, "unknown gradient flags");
    if (g->grads && !g->partials) return fail(ACFDTD_HIP_EINVAL, "gradient partials are NULL");
    if ((g->grads & ACFDTD_HIP_GRAD_WALLS) && (!d->wall_tape || !d->wall_table))
        return fail(ACFDTD_HIP_EINVAL, "wall gradients need the boundary tape and the wall table");
    if ((g->grads & ACFDTD_HIP_GRAD_SPONGE) && d->has_layer)
        for (int q = 0; q < 3; ++q)
            if (!g->weight_cell[q] || !g->weight_face[q])
                return fail(ACFDTD_HIP_EINVAL, "sponge gradients need the grading weights");
    return ACFDTD_HIP_OK;
}

template <typename real>
int face_sweep(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &gd, hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    AdjointArgs<real> g;
    fill_adjoint(gd, g);
    const FaceKernel<real> fk =
        d.has_layer ? adjoint_face_sweep<real, true> : adjoint_face_sweep<real, false>;
    hipLaunchKernelGGL(fk, dim3(a.tiles.launched), dim3(64, a.tiles.tile_j), 0, stream, a, g);
    return check_hip(hipGetLastError(), "adjoint face sweep launch");
}

template <typename real>
int cell_sweep(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &gd, int64_t step,
               int stencil, int local, hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    AdjointArgs<real> g;
    fill_adjoint(gd, g);
    hipLaunchKernelGGL(pick_cell<real>(d.n_relax, d.has_layer != 0), dim3(a.tiles.launched),
                       dim3(64, a.tiles.tile_j), 0, stream, a, g, step, stencil, local);
    return check_hip(hipGetLastError(), "adjoint cell sweep launch");
}

// One adjoint step m by the face and cell sweeps, forward states from the ring (if any).
template <typename real>
int two_sweep_adjoint_step(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &gd, int64_t m,
                           hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    AdjointArgs<real> g;
    fill_adjoint(gd, g);
    const bool ring = gd.fwd_count > 0;
    const dim3 grid(a.tiles.launched), block(64, a.tiles.tile_j);
    ForwardState<real> after{}, before{};
    if (ring) {
        after = forward_slot<real>(d, gd, m + 1 - gd.fwd_step0);
        before = forward_slot<real>(d, gd, m - gd.fwd_step0);
    }
    g.fwd_p = after.p;
    g.fwd_vx = after.vx;
    g.fwd_vy = after.vy;
    g.fwd_vz = after.vz;
    const FaceKernel<real> fk =
        d.has_layer ? adjoint_face_sweep<real, true> : adjoint_face_sweep<real, false>;
    hipLaunchKernelGGL(fk, grid, block, 0, stream, a, g);
    g.fwd_p = before.p;
    g.fwd_vx = before.vx;
    g.fwd_vy = before.vy;
    g.fwd_vz = before.vz;
    hipLaunchKernelGGL(pick_cell<real>(d.n_relax, d.has_layer != 0), grid, block, 0, stream, a, g,
                       m - 1, 1, m >= 1 ? 1 : 0);
    return check_hip(hipGetLastError(), "adjoint sweep launch");
}

template <typename real>
int adjoint_run(const acfdtd_hip_desc &d, const acfdtd_hip_adjoint_desc &g, int64_t step_hi,
                int64_t step_lo, hipStream_t stream) {
    const int64_t n_steps = step_hi - step_lo + 1;
    int64_t done = 0;
    if (has_adjoint_swap(g)) {
        done = launch_adjoint2_quads<real>(d, g, step_hi, n_steps, stream);
        int status = check_hip(hipGetLastError(), "two-step adjoint launch");
        if (status != ACFDTD_HIP_OK) return status;
        done += launch_adjoint_pairs<real>(d, g, step_hi - done, n_steps - done, stream);
        status = check_hip(hipGetLastError(), "fused adjoint launch");
        if (status != ACFDTD_HIP_OK) return status;
    }
    for (int64_t m = step_hi - done; m >= step_lo; --m) {
        const int status = two_sweep_adjoint_step<real>(d, g, m, stream);
        if (status != ACFDTD_HIP_OK) return status;
    }
    return ACFDTD_HIP_OK;
}

}  // namespace acfdtd

using namespace acfdtd;

extern "C" size_t acfdtd_hip_adjoint_partials_count(const acfdtd_hip_desc *desc) {
    if (validate(desc) != ACFDTD_HIP_OK) return 0;
    const Tiling t = tiling_for(*desc);
    int64_t waves = std::max<int64_t>((int64_t)t.tiles * t.tile_j, adjoint_fused_waves(*desc));
    Tiling2 t2;
    if (make_tiling2(*desc, desc->real_bytes == 8 ? 2 : 4, t2)) {
        // The two-step adjoint runs balanced chunks (launch_adjoint2_quads): as many tiles or
        // more.
        balance_tiling2(t2, desc->nx);
        waves = std::max<int64_t>(waves, (int64_t)t2.tiles * (t2.threads / 64));
    }
    return (size_t)waves * kGradSlots;
}

extern "C" int acfdtd_hip_adjoint_face_sweep(const acfdtd_hip_desc *desc,
                                             const acfdtd_hip_adjoint_desc *adjoint,
                                             void *stream) {
    const int status = validate_adjoint(desc, adjoint);
    if (status != ACFDTD_HIP_OK) return status;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return face_sweep<double>(*desc, *adjoint, s);
    return face_sweep<float>(*desc, *adjoint, s);
}

extern "C" int acfdtd_hip_adjoint_cell_sweep(const acfdtd_hip_desc *desc,
                                             const acfdtd_hip_adjoint_desc *adjoint, int64_t step,
                                             int stencil, int local, void *stream) {
    const int status = validate_adjoint(desc, adjoint);
    if (status != ACFDTD_HIP_OK) return status;
    if (local) {
        if (step < 0) return fail(ACFDTD_HIP_EINVAL, "step must be >= 0");
        if (desc->n_receivers > 0 && (step < adjoint->grad_recording_step0 ||
                                      step - adjoint->grad_recording_step0 >=
                                          adjoint->grad_recording_stride))
            return fail(ACFDTD_HIP_EINVAL, "recording gradient does not cover this step");
        if ((adjoint->grads & ACFDTD_HIP_GRAD_WALLS) &&
            (step < desc->wall_tape_step0 || step - desc->wall_tape_step0 >= desc->wall_tape_rows))
            return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover this step");
        if ((adjoint->grads & ACFDTD_HIP_GRAD_AIR) && !div_tape_covers(*desc, step, step))
            return fail(ACFDTD_HIP_EINVAL, "divergence tape does not cover this step");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8)
        return cell_sweep<double>(*desc, *adjoint, step, stencil != 0, local != 0, s);
    return cell_sweep<float>(*desc, *adjoint, step, stencil != 0, local != 0, s);
}

extern "C" int acfdtd_hip_adjoint_two_step(const acfdtd_hip_desc *desc,
                                          const acfdtd_hip_adjoint_desc *adjoint) {
    if (validate_adjoint(desc, adjoint) != ACFDTD_HIP_OK) return 0;
    Tiling2 t2;
    return adjoint2_applies(*desc, *adjoint) &&
           make_tiling2(*desc, desc->real_bytes == 8 ? 2 : 4, t2);
}

extern "C" int acfdtd_hip_adjoint_run(const acfdtd_hip_desc *desc,
                                      const acfdtd_hip_adjoint_desc *adjoint, int64_t step_hi,
                                      int64_t step_lo, void *stream) {
    const int status = validate_adjoint(desc, adjoint);
    if (status != ACFDTD_HIP_OK) return status;
    if (step_lo < 0 || step_hi < step_lo)
        return fail(ACFDTD_HIP_EINVAL, "need 0 <= step_lo <= step_hi");
    // Steps whose local part runs: m - 1 for m in [max(step_lo, 1), step_hi].
    const int64_t local_lo = step_lo > 0 ? step_lo - 1 : 0, local_hi = step_hi - 1;
    if (local_hi >= local_lo) {
        if (desc->n_receivers > 0 &&
            (local_lo < adjoint->grad_recording_step0 ||
             local_hi - adjoint->grad_recording_step0 >= adjoint->grad_recording_stride))
            return fail(ACFDTD_HIP_EINVAL, "recording gradient does not cover these steps");
        if ((adjoint->grads & ACFDTD_HIP_GRAD_WALLS) &&
            (local_lo < desc->wall_tape_step0 ||
             local_hi - desc->wall_tape_step0 >= desc->wall_tape_rows))
            return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover these steps");
    }
    const bool sponge = (adjoint->grads & ACFDTD_HIP_GRAD_SPONGE) && desc->has_layer;
    const bool air = (adjoint->grads & ACFDTD_HIP_GRAD_AIR) && desc->n_relax > 0;
    if (adjoint->fwd_count < 0) return fail(ACFDTD_HIP_EINVAL, "fwd_count must be >= 0");
    // The sponge gradient reads the layer tape when the forward descriptor has one and no
    // forward states are given (two adjoint steps per sweep only), else the forward states; the
    // air gradients read the divergence tape when there is one, else the forward states.
    const bool sponge_tape = sponge && desc->layer_tape != nullptr && adjoint->fwd_count == 0;
    const bool states = (sponge && !sponge_tape) || (air && desc->div_tape == nullptr);
    if (sponge_tape) {
        Tiling2 t2;
        if (!adjoint2_applies(*desc, *adjoint) ||
            !make_tiling2(*desc, desc->real_bytes == 8 ? 2 : 4, t2))
            return fail(ACFDTD_HIP_EINVAL,
                        "the layer tape is read by the two-step adjoint only, which does not apply "
                        "to these descriptors (give the forward states instead)");
        if (step_lo < 1 || (step_hi - step_lo + 1) % 4 != 0)
            return fail(ACFDTD_HIP_EINVAL,
                        "with the layer tape, adjoint steps run in groups of four from step 1 on "
                        "(give the forward states for the others)");
        if (!layer_tape_covers(*desc, step_lo - 1, step_hi))
            return fail(ACFDTD_HIP_EINVAL, "layer tape does not cover these steps");
    }
    if (states) {
        if (adjoint->fwd_count == 0 || !adjoint->fwd_vx || !adjoint->fwd_vy || !adjoint->fwd_vz ||
            (sponge && !adjoint->fwd_p))
            return fail(ACFDTD_HIP_EINVAL,
                        "sponge gradients, and air gradients without a divergence tape, need the "
                        "forward states");
        if (step_lo < adjoint->fwd_step0 ||
            step_hi + 1 - adjoint->fwd_step0 >= adjoint->fwd_count)
            return fail(ACFDTD_HIP_EINVAL, "forward states do not cover these steps");
    }
    if (air && desc->div_tape && !div_tape_covers(*desc, local_lo, local_hi))
        return fail(ACFDTD_HIP_EINVAL, "divergence tape does not cover these steps");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return adjoint_run<double>(*desc, *adjoint, step_hi, step_lo, s);
    return adjoint_run<float>(*desc, *adjoint, step_hi, step_lo, s);
}
