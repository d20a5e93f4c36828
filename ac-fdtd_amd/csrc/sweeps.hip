// The time step of the staggered pressure-velocity scheme on gfx950: the two in-place sweeps,
// the energy reduction, and the single-room C-ABI (fused steps in fused.hip, batches in
// batch.hip, the adjoint in adjoint.hip).
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
#include <mutex>
#include <string>

#include "acfdtd_hip.h"
#include "marches.hpp"
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

// Compensated summation is an fp32 mode (validated on entry), so the fp64 library carries no
// compensated kernels.
template <typename real, int NR>
PressureKernel<real> pick_pressure_nr(bool walls, bool layer, bool comp) {
    if constexpr (sizeof(real) == 4) {
        if (comp) {
            if (walls) return layer ? pressure_sweep<real, NR, true, true, true>
                                    : pressure_sweep<real, NR, true, false, true>;
            return layer ? pressure_sweep<real, NR, false, true, true>
                         : pressure_sweep<real, NR, false, false, true>;
        }
    }
    (void)comp;
    if (walls) return layer ? pressure_sweep<real, NR, true, true, false>
                            : pressure_sweep<real, NR, true, false, false>;
    return layer ? pressure_sweep<real, NR, false, true, false>
                 : pressure_sweep<real, NR, false, false, false>;
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
    if constexpr (sizeof(real) == 4) {
        if (comp) return layer ? velocity_sweep<real, true, true> : velocity_sweep<real, false, true>;
    }
    (void)comp;
    return layer ? velocity_sweep<real, true, false> : velocity_sweep<real, false, false>;
}

template <typename real>
int inplace_steps(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps, int64_t column,
                  hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    const bool comp = (d.flags & ACFDTD_HIP_COMPENSATED) != 0;
    const bool walls = d.wall_mask != 0, layer = d.has_layer != 0;
    const dim3 block(64, a.tiles.tile_j);
    const dim3 grid(a.tiles.launched);
    const VelocityKernel<real> vk = pick_velocity<real>(layer, comp);
    const PressureKernel<real> pk = pick_pressure<real>(d.n_relax, walls, layer, comp);
    for (int64_t n = 0; n < n_steps; ++n) {
        hipLaunchKernelGGL(vk, grid, block, 0, stream, a);
        hipLaunchKernelGGL(pk, grid, block, 0, stream, a, step_offset + n, column + n);
    }
    return check_hip(hipGetLastError(), "kernel launch");
}

template int inplace_steps<float>(const acfdtd_hip_desc &, int64_t, int64_t, int64_t, hipStream_t);
template int inplace_steps<double>(const acfdtd_hip_desc &, int64_t, int64_t, int64_t,
                                   hipStream_t);

// Layer tape row `step` from the state fields (pack) or into them (unpack: zeros off the
// frame), one thread per cell. Owner cells move their lower faces; the last cell of each axis
// also its far face.
template <typename real, bool PACK>
__global__ __launch_bounds__(256) void layer_tape_move(const SweepArgs<real> a, real *row,
                                                       real *p, real *vx, real *vy, real *vz) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t n = (int64_t)nx * ny * nz;
    const LayerFrame &f = a.layer;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(c % nz);
        const int j = (int)((c / nz) % ny);
        const int i = (int)(c / ((int64_t)ny * nz));
        const int64_t q = f.offset(i, j, k, f.in_frame(j, k));
        const int64_t cy = ((int64_t)i * (ny + 1) + j) * nz + k;
        const int64_t cz = ((int64_t)i * ny + j) * (nz + 1) + k;
        if (PACK) {
            if (q >= 0) {
                row[q] = p[c];
                row[f.cells + q] = vx[c];
                row[2 * f.cells + q] = vy[cy];
                row[3 * f.cells + q] = vz[cz];
            }
            if (i == nx - 1) row[f.far_x() + (int64_t)j * nz + k] = vx[c + (int64_t)ny * nz];
            if (j == ny - 1) row[f.far_y() + (int64_t)i * nz + k] = vy[cy + nz];
            if (k == nz - 1) row[f.far_z() + (int64_t)i * ny + j] = vz[cz + 1];
        } else {
            p[c] = q >= 0 ? row[q] : real(0);
            vx[c] = q >= 0 ? row[f.cells + q] : real(0);
            vy[cy] = q >= 0 ? row[2 * f.cells + q] : real(0);
            vz[cz] = q >= 0 ? row[3 * f.cells + q] : real(0);
            if (i == nx - 1) vx[c + (int64_t)ny * nz] = row[f.far_x() + (int64_t)j * nz + k];
            if (j == ny - 1) vy[cy + nz] = row[f.far_y() + (int64_t)i * nz + k];
            if (k == nz - 1) vz[cz + 1] = row[f.far_z() + (int64_t)i * ny + j];
        }
    }
}

template <typename real, bool PACK>
int layer_tape_launch(const acfdtd_hip_desc &d, int64_t step, real *p, real *vx, real *vy,
                      real *vz, hipStream_t stream) {
    SweepArgs<real> a;
    fill_args(d, a);
    real *row = a.layer_tape + (step - a.layer_tape_step0) * a.layer_tape_stride;
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
    hipLaunchKernelGGL((layer_tape_move<real, PACK>), dim3((unsigned)blocks), dim3(256), 0,
                       stream, a, row, p, vx, vy, vz);
    return check_hip(hipGetLastError(), "layer tape launch");
}

template <typename real>
int run_steps(const acfdtd_hip_desc &d, int64_t step_offset, int64_t n_steps, int64_t column,
              hipStream_t stream) {
    int64_t n = 0;
    if (n_steps > 3) n = launch_fused2_quads<real>(d, step_offset, n_steps, column, stream);
    if (d.layer_tape) {
        // The two-step sweeps write their rows; the other steps run one by one in place, each
        // followed by its row copied out of the state.
        for (; n < n_steps; ++n) {
            int rc = inplace_steps<real>(d, step_offset + n, 1, column + n, stream);
            if (rc == ACFDTD_HIP_OK)
                rc = layer_tape_launch<real, true>(d, step_offset + n, static_cast<real *>(d.p),
                                                   static_cast<real *>(d.vx),
                                                   static_cast<real *>(d.vy),
                                                   static_cast<real *>(d.vz), stream);
            if (rc != ACFDTD_HIP_OK) return rc;
        }
        return check_hip(hipGetLastError(), "kernel launch");
    }
    if (has_swap(d) && n_steps - n > 1)
        n += launch_fused_pairs<real>(d, step_offset + n, n_steps - n, column + n, stream);
    if (n < n_steps) return inplace_steps<real>(d, step_offset + n, n_steps - n, column + n, stream);
    return check_hip(hipGetLastError(), "kernel launch");
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
    if (a.tiles.launched > INT32_MAX / 3) return fail(ACFDTD_HIP_EINVAL, "too many energy tiles");
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
    const int status = validate(desc, true);
    if (status != ACFDTD_HIP_OK) return status;
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + n_steps > desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording is too short for this many steps");
    if (desc->wall_tape && (step_offset < desc->wall_tape_step0 ||
                            step_offset + n_steps - desc->wall_tape_step0 > desc->wall_tape_rows))
        return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover these steps");
    if (!div_tape_covers(*desc, step_offset, step_offset + n_steps - 1))
        return fail(ACFDTD_HIP_EINVAL, "divergence tape does not cover these steps");
    if (!layer_tape_covers(*desc, step_offset, step_offset + n_steps - 1))
        return fail(ACFDTD_HIP_EINVAL, "layer tape does not cover these steps");
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
    if (desc->layer_tape) return fail(ACFDTD_HIP_EINVAL, "the layer tape is written by acfdtd_hip_run only");
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
    if (!div_tape_covers(*desc, step, step))
        return fail(ACFDTD_HIP_EINVAL, "divergence tape does not cover this step");
    if (desc->layer_tape) return fail(ACFDTD_HIP_EINVAL, "the layer tape is written by acfdtd_hip_run only");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8) return single_sweep<double>(*desc, true, step, recording_column, s);
    return single_sweep<float>(*desc, true, step, recording_column, s);
}

extern "C" int64_t acfdtd_hip_layer_tape_stride(const acfdtd_hip_desc *desc) {
    if (desc == nullptr || desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return -1;
    return LayerFrame(desc->nx, desc->ny, desc->nz, desc->layer_lo, desc->layer_hi).stride;
}

extern "C" int acfdtd_hip_layer_tape_unpack(const acfdtd_hip_desc *desc, int64_t step, void *p,
                                            void *vx, void *vy, void *vz, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (!desc->layer_tape) return fail(ACFDTD_HIP_EINVAL, "the descriptor has no layer tape");
    if (!p || !vx || !vy || !vz) return fail(ACFDTD_HIP_EINVAL, "field pointer is NULL");
    if (step < 0 || !layer_tape_covers(*desc, step, step))
        return fail(ACFDTD_HIP_EINVAL, "layer tape does not cover this step");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8)
        return layer_tape_launch<double, false>(*desc, step, static_cast<double *>(p),
                                                static_cast<double *>(vx),
                                                static_cast<double *>(vy),
                                                static_cast<double *>(vz), s);
    return layer_tape_launch<float, false>(*desc, step, static_cast<float *>(p),
                                           static_cast<float *>(vx), static_cast<float *>(vy),
                                           static_cast<float *>(vz), s);
}

extern "C" int64_t acfdtd_hip_wall_tape_stride(int32_t nx, int32_t ny, int32_t nz) {
    // Rounded up to whole 4-element groups, so that rows of a 16-byte aligned tape stay aligned.
    return (TapePlanes(nx, ny, nz).stride + 3) / 4 * 4;
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

// ------------------------------------------------------------- energy without allocation --
// acfdtd_hip_energy_noalloc keeps the per-workgroup partials in this library-owned array
// instead of a caller workspace. Each device holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for the current device
// (the symbol's name is a host-side stand-in, not a device address).
namespace acfdtd {

static __device__ __align__(256) double scratch[3 * ACFDTD_HIP_ENERGY_SCRATCH_TILES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int scratch_address(double **out) {
    int device = -1;
    const int status = check_hip(hipGetDevice(&device), "hipGetDevice");
    if (status != ACFDTD_HIP_OK) return status;
    if (device < 0 || device >= kMaxDevices)
        return fail(ACFDTD_HIP_EDEVICE, "device index " + std::to_string(device) + " out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device] == nullptr) {
        void *address = nullptr;
        const int rc = check_hip(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)),
                                 "hipGetSymbolAddress(scratch)");
        if (rc != ACFDTD_HIP_OK) return rc;
        g_scratch[device] = static_cast<double *>(address);
    }
    *out = g_scratch[device];
    return ACFDTD_HIP_OK;
}

// The descriptor's own tiling when its partials fit the scratch (then the sums are those of
// acfdtd_hip_energy bit for bit), else 8-row blocks with the shortest marches that fit.
int scratch_tiling(acfdtd_hip_desc &d) {
    if (tiling_for(d).launched <= ACFDTD_HIP_ENERGY_SCRATCH_TILES) return ACFDTD_HIP_OK;
    d.tile_j = 8;
    const Tiling t = make_tiling(d.nx, d.ny, d.nz, d.tile_j, 1);
    const int64_t columns = (int64_t)t.nblk_k * t.nblk_j;
    const int64_t chunks = ACFDTD_HIP_ENERGY_SCRATCH_TILES / columns;
    if (chunks < 1)
        return fail(ACFDTD_HIP_EINVAL,
                    "grid too large for the energy scratch: ceil(ny / 8) * ceil(nz / 64) = " +
                        std::to_string(columns) + " workgroup columns exceed " +
                        std::to_string(ACFDTD_HIP_ENERGY_SCRATCH_TILES));
    d.chunk_x = (int)((d.nx + chunks - 1) / chunks);
    return tiling_for(d).launched <= ACFDTD_HIP_ENERGY_SCRATCH_TILES
               ? ACFDTD_HIP_OK
               : fail(ACFDTD_HIP_EINVAL, "energy tiling does not fit the scratch");
}

}  // namespace acfdtd

extern "C" int acfdtd_hip_energy_noalloc(const acfdtd_hip_desc *desc, double *out, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (out == nullptr) return fail(ACFDTD_HIP_EINVAL, "energy needs an output buffer");
    acfdtd_hip_desc d = *desc;
    int rc = scratch_tiling(d);
    if (rc != ACFDTD_HIP_OK) return rc;
    double *partials = nullptr;
    rc = scratch_address(&partials);
    if (rc != ACFDTD_HIP_OK) return rc;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (d.real_bytes == 8) return energy_terms<double>(d, partials, out, s);
    return energy_terms<float>(d, partials, out, s);
}

namespace acfdtd {

__global__ __launch_bounds__(256) void stream_copy(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace acfdtd

extern "C" int acfdtd_hip_stream_copy(const void *src, void *dst, size_t bytes, void *stream) {
    if (!src || !dst) return fail(ACFDTD_HIP_EINVAL, "stream copy pointer is NULL");
    if (bytes % 16 || reinterpret_cast<uintptr_t>(src) % 16 || reinterpret_cast<uintptr_t>(dst) % 16)
        return fail(ACFDTD_HIP_EINVAL, "stream copy needs 16-byte sizes and alignment");
    const size_t n = bytes / 16;
    if (n == 0) return ACFDTD_HIP_OK;
    if ((n + 255) / 256 > 0x7fffffffu) return fail(ACFDTD_HIP_EINVAL, "stream copy too large");
    g_last_error.clear();
    hipLaunchKernelGGL(stream_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const float4 *>(src),
                       static_cast<float4 *>(dst), n);
    return check_hip(hipGetLastError(), "stream copy launch");
}

extern "C" int acfdtd_hip_fused2_step(const acfdtd_hip_desc *desc, int64_t step,
                                      int64_t recording_column, int reverse, void *stream) {
    const int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (step < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "step and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + 1 >= desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording columns out of range");
    // The sweep writes tape rows for steps `step` and `step + 1`.
    if (desc->wall_tape && (step < desc->wall_tape_step0 ||
                            step + 1 - desc->wall_tape_step0 >= desc->wall_tape_rows))
        return fail(ACFDTD_HIP_EINVAL, "wall tape does not cover these two steps");
    if (!div_tape_covers(*desc, step, step + 1))
        return fail(ACFDTD_HIP_EINVAL, "divergence tape does not cover these two steps");
    if (desc->layer_tape) return fail(ACFDTD_HIP_EINVAL, "the layer tape is written by acfdtd_hip_run only");
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8)
        return launch_fused2_once<double>(*desc, step, recording_column, reverse != 0, s);
    return launch_fused2_once<float>(*desc, step, recording_column, reverse != 0, s);
}
