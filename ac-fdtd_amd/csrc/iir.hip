// Frequency-dependent (IIR) walls on gfx950: the wall update that follows each step.
//
// The step itself is the first-order scheme's fused step (fused.hip: one sweep from one set of
// fields to the other, in pairs A -> B -> A, when the descriptor has the swap set; an odd last
// step, or a descriptor without it, takes the two in-place sweeps of sweeps.hip) with the
// walls' instantaneous admittance in the wall table and the boundary tape written; this kernel
// then runs each face cell's branch filters on the step's mean pressure and writes the branches'
// memory as the wall face's velocity for the next step, into the set that holds the step's
// result (include/acfdtd_hip.h, acfdtd_hip_iir_desc). The fused step copies every wall face
// from one set to the other, so faces without a filter keep their values either way. One lane per face cell in the tape's face-plane order, so the tape
// reads, the state reads and writes (2 per branch, contiguous per cell) and the wall-face
// stores of x and y faces are coalesced; 2 (ny nz + nx nz + nx ny) cells per step, a few
// hundred kilobytes against the sweeps' hundreds of megabytes. The CPU oracle restates this
// arithmetic, which the kernel reproduces bit for bit (built with -ffp-contract=off).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "acfdtd_hip.h"
#include "sweeps.hpp"

namespace acfdtd {

namespace {

template <typename real>
struct IIRArgs {
    int nx, ny, nz;
    int n_branches;
    unsigned face_mask;
    uint64_t wall_mask;
    real velocity_scale;
    real coef[6][ACFDTD_HIP_MAX_BRANCHES][4];
    const real *wall_table;
    const real *tape;  // row 0
    real *state;
    real *vx, *vy, *vz;
    // Face-plane offsets of a tape row (TapePlanes).
    int64_t x_hi, y_lo, y_hi, z_lo, z_hi, stride;
};

__device__ __forceinline__ int class_of(int i, int n) { return (i == 0 ? 1 : 0) | (i == n - 1 ? 2 : 0); }

template <typename real>
__global__ __launch_bounds__(256) void iir_wall_update(const IIRArgs<real> a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.stride) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    int face, i, j, k;
    if (e < a.y_lo) {
        face = e < a.x_hi ? 0 : 1;
        const int64_t q = e - (face == 0 ? 0 : a.x_hi);
        i = face == 0 ? 0 : nx - 1;
        j = (int)(q / nz);
        k = (int)(q % nz);
    } else if (e < a.z_lo) {
        face = e < a.y_hi ? 2 : 3;
        const int64_t q = e - (face == 2 ? a.y_lo : a.y_hi);
        j = face == 2 ? 0 : ny - 1;
        i = (int)(q / nz);
        k = (int)(q % nz);
    } else {
        face = e < a.z_hi ? 4 : 5;
        const int64_t q = e - (face == 4 ? a.z_lo : a.z_hi);
        k = face == 4 ? 0 : nz - 1;
        i = (int)(q / ny);
        j = (int)(q % ny);
    }
    if (!((a.face_mask >> face) & 1u)) return;
    const int code = class_of(i, nx) | (class_of(j, ny) << 2) | (class_of(k, nz) << 4);
    const real fu = ((a.wall_mask >> code) & 1ull) ? a.wall_table[code] : real(1);
    const real x = (fu * a.tape[e]) * real(0.5);
    real *state = a.state + 2 * a.n_branches * e;
    real memory = real(0);
    for (int b = 0; b < a.n_branches; ++b) {
        const real *c = a.coef[face][b];
        const real y = c[0] * x + state[2 * b];
        const real first = state[2 * b + 1] - c[2] * y;
        const real second = c[1] * x - c[3] * y;
        state[2 * b] = first;
        state[2 * b + 1] = second;
        memory = memory + first;
    }
    real velocity = memory * a.velocity_scale;
    if ((face & 1) == 0) velocity = -velocity;
    switch (face) {
        case 0: a.vx[(int64_t)j * nz + k] = velocity; break;
        case 1: a.vx[((int64_t)nx * ny + j) * nz + k] = velocity; break;
        case 2: a.vy[((int64_t)i * (ny + 1)) * nz + k] = velocity; break;
        case 3: a.vy[((int64_t)i * (ny + 1) + ny) * nz + k] = velocity; break;
        case 4: a.vz[((int64_t)i * ny + j) * (nz + 1)] = velocity; break;
        default: a.vz[((int64_t)i * ny + j) * (nz + 1) + nz] = velocity; break;
    }
}

template <typename real>
int iir_run(const acfdtd_hip_desc &desc, const acfdtd_hip_iir_desc &iir, int64_t step_offset,
            int64_t n_steps, int64_t column, hipStream_t s) {
    IIRArgs<real> a{};
    a.nx = desc.nx;
    a.ny = desc.ny;
    a.nz = desc.nz;
    a.n_branches = iir.n_branches;
    a.face_mask = iir.face_mask;
    a.wall_mask = desc.wall_mask;
    a.velocity_scale = (real)iir.velocity_scale;
    for (int f = 0; f < 6; ++f)
        for (int b = 0; b < ACFDTD_HIP_MAX_BRANCHES; ++b)
            for (int q = 0; q < 4; ++q) a.coef[f][b][q] = (real)iir.coef[f][b][q];
    a.wall_table = static_cast<const real *>(desc.wall_table);
    a.tape = static_cast<const real *>(desc.wall_tape);
    a.state = static_cast<real *>(iir.state);
    a.vx = static_cast<real *>(desc.vx);
    a.vy = static_cast<real *>(desc.vy);
    a.vz = static_cast<real *>(desc.vz);
    const TapePlanes planes(desc.nx, desc.ny, desc.nz);
    a.x_hi = planes.x_hi;
    a.y_lo = planes.y_lo;
    a.y_hi = planes.y_hi;
    a.z_lo = planes.z_lo;
    a.z_hi = planes.z_hi;
    a.stride = planes.stride;
    const unsigned blocks = (unsigned)((planes.stride + 255) / 256);
    const bool filters = a.face_mask != 0 && a.n_branches > 0;
    acfdtd_hip_desc d = desc;
    d.wall_tape_rows = 1;
    int64_t n = 0;
    if (desc.p_swap && desc.vx_swap && desc.vy_swap && desc.vz_swap) {
        real *const first[3] = {static_cast<real *>(desc.vx), static_cast<real *>(desc.vy),
                                static_cast<real *>(desc.vz)};
        real *const second[3] = {static_cast<real *>(desc.vx_swap),
                                 static_cast<real *>(desc.vy_swap),
                                 static_cast<real *>(desc.vz_swap)};
        for (; n + 1 < n_steps; n += 2) {
            for (int h = 0; h < 2; ++h) {
                d.wall_tape_step0 = step_offset + n + h;
                const int rc = acfdtd_hip_fused_step(&d, step_offset + n + h, column + n + h, h, s);
                if (rc != ACFDTD_HIP_OK) return rc;
                real *const *out = h == 0 ? second : first;  // the step's result
                a.vx = out[0];
                a.vy = out[1];
                a.vz = out[2];
                if (filters) hipLaunchKernelGGL(iir_wall_update<real>, dim3(blocks), dim3(256), 0, s, a);
            }
        }
        a.vx = first[0];
        a.vy = first[1];
        a.vz = first[2];
    }
    for (; n < n_steps; ++n) {
        d.wall_tape_step0 = step_offset + n;
        int rc = acfdtd_hip_velocity_sweep(&d, s);
        if (rc == ACFDTD_HIP_OK) rc = acfdtd_hip_pressure_sweep(&d, step_offset + n, column + n, s);
        if (rc != ACFDTD_HIP_OK) return rc;
        if (filters) hipLaunchKernelGGL(iir_wall_update<real>, dim3(blocks), dim3(256), 0, s, a);
    }
    return check_hip(hipGetLastError(), "wall update launch");
}

}  // namespace

}  // namespace acfdtd

using namespace acfdtd;

extern "C" int acfdtd_hip_iir_run(const acfdtd_hip_desc *desc, const acfdtd_hip_iir_desc *iir,
                                  int64_t step_offset, int64_t n_steps, int64_t recording_column,
                                  void *stream) {
    int status = validate(desc);
    if (status != ACFDTD_HIP_OK) return status;
    if (desc->layer_tape) return fail(ACFDTD_HIP_EINVAL, "IIR walls write no layer tape");
    if (iir == nullptr) return fail(ACFDTD_HIP_EINVAL, "IIR wall descriptor is NULL");
    if (iir->abi_version != ACFDTD_HIP_ABI_VERSION || iir->struct_size != sizeof(acfdtd_hip_iir_desc))
        return fail(ACFDTD_HIP_EABI, "IIR wall descriptor version or size mismatch");
    if (iir->n_branches < 0 || iir->n_branches > ACFDTD_HIP_MAX_BRANCHES)
        return fail(ACFDTD_HIP_EINVAL, "n_branches out of range");
    if (iir->face_mask > 63u) return fail(ACFDTD_HIP_EINVAL, "face_mask has bits beyond 6 faces");
    if (iir->face_mask != 0 && iir->n_branches > 0 && iir->state == nullptr)
        return fail(ACFDTD_HIP_EINVAL, "IIR wall states are NULL");
    if (desc->has_layer) return fail(ACFDTD_HIP_EINVAL, "IIR walls run without a sponge");
    if (desc->wall_tape == nullptr || desc->wall_tape_rows < 1)
        return fail(ACFDTD_HIP_EINVAL, "IIR walls need a wall tape of at least one row");
    if (n_steps < 0 || step_offset < 0 || recording_column < 0)
        return fail(ACFDTD_HIP_EINVAL, "n_steps, step_offset and recording_column must be >= 0");
    if (desc->n_receivers > 0 && recording_column + n_steps > desc->recording_stride)
        return fail(ACFDTD_HIP_EINVAL, "recording is too short for this many steps");
    if (n_steps == 0) return ACFDTD_HIP_OK;
    g_last_error.clear();
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->real_bytes == 8)
        return iir_run<double>(*desc, *iir, step_offset, n_steps, recording_column, s);
    return iir_run<float>(*desc, *iir, step_offset, n_steps, recording_column, s);
}
