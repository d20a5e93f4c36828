/* C-ABI of the MI355X (gfx950) time loop for ac-fdtd's AcousticFDTD.
 *
 * The reference compiles its fused two-sweep loop into a host library with five exports
 * (src/ac_fdtd/c/acfdtd.c:99-376) and drives it from Python through ctypes
 * (src/ac_fdtd/c_backend.py:83-96, :302-356). This library is the device counterpart: the
 * same struct-and-run shape, plain C types only, device pointers instead of host pointers,
 * and an explicit HIP stream. Callers pass memory they own (torch-ROCm tensors, hipMalloc),
 * including the energy workspace; these entry points allocate nothing and keep no state.
 *
 * Field layout (shared with the reference): C order (nx, ny, nz), z unit-stride.
 *   p   : nx*ny*nz            vx : (nx+1)*ny*nz
 *   vy  : nx*(ny+1)*nz        vz : nx*ny*(nz+1)
 *   psi : n_relax fields of nx*ny*nz, back to back.
 * State invariant between calls: (p^n, v^{n-1/2}, psi^n) with n the absolute step index.
 *
 * Errors: every entry point returns 0 on success and a negative ACFDTD_HIP_E* code otherwise;
 * acfdtd_hip_last_error() returns the message of the last failure on the calling thread.
 * Calls that share device buffers must not overlap (the reference is not re-entrant either,
 * acfdtd.c:196-206).
 */
#ifndef ACFDTD_HIP_H
#define ACFDTD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACFDTD_HIP_ABI_VERSION 1
#define ACFDTD_HIP_MAX_RELAXATION 4 /* acfdtd.c:52 */
#define ACFDTD_HIP_WALL_CODES 64    /* 4 classes per axis: interior, low, high, both */

/* Error codes. */
#define ACFDTD_HIP_OK 0
#define ACFDTD_HIP_EINVAL (-1)   /* malformed descriptor or argument */
#define ACFDTD_HIP_EABI (-2)     /* ABI version or struct size mismatch */
#define ACFDTD_HIP_EDEVICE (-3)  /* a HIP runtime call failed */

/* Flags. */
#define ACFDTD_HIP_COMPENSATED 1u /* fp32 only: error-free summation of the pressure increment */

/* One room's time loop. Scalar coefficients are doubles holding values already rounded to the
 * working precision (real_bytes), so the kernels reproduce the host's bits exactly. Replaces
 * the reference struct `acfdtd` (acfdtd.c:56-97). */
typedef struct acfdtd_hip_desc {
    uint32_t abi_version; /* ACFDTD_HIP_ABI_VERSION */
    uint32_t struct_size; /* sizeof(acfdtd_hip_desc) */
    int32_t nx, ny, nz;
    int32_t real_bytes; /* 8: fp64 kernels, 4: fp32 kernels */
    int32_t n_relax;    /* 0..ACFDTD_HIP_MAX_RELAXATION */
    int32_t has_layer;  /* sponge profiles present */
    int32_t n_sources;
    int32_t n_receivers;
    uint32_t flags;     /* ACFDTD_HIP_COMPENSATED */
    int32_t tile_j;     /* rows per workgroup (1, 2, 4 or 8); 0 = default */
    int32_t chunk_x;    /* planes marched per workgroup; 0 = default */
    int32_t reserved;

    double vel_coef;    /* dt / (rho dx) */
    double pres_coef;   /* rho c^2 dt / dx */
    double src_coef;    /* rho c^2 dt */
    double relax_scale; /* dt * sum of gains */
    double relax_decay[ACFDTD_HIP_MAX_RELAXATION];
    double relax_mean[ACFDTD_HIP_MAX_RELAXATION];
    double relax_gain[ACFDTD_HIP_MAX_RELAXATION];

    /* Walls: bit c of wall_mask says whether wall code c = cx + 4 cy + 16 cz (per axis:
     * 1 = low face, 2 = high face, 3 = both) touches an absorbing face; wall_table (device,
     * working precision) holds 1/(1+a) at [c] and a/(1+a) at [64 + c]. Replaces the sparse
     * wall_index / wall_from_* arrays (acfdtd.c:87-90). */
    uint64_t wall_mask;
    const void *wall_table;

    /* Fields (device, working precision). */
    void *p, *vx, *vy, *vz, *psi;

    /* Sponge profiles (device): cell[a] has n_a entries, face[a] has n_a + 1 (acfdtd.c:80-85).
     * Ignored when has_layer == 0. */
    const void *cell[3];
    const void *face[3];

    /* Sources (device): source_cell[2 s] = i, [2 s + 1] = j * nz + k; source_length[s];
     * source_signal rows of source_stride samples, working precision (acfdtd.c:92-93). */
    const int32_t *source_cell;
    const int64_t *source_length;
    const void *source_signal;
    int64_t source_stride;

    /* Receivers (device): receiver_cell as for sources; sample of step n (local to the call)
     * lands at recording[r * recording_stride + recording_column + n] (acfdtd.c:95-96). */
    const int32_t *receiver_cell;
    void *recording;
    int64_t recording_stride;
} acfdtd_hip_desc;

/* sizeof(acfdtd_hip_desc), for the caller's mirror check. Replaces acfdtd_struct_size
 * (acfdtd.c:99). */
size_t acfdtd_hip_struct_size(void);

/* ACFDTD_HIP_ABI_VERSION of the built library. */
int acfdtd_hip_abi_version(void);

/* Bytes of the working precision: 8 or 4 are both supported by one library. Replaces
 * acfdtd_is_double (acfdtd.c:101), which reported the single compiled precision. */
int acfdtd_hip_supports_real(int real_bytes);

/* Advance n_steps steps from absolute step `step_offset` on `stream` (a hipStream_t, NULL for
 * the default stream); receivers record into columns recording_column .. + n_steps - 1.
 * Asynchronous: returns after enqueueing. Replaces acfdtd_run (acfdtd.c:355-376). */
int acfdtd_hip_run(const acfdtd_hip_desc *desc, int64_t step_offset, int64_t n_steps,
                   int64_t recording_column, void *stream);

/* One velocity sweep, or one pressure sweep for absolute step `step` recording into
 * `recording_column`: the two kernels acfdtd_hip_run issues per step, exposed for per-kernel
 * timing and tile tuning. No reference counterpart (the reference's sweeps are static,
 * acfdtd.c:110-178). */
int acfdtd_hip_velocity_sweep(const acfdtd_hip_desc *desc, void *stream);
int acfdtd_hip_pressure_sweep(const acfdtd_hip_desc *desc, int64_t step, int64_t recording_column,
                              void *stream);

/* Energy terms of the stored state, in fp64 on the device: out[0] = sum p^2,
 * out[1] = sum over all faces of v^2, out[2] = sum over interior faces of v_a (p_hi - p_lo).
 * `out` is a device pointer to 3 doubles; `workspace` a device buffer of at least
 * acfdtd_hip_energy_workspace_bytes(desc) bytes. Replaces the host reductions of
 * scheme.py:365-389 / c_backend.py:361-382. */
size_t acfdtd_hip_energy_workspace_bytes(const acfdtd_hip_desc *desc);
int acfdtd_hip_energy(const acfdtd_hip_desc *desc, void *workspace, double *out, void *stream);

/* Message of the last failed call on this thread ("" if none). */
const char *acfdtd_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ACFDTD_HIP_H */
