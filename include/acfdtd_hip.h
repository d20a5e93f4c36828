/* C-ABI of the MI355X (gfx950) time loop for ac-fdtd's AcousticFDTD.
 *
 * The reference compiles its fused two-sweep loop into a host library with five exports
 * (src/ac_fdtd/c/acfdtd.c:99-376) and drives it from Python through ctypes
 * (src/ac_fdtd/c_backend.py:83-96, :302-356). This library is the device counterpart: the
 * same struct-and-run shape, plain C types only, device pointers instead of host pointers,
 * and an explicit HIP stream. Callers pass memory they own (torch-ROCm tensors, hipMalloc),
 * including every workspace; no entry point allocates device memory, and only the batch API
 * keeps state (a small host handle).
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
/* time_block = 2: requests each plane's first-step inputs one plane ahead, straight into LDS
 * (direct-to-LDS loads), while the previous plane's second step runs. Same results; used when
 * p, vx, vy, psi and their swap buffers are 16-byte aligned. */
#define ACFDTD_HIP_STAGED 2u

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
    uint32_t flags;     /* ACFDTD_HIP_COMPENSATED, ACFDTD_HIP_STAGED */
    int32_t tile_j;     /* wavefronts of lanes per workgroup, 1 to 8; 0 = default */
    int32_t chunk_x;    /* planes marched per workgroup; 0 = default */
    int32_t time_block; /* steps per sweep of acfdtd_hip_run: 0 or 1 = one, 2 = two (see psi_swap) */

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

    /* Optional second set of p / vx / vy / vz buffers (device, same shapes). When all four are
     * given, acfdtd_hip_run advances steps in pairs with one fused sweep per step, reading one
     * set and writing the other (A -> B -> A), which moves (8 + 2 n_relax) elements per
     * cell-update instead of (12 + 2 n_relax); a trailing odd step uses the two in-place sweeps.
     * Either way the state ends in p / vx / vy / vz and the results are bit-identical. The
     * swap contents are scratch on entry and exit. */
    void *p_swap, *vx_swap, *vy_swap, *vz_swap;

    /* Optional boundary tape for wall-admittance gradients (device, working precision). For
     * absolute step n, row n - wall_tape_step0 (of wall_tape_rows rows, wall_tape_stride
     * elements each) receives p_tmp + p_old of every boundary cell, where p_tmp is the
     * pressure before the wall closed form, once per face plane the cell lies on. Planes in
     * row order: x-low, x-high (ny*nz, index j*nz+k), y-low, y-high (nx*nz, i*nz+k), z-low,
     * z-high (nx*ny, i*ny+j), 2 (ny*nz + nx*nz + nx*ny) elements in all. Any stride of at
     * least that is accepted; acfdtd_hip_wall_tape_stride() gives it rounded up to a multiple
     * of 4, which with a 16-byte aligned tape lets the two-step sweep write x and y face rows
     * as vector stores. */
    void *wall_tape;
    int64_t wall_tape_stride;
    int64_t wall_tape_step0;
    int64_t wall_tape_rows;

    /* Second set of relaxation states (device, same shape as psi), for time_block = 2: a sweep
     * that advances two steps reads p, v, psi from one set and writes them to the other (its
     * halo cells read their neighbours' step-n states), in pairs A -> B -> A, so four steps
     * move (8 + 2 n_relax) elements per cell twice instead of four times. Used when time_block
     * is 2, the four swap buffers are given, psi_swap is given (or n_relax is 0) and nz is a
     * multiple of the lane width (4 cells fp32, 2 fp64); the remaining steps run one per
     * sweep. Results (boundary tape included) are bit-identical either way. */
    void *psi_swap;

    /* Padded rows, accepted by acfdtd_hip_run and the batch API only (the other entry points
     * return ACFDTD_HIP_EINVAL unless it is 0). 0: the layout above. Otherwise z_pitch > nz is the row
     * length, in elements, of every field and swap buffer, vz included: p and each psi field
     * nx*ny*z_pitch, vx (nx+1)*ny*z_pitch, vy nx*(ny+1)*z_pitch, vz nx*ny*z_pitch, cell (i, j, k)
     * at (i*ny + j)*z_pitch + k, vz face k of row (i, j) at (i*ny + j)*z_pitch + k; source and
     * receiver cells are given as j * z_pitch + k. The elements past the room's own (k >= nz,
     * k >= nz + 1 for vz) belong to no cell: the sweeps read and write them but never let them
     * reach a cell, so their contents are arbitrary. It lets a room whose nz is not a multiple
     * of the lane width (4 cells fp32, 2 fp64) take full-width lanes: z_pitch a multiple of
     * that width (with another pitch acfdtd_hip_run takes the two in-place sweeps). No boundary
     * or divergence tape, no staged sweep with padded rows. */
    int32_t z_pitch;
    int32_t reserved; /* 0 */

    /* Optional divergence tape, for the air-strength gradients (device, working precision).
     * For absolute step n, row n - div_tape_step0 (of div_tape_rows rows of nx*ny*nz elements,
     * C order like p) receives the velocity divergence every cell's pressure update of step n
     * uses, ((vx_hi - vx_lo) + (vy_hi - vy_lo)) + (vz_hi - vz_lo) of v^{n+1/2}, bit for bit.
     * The adjoint reads it in place of recomputed forward velocities (acfdtd_hip_adjoint_run),
     * so the air gradients need no recomputation and run two adjoint steps per sweep. Needs
     * n_relax > 0; not with padded rows. With a tape the two-step sweep runs on the widest
     * lanes only (nz a multiple of 4 fp32, 2 fp64); other grids take one step per sweep. */
    void *div_tape;
    int64_t div_tape_step0;
    int64_t div_tape_rows;

    /* Optional layer tape, for the sponge-strength gradient (device, working precision, 16-byte
     * aligned). For absolute step n, row n - layer_tape_step0 (of layer_tape_rows rows,
     * layer_tape_stride elements each, at least acfdtd_hip_layer_tape_stride(desc) and a
     * multiple of 4) receives the fields after step n of the cells of the absorbing layer's
     * frame: p (after the step's source injection), and the velocities of each cell's lower x,
     * y and z faces. The frame is given by layer_lo / layer_hi, the number of layer cells at
     * the low and high end of each axis (0 where that face has no layer): planes i < lo[0] or
     * i >= nx - hi[0] are whole; in the other planes rows j < lo[1] or j >= ny - hi[1] are
     * whole, and the other rows hold the cells k < zl or k >= nz - zh, zl and zh being lo[2]
     * and hi[2] rounded up to multiples of 4 (a side whose counts cover the axis makes
     * everything whole). Row layout: p of the frame cells (C of them, plane by plane, row by row,
     * k ascending), then vx, vy, vz of the same cells' lower faces (C each), then the far faces:
     * vx of face nx (ny*nz, j*nz + k), vy of face ny (nx*nz, i*nz + k), vz of face nz (nx*ny,
     * i*ny + j). Written by acfdtd_hip_run (the other stepping entry points refuse a
     * descriptor with one); read by the two-step adjoint for the sponge gradient
     * (acfdtd_hip_adjoint_run) and unpacked by acfdtd_hip_layer_tape_unpack. Needs has_layer;
     * not with padded rows. No reference counterpart: the reference has no gradients. */
    void *layer_tape;
    int64_t layer_tape_step0;
    int64_t layer_tape_rows;
    int64_t layer_tape_stride;
    int32_t layer_lo[3];
    int32_t layer_hi[3];
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

/* One fused step (velocity and pressure in one sweep) for absolute step `step`, reading
 * p / vx / vy / vz and writing p_swap / vx_swap / vy_swap / vz_swap (psi in place); with
 * `reverse` nonzero the two sets swap roles. For per-kernel timing and tuning; no reference
 * counterpart. */
int acfdtd_hip_fused_step(const acfdtd_hip_desc *desc, int64_t step, int64_t recording_column,
                          int reverse, void *stream);

/* One two-step sweep (steps `step` and `step` + 1, recording into `recording_column` and the
 * next), from p / vx / vy / vz / psi to the swap set, or back with `reverse` nonzero; needs
 * what time_block = 2 needs (see psi_swap) and fails with ACFDTD_HIP_EINVAL otherwise. For
 * per-kernel timing and tuning; no reference counterpart. */
int acfdtd_hip_fused2_step(const acfdtd_hip_desc *desc, int64_t step, int64_t recording_column,
                           int reverse, void *stream);

/* ----------------------------------------------------------------------- batched rooms --
 * Many rooms advance together, one launch per step (one fused launch per step with swap
 * buffers, else two sweeps): a tile table covering every room is built once in a caller-owned
 * device workspace, with each room's arguments beside it. Rooms must share real_bytes,
 * n_relax, has_layer, flags, tile_j and whether swap buffers are given; sizes, walls, air,
 * sources and receivers may differ. The lanes are as wide as every room's row length allows
 * (4 cells fp32, 2 fp64 at most): a room whose nz is not a multiple of that width can be given
 * in padded rows (z_pitch), so that it does not narrow the whole batch's lanes. No reference
 * counterpart (the reference runs one room). */
typedef struct acfdtd_hip_batch acfdtd_hip_batch; /* opaque host handle */

size_t acfdtd_hip_batch_workspace_bytes(const acfdtd_hip_desc *rooms, int32_t n_rooms);
int acfdtd_hip_batch_create(const acfdtd_hip_desc *rooms, int32_t n_rooms, void *workspace,
                            size_t workspace_bytes, void *stream, acfdtd_hip_batch **out);
/* Re-upload the rooms' descriptors (new recording buffers, edited coefficients); the rooms'
 * sizes and tiling must be those the batch was created with. Synchronises `stream`. */
int acfdtd_hip_batch_update(acfdtd_hip_batch *batch, const acfdtd_hip_desc *rooms,
                            int32_t n_rooms, void *stream);
int acfdtd_hip_batch_run(acfdtd_hip_batch *batch, int64_t step_offset, int64_t n_steps,
                         int64_t recording_column, void *stream);
/* 1 when the batch advances its rooms two steps per sweep (every room's descriptor allows it:
 * time_block = 2 with the swap set and psi_swap, at most 64 sources and 256 receivers), the
 * leading multiple of four steps of each run; else 0. No reference counterpart. */
int acfdtd_hip_batch_two_step(const acfdtd_hip_batch *batch);
void acfdtd_hip_batch_destroy(acfdtd_hip_batch *batch);

/* Stride of a boundary tape row: 2 (ny nz + nx nz + nx ny) elements rounded up to a multiple
 * of 4. */
int64_t acfdtd_hip_wall_tape_stride(int32_t nx, int32_t ny, int32_t nz);

/* Elements per row of desc's layer tape (see acfdtd_hip_desc::layer_tape): 4 C + ny nz + nx nz
 * + nx ny rounded up to a multiple of 4, C the frame's cell count from nx, ny, nz, layer_lo
 * and layer_hi; -1 for a malformed descriptor. No reference counterpart. */
int64_t acfdtd_hip_layer_tape_stride(const acfdtd_hip_desc *desc);

/* The fields of layer tape row `step` - layer_tape_step0 as full state fields (device, the
 * shapes of p, vx, vy, vz): the frame cells' p and lower faces and the far faces from the
 * tape, every other element 0. The sponge gradient reads them only where the grading weights
 * are nonzero, so such a state stands in for the forward state of that step
 * (acfdtd_hip_adjoint_desc::fwd_p ...). No reference counterpart. */
int acfdtd_hip_layer_tape_unpack(const acfdtd_hip_desc *desc, int64_t step, void *p, void *vx,
                                 void *vy, void *vz, void *stream);

/* Energy terms of the stored state, in fp64 on the device: out[0] = sum p^2,
 * out[1] = sum over all faces of v^2, out[2] = sum over interior faces of v_a (p_hi - p_lo).
 * `out` is a device pointer to 3 doubles; `workspace` a device buffer of at least
 * acfdtd_hip_energy_workspace_bytes(desc) bytes. Replaces the host reductions of
 * scheme.py:365-389 / c_backend.py:361-382. */
size_t acfdtd_hip_energy_workspace_bytes(const acfdtd_hip_desc *desc);
int acfdtd_hip_energy(const acfdtd_hip_desc *desc, void *workspace, double *out, void *stream);

/* acfdtd_hip_energy without a workspace and without any allocation: the per-workgroup partials
 * live in a library-owned device array of 3 * ACFDTD_HIP_ENERGY_SCRATCH_TILES doubles (768 KiB,
 * held by each device the library runs on until the process ends; its address is looked up
 * once per device, on that device's first call). Same results as acfdtd_hip_energy bit for bit
 * when the descriptor's tiling has at most ACFDTD_HIP_ENERGY_SCRATCH_TILES workgroups; larger
 * tilings are replaced by 8-row blocks with longer marches (the fp64 sums then differ in
 * order only), and grids with more than ACFDTD_HIP_ENERGY_SCRATCH_TILES workgroup columns
 * (ceil(ny / 8) * ceil(nz / 64)) are refused with ACFDTD_HIP_EINVAL. Every workgroup writes its
 * own partials, so nothing needs clearing between calls; calls that overlap on one device
 * (other streams or threads) share the array and race. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation). */
#define ACFDTD_HIP_ENERGY_SCRATCH_TILES 32768
int acfdtd_hip_energy_noalloc(const acfdtd_hip_desc *desc, double *out, void *stream);

/* Stream ceiling for the roofline: copy `bytes` (a multiple of 16, both pointers 16-byte
 * aligned) from src to dst, one 16-byte vector per lane. Measures the HBM bandwidth a plain
 * streaming kernel reaches on this device; no reference counterpart (the reference times a
 * NumPy add, scripts/backend_benchmark.py:86-115). */
int acfdtd_hip_stream_copy(const void *src, void *dst, size_t bytes, void *stream);

/* ------------------------------------------------------------------------------ adjoint --
 * The transposed, time-reversed step and the parameter gradients it yields (no reference
 * counterpart: the reference is not differentiable). An adjoint step n is a face sweep
 * followed by a cell sweep that applies the stencil part of step n and the cell-local part of
 * step n - 1; the local part of the last step runs alone first. Gradients accumulate in fp64
 * into `partials` (acfdtd_hip_adjoint_partials_count doubles, zeroed by the caller), 12 slots
 * per wavefront: [0..5] dL/da summed over the cells of each face plane (FACES order),
 * [6] dL/dsigma_max, [7] dL/drelax_scale, [8..11] dL/dgain_r. */
#define ACFDTD_HIP_GRAD_WALLS 1u  /* needs the forward boundary tape (desc->wall_tape) */
#define ACFDTD_HIP_GRAD_SPONGE 2u /* needs the grading weights, and fwd_* or desc->layer_tape */
#define ACFDTD_HIP_GRAD_AIR 4u    /* needs fwd_vx / fwd_vy / fwd_vz or desc->div_tape */

typedef struct acfdtd_hip_adjoint_desc {
    uint32_t abi_version; /* ACFDTD_HIP_ABI_VERSION */
    uint32_t struct_size; /* sizeof(acfdtd_hip_adjoint_desc) */
    uint32_t grads;       /* ACFDTD_HIP_GRAD_* */
    /* acfdtd_hip_adjoint_run only: 2 advances two adjoint steps per sweep where allowed (see
     * lp_swap), 0 or 1 one. */
    int32_t time_block;
    /* Adjoint state, same shapes as p, vx, vy, vz, psi; ldiv is one cell field of scratch. */
    void *lp, *lvx, *lvy, *lvz, *lpsi, *ldiv;
    /* dL/d(recording): column (step - grad_recording_step0) of rows of
     * grad_recording_stride samples, one row per receiver. */
    const void *grad_recording;
    int64_t grad_recording_stride;
    int64_t grad_recording_step0;
    /* dL/d(source signal), rows of desc->source_stride samples (written, not accumulated);
     * NULL to skip. */
    void *grad_signal;
    /* Forward state after the step whose local part the call transposes (cell sweep) or
     * after the step it transposes (face sweep): p (before nothing; sources are removed in the
     * kernel), vx, vy, vz. Only read for sponge and air gradients. */
    const void *fwd_p, *fwd_vx, *fwd_vy, *fwd_vz;
    /* Sponge grading weights inside^order at cell centres and faces (the exponents of the
     * profiles in desc->cell / desc->face divided by -sigma_max dt). */
    const void *weight_cell[3];
    const void *weight_face[3];
    double dt;
    double *partials;
    /* acfdtd_hip_adjoint_run only. A second set of ldiv, lvx, lvy, lvz (same shapes) lets it
     * run adjoint steps as one fused sweep each, in pairs from the first set to the second and
     * back; with any of them NULL it runs the face and cell sweeps. */
    void *ldiv_swap, *lvx_swap, *lvy_swap, *lvz_swap;
    /* acfdtd_hip_adjoint_run only: fwd_p / fwd_vx / fwd_vy / fwd_vz are the bases of
     * fwd_count consecutive forward states (each field's states back to back): slot q holds
     * (p^n, v^(n-1/2)) with n = fwd_step0 + q. Needed for sponge and air gradients. */
    int64_t fwd_step0;
    int64_t fwd_count;
    /* acfdtd_hip_adjoint_run with time_block = 2: second lp and lpsi (same shapes). With them
     * and the swap set above, the leading multiple of four adjoint steps whose local parts all
     * exist (steps >= 1) run two per sweep when no forward states are given (fwd_count 0: wall
     * and signal gradients, air gradients from desc->div_tape, the sponge gradient from
     * desc->layer_tape), nz is a multiple of 4 (fp32) or 2 (fp64) and there are at most 64
     * sources and 256 receivers; the other steps one per sweep. With a layer tape and no
     * forward states every step of the call must run two per sweep: step_lo >= 1 and a
     * multiple of four steps, else ACFDTD_HIP_EINVAL. The results are the same bits (the fp64
     * partials summed in another order); the state ends in the first set. */
    void *lp_swap, *lpsi_swap;
} acfdtd_hip_adjoint_desc;

size_t acfdtd_hip_adjoint_partials_count(const acfdtd_hip_desc *desc);
int acfdtd_hip_adjoint_face_sweep(const acfdtd_hip_desc *desc,
                                  const acfdtd_hip_adjoint_desc *adjoint, void *stream);
int acfdtd_hip_adjoint_cell_sweep(const acfdtd_hip_desc *desc,
                                  const acfdtd_hip_adjoint_desc *adjoint, int64_t step,
                                  int stencil, int local, void *stream);

/* Adjoint steps step_hi, step_hi - 1, ..., step_lo in one call: each is the face sweep and
 * stencil of step m and the local part of step m - 1 (none for m = 0), as the two entry
 * points above would do it, with the same results bit for bit in lp, lvx, lvy, lvz, lpsi and
 * grad_signal (the fp64 gradient partials are summed in another order). The local part of
 * step step_hi + 1 must already have been applied (the cell sweep with stencil = 0 for the
 * last step of a run). The state ends in the first set (lp, ldiv, lvx, lvy, lvz). */
int acfdtd_hip_adjoint_run(const acfdtd_hip_desc *desc, const acfdtd_hip_adjoint_desc *adjoint,
                           int64_t step_hi, int64_t step_lo, void *stream);

/* 1 when acfdtd_hip_adjoint_run would advance these descriptors two adjoint steps per sweep
 * (see acfdtd_hip_adjoint_desc::lp_swap), else 0. No reference counterpart. */
int acfdtd_hip_adjoint_two_step(const acfdtd_hip_desc *desc,
                                const acfdtd_hip_adjoint_desc *adjoint);

/* Forward recomputation into a ring of states: slot 0 of p_states, vx_states, vy_states,
 * vz_states (each field's states back to back) holds (p, v) at step step_offset; advances
 * n_steps fused steps, writing the state after step step_offset + q into slot q + 1, and the
 * relaxation states in place in desc->psi. desc->p, vx, vy, vz and the swap buffers are not
 * used, nothing is recorded and no boundary tape is written. Used to recompute a checkpoint
 * segment for the adjoint. */
int acfdtd_hip_run_states(const acfdtd_hip_desc *desc, int64_t step_offset, int64_t n_steps,
                          void *p_states, void *vx_states, void *vy_states, void *vz_states,
                          void *stream);

/* ------------------------------------------------------ pressure-only (wave) scheme --
 * The second-order wave equation on pressure alone: the variant the reference plans and does
 * not implement ("two fields instead of four", PLAN.md:52-56, milestone M8 at PLAN.md:116).
 * It is the first-order step (scheme.py:285-363) with the velocity eliminated, exactly, for
 * rigid and locally reacting walls and soft sources (no sponge and no air absorption: those
 * attach to the velocity and the relaxation states). With lambda^2 = (c dt / dx)^2, a the
 * summed wall coefficient of the cell (0 inside) and L the 7-point Laplacian over the cell's
 * interior faces:
 *   p^{n+1} = ((2 p^n - (1 - a) p^{n-1}) + lambda^2 L p^n) / (1 + a)
 *   then, per source in insertion order, p^{n+1}[cell] += src s[n] - src s[n - 1].
 * One sweep per step reads p^n and p^{n-1} and writes p^{n+1} over p^{n-1}: 3 elements per
 * cell-update. No reference counterpart (the reference has no such scheme). */
typedef struct acfdtd_hip_wave_desc {
    uint32_t abi_version; /* ACFDTD_HIP_ABI_VERSION */
    uint32_t struct_size; /* sizeof(acfdtd_hip_wave_desc) */
    int32_t nx, ny, nz;
    int32_t real_bytes; /* 4 or 8 */
    int32_t n_sources;
    int32_t n_receivers;
    int32_t tile_j;  /* wavefronts per workgroup, 1 to 8; 0 = 4 */
    int32_t chunk_x; /* planes marched per workgroup; 0 = default */
    double courant2; /* lambda^2, already rounded to the working precision */
    double src_coef; /* rho c^2 dt, already rounded */
    /* Bit c of wall_mask: wall code c (as in acfdtd_hip_desc) touches an absorbing face;
     * wall_table (device, working precision) holds 1 / (1 + a) at [c] and 1 - a at [64 + c]. */
    uint64_t wall_mask;
    const void *wall_table;
    /* p^n and p^{n-1} of the call's first step (device, nx*ny*nz each, distinct). The two
     * swap roles every step: after an even n_steps p holds the newest field, after an odd one
     * p_prev does. */
    void *p, *p_prev;
    /* Sources and receivers, laid out as in acfdtd_hip_desc. */
    const int32_t *source_cell;
    const int64_t *source_length;
    const void *source_signal;
    int64_t source_stride;
    const int32_t *receiver_cell;
    void *recording;
    int64_t recording_stride;
    /* Two steps per sweep (time_block = 2, both swap buffers given, distinct, same shape):
     * a sweep reads (p^n, p^{n-1}) from one pair and writes (p^{n+2}, p^{n+1}) to the other,
     * 4 elements per two cell-updates instead of 6; sweeps run in pairs, p / p_prev ->
     * p_swap / p_prev_swap -> p / p_prev, for the leading multiple of four steps, the rest one
     * step per sweep. The swap pair is scratch; the roles of p and p_prev on return follow the
     * rule above, and the results are the same bits as one step per sweep. A workgroup owns
     * whole rows, or equal row segments when a row is too long, plus one halo row each side
     * (tile_j > 0: 2 * tile_j owned rows; 0: as many as fill 512 lanes). */
    int32_t time_block; /* 0 or 1: one step per sweep; 2: two */
    int32_t reserved;
    void *p_swap, *p_prev_swap;
} acfdtd_hip_wave_desc;

size_t acfdtd_hip_wave_struct_size(void);
/* Advance n_steps from absolute step step_offset on `stream`; receivers record p^{n+1} of
 * each step into columns recording_column onwards. */
int acfdtd_hip_wave_run(const acfdtd_hip_wave_desc *desc, int64_t step_offset, int64_t n_steps,
                        int64_t recording_column, void *stream);

/* ------------------------------------------------------ fourth order in space --
 * The staggered scheme with the fourth-order difference D u = c1 (u[i] - u[i-1]) -
 * c2 (u[i+1] - u[i-2]) (c1 = 9/8, c2 = 1/24 for fourth order; 1, 0 gives the second-order
 * sweeps' arithmetic) in the velocity and pressure sweeps: the reference's "4th-order-in-space
 * variant as an option" (PLAN.md:17, :116), which it does not implement. Rigid walls by images
 * (pressure even, normal velocity odd across each wall face); sponge, air absorption, sources
 * and receivers as in acfdtd_hip_run. The descriptor is acfdtd_hip_desc with wall_mask 0, no
 * flags and no tape (refused otherwise); swap buffers and time_block are ignored: each step is
 * one velocity sweep and one pressure sweep in place, (12 + 2 n_relax) elements per
 * cell-update. The caller picks dt within the fourth-order limit c dt / dx <= 1 / (sqrt(3)
 * (c1 + c2)). No reference counterpart. */
int acfdtd_hip_fourth_run(const acfdtd_hip_desc *desc, double c1, double c2, int64_t step_offset,
                          int64_t n_steps, int64_t recording_column, void *stream);

/* ------------------------------------------------- frequency-dependent (IIR) walls --
 * Walls whose normalised admittance is a conductance plus mass-spring-damper branches, each
 * branch a biquad from the bilinear transform: the "frequency-dependent impedance via an IIR
 * filter fitted to octave-band absorption data" the reference plans (PLAN.md:21-24) and does
 * not implement. The walls' instantaneous admittance sits in desc's wall table as for a real
 * admittance; after each step every face cell of a frequency-dependent face (bit f of
 * face_mask, faces in the order x-low, x-high, y-low, y-high, z-low, z-high) takes its step's
 * mean pressure x = (f_u T) * 0.5, with T the boundary tape value p_tmp + p_old and f_u the
 * cell's 1 / (1 + a), runs each branch b (TDF-II: y = b0 x + s1; s1 <- s2 - a1 y;
 * s2 <- b2 x - a2 y; coefficients coef[f][b] = {b0, b2, a1, a2}) and writes
 * (sum_b s1) * velocity_scale on its wall face (negated on low faces), which the next
 * pressure sweep reads as the wall's outward velocity. No reference counterpart. */
#define ACFDTD_HIP_MAX_BRANCHES 8
typedef struct acfdtd_hip_iir_desc {
    uint32_t abi_version; /* ACFDTD_HIP_ABI_VERSION */
    uint32_t struct_size; /* sizeof(acfdtd_hip_iir_desc) */
    int32_t n_branches;   /* 0 .. ACFDTD_HIP_MAX_BRANCHES, the same count on every face */
    uint32_t face_mask;
    double velocity_scale; /* 1 / (rho c), working precision */
    double coef[6][ACFDTD_HIP_MAX_BRANCHES][4]; /* working precision */
    /* Branch states (device, working precision, zero at the start): element e of a tape row
     * (the face-plane layout of acfdtd_hip_desc::wall_tape) owns 2 * n_branches values at
     * state[2 * n_branches * e]. */
    void *state;
} acfdtd_hip_iir_desc;

/* Advance n_steps: per step the first-order step writing the tape, then the wall update. With
 * the swap set in desc the steps go in fused pairs (one sweep from one set to the other, the
 * wall update writing into the set that holds the result, the state back in the first set
 * after each pair); an odd last step, or desc without the swap set, is the velocity sweep and
 * the pressure sweep in place. Same results either way. desc needs a wall tape of at least one
 * row (used as scratch: the call writes row 0 every step), no sponge, and the wall table of
 * the instantaneous admittances. */
int acfdtd_hip_iir_run(const acfdtd_hip_desc *desc, const acfdtd_hip_iir_desc *iir,
                       int64_t step_offset, int64_t n_steps, int64_t recording_column,
                       void *stream);

/* ------------------------------------------------------ split-field perfectly matched layer --
 * Berenger's split-field layer for the first-order scheme: the "true PML held in reserve"
 * the reference plans (PLAN.md:21-24) and contrasts with its sponge (boundaries.py:42-46),
 * not implemented there. Each field is damped only along its own axis, the pressure split in
 * p = p_x + p_y + p_z inside the layer; the damping is centred in time, with per-axis
 * coefficients A = (1 - sigma dt / 2) / (1 + sigma dt / 2) and B = 1 / (1 + sigma dt / 2) at
 * cell centres and on faces. Per step:
 *   interior faces: v_a = A_f v_a - B_f (vel (p[i] - p[i-1]));
 *   cells with A_c = 1 on all three axes: p = p - ((Dx + Dy) + Dz) pres, Da = v_a[hi] - v_a[lo];
 *   other cells: p_x = (p - p_y) - p_z; p_a = A_c,a p_a - B_c,a (pres Da) for each axis;
 *                p = (p_x + p_y) + p_z (p_y and p_z stored);
 *   then sources and receivers as acfdtd_hip_run. desc must have rigid walls (wall_mask 0),
 * no relaxation states, no sponge, no flags and no tape. With the swap set in desc the steps go
 * in fused pairs (one sweep per step from one set of fields to the other, A -> B -> A, the split
 * parts updated in place); an odd last step, or desc without the swap set, is one velocity and
 * one pressure sweep in place. Same results either way; time_block is ignored. No reference
 * counterpart. */
typedef struct acfdtd_hip_pml_desc {
    uint32_t abi_version; /* ACFDTD_HIP_ABI_VERSION */
    uint32_t struct_size; /* sizeof(acfdtd_hip_pml_desc) */
    /* Device, working precision: cell_* have n_a entries, face_* n_a + 1. */
    const void *cell_a[3], *cell_b[3];
    const void *face_a[3], *face_b[3];
    /* Split parts of the pressure in the layer (device, nx*ny*nz each, zero at the start). */
    void *p_y, *p_z;
} acfdtd_hip_pml_desc;

int acfdtd_hip_pml_run(const acfdtd_hip_desc *desc, const acfdtd_hip_pml_desc *pml,
                       int64_t step_offset, int64_t n_steps, int64_t recording_column,
                       void *stream);

/* Message of the last failed call on this thread ("" if none). */
const char *acfdtd_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ACFDTD_HIP_H */
