// Two steps per sweep (temporal blocking): the body of the fused2 kernels.
//
// One sweep reads (p^n, v^(n-1/2), psi^n) from the `in` set and writes
// (p^(n+2), v^(n+3/2), psi^(n+2)) to the `out` set, so each field crosses HBM once per two
// steps. A workgroup owns a block of rows (y) by lane groups (z, V cells per lane, as in
// fused_march_vec) and marches a chunk of x planes. Along x it runs the first step one plane
// ahead of the second: step 1 of plane i + 1, then step 2 of plane i, which needs p^(n+1) of
// planes i - 1, i, i + 1 (registers of the same lane) and of its four y and z neighbours in
// plane i (other lanes: exchanged through LDS). So the workgroup also runs step 1, without
// storing anything, on a halo row above and below the block, on a halo lane group at each end
// of a row segment shorter than the row, and on one halo plane at each end of its chunk.
//
// Owner lanes also write the boundary tape of both steps when one is attached.
//
// Every cell's arithmetic is that of fused_march_vec term by term, and a halo cell's first
// step reads the same unmodified inputs as its owner's, so p^(n+1), v^(n+1/2) and psi^(n+1)
// are the same bits in every lane that computes them, and the sweep is bit-identical to two
// single-step sweeps. The relaxation states are read by halo lanes too, so unlike the
// single-step sweep they ping-pong with the other fields.
//
// LDS holds p^(n+1) of three planes (written for plane i + 1 while plane i is read, and one
// more so that one barrier per plane suffices): 3 x 512 lanes x V values, 24 KiB.

#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "acfdtd_hip.h"
#include "marches.hpp"
#include "sweeps.hpp"

namespace acfdtd {

// Largest workgroup of the two-step sweep (launch bound and LDS size).
constexpr int kFused2Threads = 512;
// Source and receiver tables up to these sizes are copied to LDS (larger ones are read from
// global memory).
constexpr int kFused2Sources = 64;
constexpr int kFused2Receivers = 256;

template <typename real>
struct Fused2IO {
    const real *p_in, *vx_in, *vy_in, *vz_in, *psi_in;
    real *p_out, *vx_out, *vy_out, *vz_out, *psi_out;
};

// The workgroup barrier of the plane exchange. It orders LDS only: __syncthreads is also a
// workgroup release fence on global memory, which makes every wave wait for its outstanding
// global stores (the fields, the tape, the recording) at each plane, and no wave of a sweep
// reads what another writes to global memory.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// s_waitcnt vmcnt(0): wait for every outstanding vector memory operation (gfx9 encoding:
// vmcnt in bits 3:0 and 15:14, expcnt in 6:4, lgkmcnt in 11:8; the others at their maximum).
__device__ __forceinline__ void wait_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// s_waitcnt lgkmcnt(0): wait for the wave's outstanding LDS (and scalar memory) operations.
__device__ __forceinline__ void wait_lds() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// s_waitcnt vmcnt(N): wait until at most the N most recently issued vector memory operations
// of the wave are outstanding (the counter completes in issue order).
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt is six bits");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | 0x70 | 0xF00 | ((N >> 4) << 14));
}

// The stage's reads for one lane, in asm blocks that also wait for them. The compiler's own
// bookkeeping of direct-to-LDS loads would put a wait for every earlier memory operation (the
// previous plane's stores too) before these reads; the blocks keep it from seeing them, and
// the caller's wait_vm covers the stage. Each block's results are complete when it ends (its
// last instruction waits for them), and its outputs are early-clobber so that no result
// register overlaps an address still to be read. The blocks declare no memory clobber (that
// alone makes the compiler wait for all memory operations first); they stay between the
// caller's wait_vm and the next stage request, side-effecting builtins a volatile asm block is
// not moved across.
//
// Layout (bytes from the stage's start): 16-byte field f of lane t at f * kStageField + 16 t,
// f < NF; 4-byte word w of lane t at NF * kStageField + w * kStageWord + 4 t. Reads go through
// two base addresses (lane's slot of field 0, and of field 4) and immediate offsets (< 64 KiB).
constexpr int kStageField = 16 * kFused2Threads;
constexpr int kStageWord = 4 * kFused2Threads;
using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
template <int NF>
struct StageRead16;
template <int NF, int NS>
struct StageRead4;

template <>
struct StageRead16<6> {
    __device__ __forceinline__ static void run(const uint32_t base0, const uint32_t base4,
                                               u32x4 (&o)[6]) {
        asm volatile(
            "ds_read_b128 %0, %6 offset:0\n\t"
            "ds_read_b128 %1, %6 offset:8192\n\t"
            "ds_read_b128 %2, %6 offset:16384\n\t"
            "ds_read_b128 %3, %6 offset:24576\n\t"
            "ds_read_b128 %4, %7 offset:0\n\t"
            "ds_read_b128 %5, %7 offset:8192\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5])
                     : "v"(base0), "v"(base4));
    }
};

template <>
struct StageRead16<7> {
    __device__ __forceinline__ static void run(const uint32_t base0, const uint32_t base4,
                                               u32x4 (&o)[7]) {
        asm volatile(
            "ds_read_b128 %0, %7 offset:0\n\t"
            "ds_read_b128 %1, %7 offset:8192\n\t"
            "ds_read_b128 %2, %7 offset:16384\n\t"
            "ds_read_b128 %3, %7 offset:24576\n\t"
            "ds_read_b128 %4, %8 offset:0\n\t"
            "ds_read_b128 %5, %8 offset:8192\n\t"
            "ds_read_b128 %6, %8 offset:16384\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5]), "=&v"(o[6])
                     : "v"(base0), "v"(base4));
    }
};

template <>
struct StageRead16<8> {
    __device__ __forceinline__ static void run(const uint32_t base0, const uint32_t base4,
                                               u32x4 (&o)[8]) {
        asm volatile(
            "ds_read_b128 %0, %8 offset:0\n\t"
            "ds_read_b128 %1, %8 offset:8192\n\t"
            "ds_read_b128 %2, %8 offset:16384\n\t"
            "ds_read_b128 %3, %8 offset:24576\n\t"
            "ds_read_b128 %4, %9 offset:0\n\t"
            "ds_read_b128 %5, %9 offset:8192\n\t"
            "ds_read_b128 %6, %9 offset:16384\n\t"
            "ds_read_b128 %7, %9 offset:24576\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5]), "=&v"(o[6]), "=&v"(o[7])
                     : "v"(base0), "v"(base4));
    }
};

template <>
struct StageRead16<9> {
    __device__ __forceinline__ static void run(const uint32_t base0, const uint32_t base4,
                                               u32x4 (&o)[9]) {
        asm volatile(
            "ds_read_b128 %0, %9 offset:0\n\t"
            "ds_read_b128 %1, %9 offset:8192\n\t"
            "ds_read_b128 %2, %9 offset:16384\n\t"
            "ds_read_b128 %3, %9 offset:24576\n\t"
            "ds_read_b128 %4, %10 offset:0\n\t"
            "ds_read_b128 %5, %10 offset:8192\n\t"
            "ds_read_b128 %6, %10 offset:16384\n\t"
            "ds_read_b128 %7, %10 offset:24576\n\t"
            "ds_read_b128 %8, %10 offset:32768\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5]), "=&v"(o[6]), "=&v"(o[7]), "=&v"(o[8])
                     : "v"(base0), "v"(base4));
    }
};

template <>
struct StageRead16<10> {
    __device__ __forceinline__ static void run(const uint32_t base0, const uint32_t base4,
                                               u32x4 (&o)[10]) {
        asm volatile(
            "ds_read_b128 %0, %10 offset:0\n\t"
            "ds_read_b128 %1, %10 offset:8192\n\t"
            "ds_read_b128 %2, %10 offset:16384\n\t"
            "ds_read_b128 %3, %10 offset:24576\n\t"
            "ds_read_b128 %4, %11 offset:0\n\t"
            "ds_read_b128 %5, %11 offset:8192\n\t"
            "ds_read_b128 %6, %11 offset:16384\n\t"
            "ds_read_b128 %7, %11 offset:24576\n\t"
            "ds_read_b128 %8, %11 offset:32768\n\t"
            "ds_read_b128 %9, %11 offset:40960\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5]), "=&v"(o[6]), "=&v"(o[7]), "=&v"(o[8]), "=&v"(o[9])
                     : "v"(base0), "v"(base4));
    }
};

template <int NF>
struct StageRead4<NF, 7> {
    __device__ __forceinline__ static void run(const uint32_t base, uint32_t (&o)[7]) {
        asm volatile(
            "ds_read_b32 %0, %7 offset:0\n\t"
            "ds_read_b32 %1, %7 offset:2048\n\t"
            "ds_read_b32 %2, %7 offset:4096\n\t"
            "ds_read_b32 %3, %7 offset:6144\n\t"
            "ds_read_b32 %4, %7 offset:8192\n\t"
            "ds_read_b32 %5, %7 offset:10240\n\t"
            "ds_read_b32 %6, %7 offset:12288\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5]), "=&v"(o[6])
                     : "v"(base));
    }
};

template <int NF>
struct StageRead4<NF, 10> {
    __device__ __forceinline__ static void run(const uint32_t base, uint32_t (&o)[10]) {
        asm volatile(
            "ds_read_b32 %0, %10 offset:0\n\t"
            "ds_read_b32 %1, %10 offset:2048\n\t"
            "ds_read_b32 %2, %10 offset:4096\n\t"
            "ds_read_b32 %3, %10 offset:6144\n\t"
            "ds_read_b32 %4, %10 offset:8192\n\t"
            "ds_read_b32 %5, %10 offset:10240\n\t"
            "ds_read_b32 %6, %10 offset:12288\n\t"
            "ds_read_b32 %7, %10 offset:14336\n\t"
            "ds_read_b32 %8, %10 offset:16384\n\t"
            "ds_read_b32 %9, %10 offset:18432\n\t"
            "s_waitcnt lgkmcnt(0)"
                     : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]),
                       "=&v"(o[5]), "=&v"(o[6]), "=&v"(o[7]), "=&v"(o[8]), "=&v"(o[9])
                     : "v"(base));
    }
};

// The plane exchange's barrier in the staged sweep: this wave's LDS operations complete
// (lgkmcnt), then s_barrier. No memory fence: an LDS release fence would also wait for the
// stage's direct-to-LDS loads still in flight (they write LDS), which is the latency the
// staging hides; no wave reads another wave's stage slots. The signal fences keep the compiler
// from moving memory accesses across the barrier.
__device__ __forceinline__ void exchange_barrier() {
    wait_lds();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Direct-to-LDS load of 4 bytes per lane: lane l writes the word at `src` to wave_base + 4 l.
__device__ __forceinline__ void load_to_lds4(const void *src, void *wave_base) {
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)src,
                                     (void __attribute__((address_space(3))) *)wave_base, 4, 0, 0);
}

// The LDS byte address of a pointer into a __shared__ array.
__device__ __forceinline__ uint32_t lds_address(const void *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

// Direct-to-LDS load of 16 bytes per lane: lane l of the wavefront writes the 16 bytes at
// `src` (its own address, 16-byte aligned) to wave_base + 16 l. `wave_base` is the same for
// the whole wavefront. Nothing passes through registers; the load counts in the memory
// counter like any other.
__device__ __forceinline__ void load_to_lds16(const void *src, void *wave_base) {
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)src,
                                     (void __attribute__((address_space(3))) *)wave_base, 16, 0, 0);
}

// Workgroup shape of the two-step sweep: a block of `rows` rows (y) by `seg` lane groups of a
// row (z, `vec` cells each), plus one halo row above and below and, when a segment is shorter
// than the row, one halo lane group at each end. Thread t sits at block row t / pitch - 1 and
// segment position t % pitch - zhalo, so its y neighbours are t -/+ pitch and its z
// neighbours t -/+ 1.
struct Tiling2 {
    int vec;      // z cells per lane (V)
    int groups;   // lane groups per row (nz / V)
    int rows;     // own rows per block (R)
    int seg;      // own lane groups per block row (S)
    int zhalo;    // 1 when a segment is shorter than the row, else 0
    int pitch;    // S + 2 zhalo: threads per block row
    int threads;  // (R + 2) * pitch, rounded up to whole wavefronts (<= kFused2Threads)
    int chunk;    // x planes per march (fixed chunks)
    int nblk_g, nblk_j, nblk, nchunk, tiles, launched;
    int tape_vec;  // boundary tape rows start V-aligned: x and y face rows as vector stores
    // Balanced chunks (balance_tiling2) when >= 0: row blocks below `extra` split x into
    // nchunk + 1 chunks, the others into nchunk, each of nearly equal length; -1: fixed chunks
    // of `chunk` planes (the last one shorter).
    int extra;
};

// The x planes [i0, i1) that chunk xc of row block blk marches.
__host__ __device__ inline void chunk_planes(const Tiling2 &t, const int blk, const int xc,
                                             const int nx, int &i0, int &i1) {
    if (t.extra < 0) {
        i0 = xc * t.chunk;
        i1 = i0 + t.chunk < nx ? i0 + t.chunk : nx;
        return;
    }
    const int64_t n = t.nchunk + (blk < t.extra ? 1 : 0);
    i0 = (int)((int64_t)xc * nx / n);
    i1 = (int)((int64_t)(xc + 1) * nx / n);
}

// Whether acfdtd_hip_run may use the two-step sweep for this descriptor.
inline bool fused2_applies(const acfdtd_hip_desc &d) {
    return d.time_block == 2 && has_swap(d) && (d.n_relax == 0 || d.psi_swap != nullptr) &&
           d.n_sources <= kFused2Sources && d.n_receivers <= kFused2Receivers;
}

// Compute units of the current device, read once per device (256 when there is none).
inline int device_compute_units() {
    static int known[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 256;
    }
    if (known[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0) {
            (void)hipGetLastError();
            n = 256;
        }
        known[dev] = n;
    }
    return known[dev];
}

// Default planes per march of a launch of `nblk` (y, z) blocks of `waves` wavefronts each, whose
// marches also step through `halo` planes that they do not own. A compute unit works through
// ceil(workgroups / units) marches of (planes + halo) steps; up to 8 wavefronts per unit overlap
// their memory latency, so a unit with fewer costs the same as one with 8 and beyond that the
// cost grows with the wavefronts. The candidate with the lowest cost wins, on ties the longer
// march (fewer halo planes read). Short marches win on small grids, where too few workgroups
// would leave units idle and each march's per-plane latency sets the time; long ones on large
// grids, where every unit has several workgroups and the halo planes are extra traffic.
inline int default_march_planes(int nx, int64_t nblk, int waves, int halo) {
    static const int kPlanes[] = {1, 2, 3, 4, 6, 8, 12, 16, 20, 24, 28, 32};
    const int64_t units = device_compute_units();
    double best = 0;
    int pick = 1;
    for (int c : kPlanes) {
        const int64_t per_unit = ceil_div(nblk * ceil_div(nx, c), units);
        const double load = (double)per_unit * waves / 8.0;
        const double cost = (double)(c + halo) * (load > 1.0 ? load : 1.0);
        if (c == 1 || cost <= best) {
            best = cost;
            pick = c;
        }
        if (c >= nx) break;
    }
    return pick;
}

// Lane groups of a block of `rows` own rows: whole rows when rows + 2 of them fit in
// kFused2Threads lanes, else as few row segments as fit, of equal length (a short last segment
// would hold a workgroup's slot for a march of mostly idle lanes). False when none fits.
inline bool rows_shape(const int groups, const int rows, Tiling2 &t) {
    t.rows = rows;
    if ((rows + 2) * groups <= kFused2Threads) {
        t.seg = groups;
        t.zhalo = 0;
        t.pitch = groups;
    } else {
        const int widest = kFused2Threads / (rows + 2) - 2;
        if (widest < 1) return false;
        t.seg = ceil_div(groups, ceil_div(groups, widest));
        t.pitch = t.seg + 2;
        t.zhalo = 1;
    }
    t.threads = ceil_div((int64_t)(rows + 2) * t.pitch, 64) * 64;
    return true;
}

// Default own rows per block for `kernel` when whole rows fit: of 4 to 12 rows, the shape that
// keeps the most own lanes resident on a CU, (workgroups per CU) x rows x lane groups, the
// workgroups per CU being what the kernel's registers and LDS allow for that block size; on
// ties the fewer rows (more workgroups). A block whose wavefronts do not divide the CU's
// wavefront slots leaves slots idle: with two wavefronts per SIMD, 4-wavefront blocks run two
// per CU and 5-wavefront blocks one, so 8 rows of 22 to 24 lane groups (4 wavefronts) beat 6
// rows (3 wavefronts, two per CU), measured 10 to 18 % faster, and 6 rows of 32 (4 wavefronts)
// beat 8 rows (5), 25 % faster. Row segments are left to the fixed rule (there the count of
// resident lanes ranked shapes wrongly: 12 rows of segments against 8 at 256^3 fp64 measured
// 9 % slower). 0 when no whole rows fit or the occupancy cannot be read.
inline int default_rows(const int groups, const void *kernel) {
    int best_rows = 0;
    int64_t best = 0;
    for (int rows = 4; rows <= 12; rows += 2) {
        Tiling2 s;
        if (!rows_shape(groups, rows, s) || s.zhalo != 0) continue;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, s.threads, 0) !=
            hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        const int64_t resident = (int64_t)per_cu * rows * s.seg;
        if (resident > best) {
            best = resident;
            best_rows = rows;
        }
    }
    return best_rows;
}

// Shape of the two-step sweep: R = 2 tile_j own rows (rows_shape) and chunk_x planes per march.
// Defaults: the rows default_rows picks for `kernel` when it is given, else 6 rows, or 4 when
// only 4 whole rows fit, or 8 in row segments; default_march_planes planes. False when nz is
// not a multiple of the lane width or no segment fits.
inline bool make_tiling2(const acfdtd_hip_desc &d, int vec, Tiling2 &t,
                         const void *kernel = nullptr) {
    if (row_cells(d) % vec != 0) return false;
    t.vec = vec;
    t.groups = row_cells(d) / vec;
    int rows = d.tile_j > 0 ? 2 * d.tile_j : 0;
    if (rows == 0 && kernel != nullptr) rows = default_rows(t.groups, kernel);
    if (rows == 0) {
        if (8 * t.groups <= kFused2Threads)
            rows = 6;  // whole rows, 6 of them
        else if (6 * t.groups <= kFused2Threads)
            rows = 4;  // whole rows, 4 of them: without z halo lanes
        else
            rows = 8;  // row segments: taller blocks spread the y halo rows thinner
    }
    if (!rows_shape(t.groups, rows, t)) return false;
    t.nblk_g = ceil_div(t.groups, t.seg);
    t.nblk_j = ceil_div(d.ny, t.rows);
    const int64_t nblk = (int64_t)t.nblk_g * t.nblk_j;
    t.chunk = d.chunk_x > 0 ? d.chunk_x : default_march_planes(d.nx, nblk, t.threads / 64, 2);
    t.nchunk = ceil_div(d.nx, t.chunk);
    const int64_t tiles = nblk * t.nchunk;
    if (tiles > (int64_t)INT32_MAX - 8) return false;
    t.nblk = (int)nblk;
    t.tiles = (int)tiles;
    t.launched = ceil_div(tiles, 8) * 8;
    t.tape_vec = d.wall_tape != nullptr &&
                 reinterpret_cast<uintptr_t>(d.wall_tape) % ((size_t)vec * d.real_bytes) == 0 &&
                 d.wall_tape_stride % vec == 0;
    t.extra = -1;
    return true;
}

// Workgroups of a two-step sweep run one per CU (8 wavefronts of up to 256 VGPRs) or a few
// (smaller blocks), in rounds. When the tiles overrun a whole number of rounds, the last round
// leaves CUs idle while its workgroups march; instead, split x into as many chunks as fill whole
// rounds, every row block into nchunk or nchunk + 1 chunks of nearly equal length (the first
// `extra` row blocks one more). The tile ids keep their order (row block fastest), so the
// workgroups of one XCD still share their halo rows in its L2. Unchanged when the tiles fit one
// round or the chunks would fall below two planes.
inline void balance_tiling2(Tiling2 &t, const int nx) {
    const int waves = t.threads / 64;
    const int64_t units = (int64_t)device_compute_units() * (waves >= 8 ? 1 : 8 / waves);
    if (t.tiles <= units) return;
    const int64_t total = (t.tiles + units - 1) / units * units;
    const int64_t q = total / t.nblk, r = total % t.nblk;
    if (q < 1 || nx / (q + (r > 0 ? 1 : 0)) < 2 || total > (int64_t)INT32_MAX - 8) return;
    t.nchunk = (int)q;
    t.extra = (int)r;
    t.tiles = (int)total;
    t.launched = ceil_div(total, 8) * 8;
}

// Per-lane constants of the step: position, neighbour existence, sponge factors, wall classes.
template <typename real, int V>
struct LaneConsts {
    // Source and receiver tables (LDS copies when they fit, see fused2_step).
    const int32_t *src_cell, *rx_cell;
    const int64_t *src_len;
    const real *src_now;  // the sources' samples of the sweep's two steps, two per source
    int ev_lo, ev_hi;     // planes of the lane's sources and receivers (empty: lo > hi)
    int j, k0, jk0;
    int ztop;      // the lane's first padding cell, its upper wall face (padded rows; else >= V)
    bool partial;  // padded rows: the lane holds padding (ztop < V)
    bool has_ylo, has_yhi, has_zlo, has_zhi, yz_boundary, special;
    int64_t dy_lo, dy_hi, dz_lo, dz_hi, dz_top;
    real fyc, fyf_lo, fyf_hi, fzc[V], fzf[V + 1];
    int code_yz[V];
    bool wall_in[V];
    real fu_in[V], fp_in[V];
};

// PADDED: rows may be padded (see fused_march_vec): a lane may end in padding cells, whose
// sponge factors are read clamped to the room's profiles.
template <typename real, int V, bool WALLS, bool LAYER, bool PADDED = false>
__device__ __forceinline__ void lane_consts(const SweepArgs<real> &a, const int j, const int k0,
                                            const int32_t *src_cell, const int64_t *src_len,
                                            const real *src_now, const int32_t *rx_cell,
                                            LaneConsts<real, V> &c) {
    const int ny = a.ny, nz = a.nz;
    c.src_cell = src_cell;
    c.src_len = src_len;
    c.src_now = src_now;
    c.rx_cell = rx_cell;
    c.j = j;
    c.k0 = k0;
    c.jk0 = j * a.zp + k0;
    c.ztop = nz - k0;
    c.partial = PADDED && c.ztop < V;
    c.has_ylo = j > 0;
    c.has_yhi = j + 1 < ny;
    c.has_zlo = k0 > 0;
    c.has_zhi = k0 + V < nz;
    c.dy_lo = c.has_ylo ? a.zp : 0;
    c.dy_hi = c.has_yhi ? a.zp : 0;
    c.dz_lo = c.has_zlo ? 1 : 0;
    c.dz_hi = c.has_zhi ? V : V - 1;
    c.dz_top = c.partial ? V - 1 : V;
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
        for (int e = 0; e < V; ++e) c.fzc[e] = a.cell[2][PADDED ? min(k0 + e, nz - 1) : k0 + e];
#pragma unroll
        for (int e = 0; e <= V; ++e) c.fzf[e] = a.face[2][PADDED ? min(k0 + e, nz) : k0 + e];
    }
    c.yz_boundary = !c.has_ylo || !c.has_yhi || !c.has_zlo || !c.has_zhi;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        c.code_yz[e] = (wall_class(j, ny) << 2) | (wall_class(k0 + e, nz) << 4);
        c.wall_in[e] = WALLS && ((a.wall_mask >> c.code_yz[e]) & 1ull);
        c.fu_in[e] = 1;
        c.fp_in[e] = 0;
    }
    if (WALLS && c.yz_boundary) {
#pragma unroll
        for (int e = 0; e < V; ++e)
            if (c.wall_in[e]) {
                c.fu_in[e] = a.wall_table[c.code_yz[e]];
                c.fp_in[e] = a.wall_table[ACFDTD_HIP_WALL_CODES + c.code_yz[e]];
            }
    }
    int lo = a.nx, hi = -1;
    for (int s = 0; s < a.n_sources; ++s) {
        const int e = c.src_cell[2 * s + 1] - c.jk0;
        if (e >= 0 && e < V) {
            lo = min(lo, c.src_cell[2 * s]);
            hi = max(hi, c.src_cell[2 * s]);
        }
    }
    for (int r = 0; r < a.n_receivers; ++r) {
        const int e = c.rx_cell[2 * r + 1] - c.jk0;
        if (e >= 0 && e < V) {
            lo = min(lo, c.rx_cell[2 * r]);
            hi = max(hi, c.rx_cell[2 * r]);
        }
    }
    c.ev_lo = lo;
    c.ev_hi = hi;
    c.special = hi >= 0;
}

// Whether plane i of the lane holds a source or a receiver.
template <typename real, int V>
__device__ __forceinline__ bool has_events(const LaneConsts<real, V> &c, const int i) {
    return c.special && i >= c.ev_lo && i <= c.ev_hi;
}

// What one step leaves in a lane for one plane: the new pressure, the new velocity on the
// cell faces the next step needs again (upper x, both y, all z), the new relaxation states.
template <typename real, int V, int NR>
struct PlaneState {
    real p[V], vx_hi[V], vy_lo[V], vy_hi[V], vz[V + 1];
    real psi[NR > 0 ? NR : 1][V];
};

// One step of the lane's V cells in plane i, the arithmetic of fused_march_vec's plane body:
// from the pressures p_c (plane i), p_up (plane i + 1), q_lo / q_hi (rows j -/+ 1), pz_lo /
// pz_hi (cells k0 - 1 and k0 + V), the new velocity on the lower x face, the old velocities on
// the other faces and the relaxation states (advanced in place), produce the new velocities,
// the new pressure after walls, sponge and sources of `step`, and the boundary tape's value
// p_tmp + p_old (the pressure before the wall closed form plus the old pressure).
template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED, bool XWALL,
          bool PLAINXY = false, bool PADDED = false>
__device__ __forceinline__ void advance_cells(
    const SweepArgs<real> &a, const LaneConsts<real, V> &c, const int i, const int64_t step,
    const real (&p_c)[V], const real (&p_up)[V], const real (&q_lo)[V], const real (&q_hi)[V],
    const real pz_lo, const real pz_hi, const real (&vx_lo)[V], const real (&vx_v)[V],
    const real (&vy_lo_in)[V], const real (&vy_hi_in)[V], const real (&vz_v)[V],
    const real vz_top, real (&states)[NR > 0 ? NR : 1][V], PlaneState<real, V, NR> &o,
    real (&tape_value)[V], real (&div_value)[V]) {
    const int nx = a.nx;
    const real vel = a.vel;
    const bool up_interior = !XWALL || i + 1 < nx;
    // PLAINXY: every x and y sponge factor of the plane and row is 1, so their products are
    // skipped (v * 1 is v exactly) and only the z factor is applied.
    const real fxc = LAYER && !PLAINXY ? uniform_load(a.cell[0] + i) : real(1);
    const real fxf_hi = LAYER && !PLAINXY ? uniform_load(a.face[0] + i + 1) : real(1);
    auto damp = [&](const real v, const real fx, const real fy, const real fz)
                    __attribute__((always_inline)) {
        if (PLAINXY) return LAYER ? v * fz : v;
        return damp3<real, LAYER>(v, fx, fy, fz);
    };

#pragma unroll
    for (int e = 0; e < V; ++e) {
        const real wx = up_interior ? vx_v[e] - vel * (p_up[e] - p_c[e]) : vx_v[e];
        o.vx_hi[e] = damp(wx, fxf_hi, c.fyc, c.fzc[e]);
        const real wl = c.has_ylo ? vy_lo_in[e] - vel * (p_c[e] - q_lo[e]) : vy_lo_in[e];
        o.vy_lo[e] = damp(wl, fxc, c.fyf_lo, c.fzc[e]);
        const real wh = c.has_yhi ? vy_hi_in[e] - vel * (q_hi[e] - p_c[e]) : vy_hi_in[e];
        o.vy_hi[e] = damp(wh, fxc, c.fyf_hi, c.fzc[e]);
        real wz;
        if (e > 0) {
            wz = vz_v[e] - vel * (p_c[e] - p_c[e - 1]);
            if (PADDED && e == c.ztop) wz = vz_v[e];  // padded rows: the upper wall face
        } else {
            wz = c.has_zlo ? vz_v[0] - vel * (p_c[0] - pz_lo) : vz_v[0];
        }
        o.vz[e] = damp(wz, fxc, c.fyc, c.fzf[e]);
    }
    {
        const real wz = c.has_zhi ? vz_top - vel * (pz_hi - p_c[V - 1]) : vz_top;
        o.vz[V] = damp(wz, fxc, c.fyc, c.fzf[V]);
    }

    const int cx = XWALL ? wall_class(i, nx) : 0;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const real div =
            ((o.vx_hi[e] - vx_lo[e]) + (o.vy_hi[e] - o.vy_lo[e])) + (o.vz[e + 1] - o.vz[e]);
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
            bool is_wall = c.wall_in[e];
            real fu = c.fu_in[e], fp = c.fp_in[e];
            if (XWALL) {
                const int code = c.code_yz[e] | cx;
                is_wall = (a.wall_mask >> code) & 1ull;
                if (is_wall) {
                    fu = a.wall_table[code];
                    fp = a.wall_table[ACFDTD_HIP_WALL_CODES + code];
                }
            }
            if (is_wall) pn = pn * fu - p_old * fp;
        }
        if (LAYER) pn = PLAINXY ? pn * c.fzc[e] : ((pn * fxc) * c.fyc) * c.fzc[e];
        o.p[e] = pn;
    }
    if (has_events(c, i)) {
        // Sources in insertion order, after the sponge (scheme.py:360-363).
        for (int s = 0; s < a.n_sources; ++s) {
            const int e = c.src_cell[2 * s + 1] - c.jk0;
            if (c.src_cell[2 * s] == i && e >= 0 && e < V && step < c.src_len[s]) {
                lane_add<real, V>(o.p, e, a.src * c.src_now[2 * s + (step & 1)]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int e = 0; e < V; ++e) o.psi[r][e] = states[r][e];
}

// Boundary tape of `step` for the lane's wall cells in plane i: the values fused_march_vec
// writes, cell for cell, with one test per face row instead of one per cell and face. A lane's
// V cells share their x and y faces, so its x and y face rows take one vector store each when
// the tape rows are V-aligned (nz, and so every face row offset, is a multiple of V); only
// the first and last lane of a row hold z-face cells, one each.
template <typename real, int V, bool XWALL>
__device__ __forceinline__ void tape_cells(const SweepArgs<real> &a, const LaneConsts<real, V> &c,
                                           const int i, const int64_t step, const bool vec_rows,
                                           const real (&tape_value)[V]) {
    if (!(XWALL || c.yz_boundary)) return;
    real *row = a.wall_tape + (step - a.wall_tape_step0) * a.wall_tape_stride;
    const TapePlanes t(a.nx, a.ny, a.nz);
    auto face_row = [&](real *dst) __attribute__((always_inline)) {
        if (vec_rows) {
            vstore<real, V>(dst, tape_value);
        } else {
#pragma unroll
            for (int e = 0; e < V; ++e) dst[e] = tape_value[e];
        }
    };
    if (XWALL) {
        const int64_t yz = (int64_t)c.j * a.nz + c.k0;
        if (i == 0) face_row(row + t.x_lo + yz);
        if (i == a.nx - 1) face_row(row + t.x_hi + yz);
    }
    const int64_t xz = (int64_t)i * a.nz + c.k0;
    if (c.j == 0) face_row(row + t.y_lo + xz);
    if (c.j == a.ny - 1) face_row(row + t.y_hi + xz);
    const int64_t xy = (int64_t)i * a.ny + c.j;
    if (c.k0 == 0) row[t.z_lo + xy] = tape_value[0];
    if (c.k0 + V == a.nz) row[t.z_hi + xy] = tape_value[V - 1];
}

template <typename real, int V>
__device__ __forceinline__ void record_cells(const SweepArgs<real> &a, const LaneConsts<real, V> &c,
                                             const int i, const int64_t column,
                                             const real (&p)[V]) {
    for (int r = 0; r < a.n_receivers; ++r) {
        const int e = c.rx_cell[2 * r + 1] - c.jk0;
        if (c.rx_cell[2 * r] == i && e >= 0 && e < V) {
            a.recording[r * a.recording_stride + column] = lane_get<real, V>(p, e);
        }
    }
}

// The march of one lane over planes [i0, i1). `valid`: the lane maps to cells of the grid;
// `owner`: it is one of the workgroup's own lanes (not halo), which alone store and record.
// Every thread of the workgroup runs the same loop and reaches every barrier.
// The plain (not staged) two-step sweep with 16-byte lanes (4 fp32 or 2 fp64 cells) loads
// plane i + 2's first-step inputs into registers right after step 1 of plane i + 1, so that
// they travel during the barrier and step 2 of plane i (measured on fp32 2 to 15 % faster from
// 146 x 117 x 88 to 384^3, one case 5 % slower; on fp64 without a sponge 10 to 19 % faster).
// It takes up to 43 more VGPRs per lane, which fit under the 256 of two wavefronts per SIMD
// with up to three relaxation states, and in fp64 only without the sponge's factors; beyond
// that the kernel would spill. On narrower lanes it measured slower (up to 32 %). Where it does
// not apply, each plane's inputs are loaded when they are needed.
#ifndef ACFDTD_FUSED2_PREFETCH
#define ACFDTD_FUSED2_PREFETCH 1
#endif

// First-step inputs of one plane for one lane.
template <typename real, int V, int NR>
struct Step1In {
    real vx_v[V], p_up[V], vy_lo[V], vy_hi[V], q_lo[V], q_hi[V], vz_v[V];
    real vz_top, pz_lo, pz_hi;
    real psi[NR > 0 ? NR : 1][V];
};

// Fields of a plane's first-step inputs staged in LDS when STAGED: vx, p above, vy below and
// above, p of the rows below and above, the relaxation states (each 16 bytes per lane).
template <int NR>
constexpr int staged_fields() {
    return 6 + NR;
}
// 4-byte words staged per lane: the V values of vz (its rows are nz + 1 long, so not 16-byte
// aligned), vz's top value and the two z-neighbour pressures.
template <typename real, int V>
constexpr int staged_words() {
    return (V + 3) * (int)(sizeof(real) / 4);
}
// Bytes of a workgroup's stage.
template <typename real, int V, int NR>
constexpr int stage_bytes() {
    return staged_fields<NR>() * kStageField + staged_words<real, V>() * kStageWord;
}

// PADDED: rows may be padded (z_pitch), see fused_march_vec; not with STAGED or a tape.
// DIVT: owner lanes write both steps' divergences to the divergence tape (a.div_tape, when
// there is one) and the fields of their layer-frame cells to the layer tape (a.layer_tape,
// when there is one).
template <typename real, int V, int NR, bool WALLS, bool LAYER, bool COMPENSATED, bool STAGED,
          bool PADDED = false, bool DIVT = false>
__device__ __forceinline__ void fused2_march(
    const SweepArgs<real> &a, const Tiling2 &t2, const int i0, const int i1, const int j,
    const int k0, const bool valid, const bool owner, const int t, const int64_t step,
    const int64_t column, const real *__restrict__ p_in, const real *__restrict__ vx_in,
    const real *__restrict__ vy_in, const real *__restrict__ vz_in,
    const real *__restrict__ psi_in, real *__restrict__ p_out, real *__restrict__ vx_out,
    real *__restrict__ vy_out, real *__restrict__ vz_out, real *__restrict__ psi_out,
    real *__restrict__ lds, const int32_t *src_cell, const int64_t *src_len,
    const real *src_now, const int32_t *rx_cell, unsigned char *__restrict__ stage) {
    using Wall = std::integral_constant<bool, true>;
    using Interior = std::integral_constant<bool, false>;
    static_assert(!(PADDED && STAGED), "the staged sweep takes unpadded rows");
    static_assert(!(DIVT && (PADDED || STAGED)), "the divergence tape is a plain-sweep option");
    const int nx = a.nx, zp = a.zp;  // zp: the row length (the room's nz unless padded)
    const Layout L = room_layout(a, j, k0);
    const int64_t n_cells = L.plane * nx;
    const real vel = a.vel;
    const int G = t2.pitch, T = t2.threads;  // G: lanes between y neighbours
    LaneConsts<real, V> c;
    if (valid)
        lane_consts<real, V, WALLS, LAYER, PADDED>(a, j, k0, src_cell, src_len, src_now, rx_cell, c);
    // The lane constants' loads complete here, on every path into the march. Lanes that skip
    // the prologue would otherwise reach the loop with those loads counted as outstanding,
    // and the compiler waits for them inside the loop with a memory count of zero, which on
    // this counter also waits for every store issued before (the boundary tape's).
    wait_memory();
    // Whether every x and y sponge factor of plane x is 1 for the whole wavefront (rows are
    // whole wavefronts or less, and outside the sponge most are): the step then skips those
    // products, which leave a value unchanged (x * 1 is x), and is the same bits with fewer
    // instructions. Uniform across the wavefront, so it selects a code path, not lanes.
    const bool lane_row_plain =
        !LAYER || !valid || (c.fyc == real(1) && c.fyf_lo == real(1) && c.fyf_hi == real(1));
    const bool wave_rows_plain = LAYER && __ballot(!lane_row_plain) == 0;
    auto plain_xy = [&](const int x) __attribute__((always_inline)) {
        if constexpr (!LAYER) {
            return false;
        } else {
            return wave_rows_plain && uniform_load(a.cell[0] + x) == real(1) &&
                   uniform_load(a.face[0] + x + 1) == real(1);
        }
    };

    auto slot = [&](const int plane, const int lane) __attribute__((always_inline)) {
        return lds + ((plane % 3) * T + lane) * V;
    };
    // Layer tape: the fields of plane s after step n (o, and vx_lo on the lower x face) for the
    // lane's cells when they lie in the layer's frame, and the far faces they own.
    const int64_t lt_pos = DIVT && valid && a.layer_tape ? a.layer.in_frame(j, k0) : -1;
    auto layer_write = [&](const int s, const int64_t n, const PlaneState<real, V, NR> &o,
                           const real (&vx_lo)[V]) __attribute__((always_inline)) {
        if constexpr (DIVT) {
            if (a.layer_tape == nullptr) return;
            real *row = a.layer_tape + (n - a.layer_tape_step0) * a.layer_tape_stride;
            const LayerFrame &f = a.layer;
            const int64_t q = f.offset(s, j, k0, lt_pos);
            if (q >= 0) {
                real lower[V];
#pragma unroll
                for (int e = 0; e < V; ++e) lower[e] = o.vz[e];
                vstore<real, V>(row + q, o.p);
                vstore<real, V>(row + f.cells + q, vx_lo);
                vstore<real, V>(row + 2 * f.cells + q, o.vy_lo);
                vstore<real, V>(row + 3 * f.cells + q, lower);
            }
            if (s == nx - 1) vstore<real, V>(row + f.far_x() + (int64_t)j * a.nz + k0, o.vx_hi);
            if (!c.has_yhi) vstore<real, V>(row + f.far_y() + (int64_t)s * a.nz + k0, o.vy_hi);
            if (!c.has_zhi) row[f.far_z() + (int64_t)s * a.ny + j] = o.vz[V];
        }
    };
    auto share = [&](const int plane, const real (&p)[V]) __attribute__((always_inline)) {
        real *d = slot(plane, t);
        vstore<real, V>(d, p);
    };

    // First step: p^n of the plane being advanced, and the new lower x face velocity.
    real pn_c[V], s1_vx_lo[V];
    auto first_face = [&](const int s) __attribute__((always_inline)) {
        real v[V], prev[V];
        vload<real, V>(p_in + L.cell(s), pn_c);
        vload<real, V>(vx_in + L.cell(s), v);
        vload<real, V>(p_in + L.cell(s > 0 ? s - 1 : s), prev);
        const real fxf = LAYER ? uniform_load(a.face[0] + s) : real(1);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real w = s > 0 ? v[e] - vel * (pn_c[e] - prev[e]) : v[e];
            s1_vx_lo[e] = damp3<real, LAYER>(w, fxf, c.fyc, c.fzc[e]);
        }
    };
    // Step 1's inputs of plane s from global memory.
    auto load_direct = [&](const int s, Step1In<real, V, NR> &in, auto xwall_tag) __attribute__((always_inline)) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        const int64_t cc = L.cell(s);
        const bool up_interior = !XWALL || s + 1 < nx;
        vload<real, V>(vx_in + cc + L.plane, in.vx_v);
        vload<real, V>(p_in + cc + (up_interior ? L.plane : 0), in.p_up);
        const int64_t cy = L.face_y(s);
        vload<real, V>(vy_in + cy, in.vy_lo);
        vload<real, V>(vy_in + cy + zp, in.vy_hi);
        vload<real, V>(p_in + cc - c.dy_lo, in.q_lo);
        vload<real, V>(p_in + cc + c.dy_hi, in.q_hi);
        const int64_t cz = L.face_z(s);
        vload<real, V>(vz_in + cz, in.vz_v);
        in.vz_top = vz_in[cz + (PADDED ? c.dz_top : V)];
        in.pz_lo = p_in[cc - c.dz_lo];
        in.pz_hi = p_in[cc + c.dz_hi];
#pragma unroll
        for (int r = 0; r < NR; ++r) vload<real, V>(psi_in + (int64_t)r * n_cells + cc, in.psi[r]);
    };
    // STAGED: the same inputs of plane s requested one plane ahead, straight into this lane's
    // LDS slots (nothing held in registers while they travel): the 16-byte aligned fields as
    // 16-byte loads, vz (its rows are nz + 1 long, so not aligned) and the single values as
    // 4-byte words. Layout: see StageRead16.
    constexpr int NF = staged_fields<NR>();
    constexpr int NS = staged_words<real, V>();
    constexpr int WPR = (int)(sizeof(real) / 4);  // 4-byte words per value
    const bool wave_owner = __ballot(owner) != 0;
    auto stage_issue = [&](const int s, auto xwall_tag) __attribute__((always_inline)) {
        if constexpr (STAGED) {
            constexpr bool XWALL = decltype(xwall_tag)::value;
            const int64_t cc = L.cell(s);
            const bool up_interior = !XWALL || s + 1 < nx;
            unsigned char *base = stage + (t & ~63) * 16;  // this wavefront's slots of field 0
            load_to_lds16(vx_in + cc + L.plane, base);
            load_to_lds16(p_in + cc + (up_interior ? L.plane : 0), base + kStageField);
            const int64_t cy = L.face_y(s);
            load_to_lds16(vy_in + cy, base + 2 * kStageField);
            load_to_lds16(vy_in + cy + zp, base + 3 * kStageField);
            load_to_lds16(p_in + cc - c.dy_lo, base + 4 * kStageField);
            load_to_lds16(p_in + cc + c.dy_hi, base + 5 * kStageField);
#pragma unroll
            for (int r = 0; r < NR; ++r)
                load_to_lds16(psi_in + (int64_t)r * n_cells + cc, base + (6 + r) * kStageField);
            unsigned char *words = stage + NF * kStageField + (t & ~63) * 4;
            auto word_load = [&](const real *src, const int w) __attribute__((always_inline)) {
#pragma unroll
                for (int h = 0; h < WPR; ++h)
                    load_to_lds4(reinterpret_cast<const unsigned char *>(src) + 4 * h,
                                 words + (w * WPR + h) * kStageWord);
            };
            const int64_t cz = L.face_z(s);
#pragma unroll
            for (int e = 0; e <= V; ++e) word_load(vz_in + cz + e, e);
            word_load(p_in + cc - c.dz_lo, V + 1);
            word_load(p_in + cc + c.dz_hi, V + 2);
        }
    };
    auto stage_read = [&](Step1In<real, V, NR> &in) __attribute__((always_inline)) {
        if constexpr (STAGED) {
            // The stage was requested before the previous plane's second step, whose 3 + NR
            // unconditional stores (vy, vz, psi, p) a wavefront with an owner lane issued after it:
            // with at most that many operations outstanding, the stage has landed. A wavefront
            // with no owner lane (halo rows) stored nothing since, and waits for everything.
            if (wave_owner) wait_vm<3 + NR>();
            else wait_memory();
            const uint32_t a16 = lds_address(stage) + (uint32_t)t * 16;
            u32x4 raw[NF];
            StageRead16<NF>::run(a16, a16 + 4 * kStageField, raw);
            uint32_t w[NS];
            StageRead4<NF, NS>::run(lds_address(stage) + NF * kStageField + (uint32_t)t * 4, w);
            static_assert(sizeof(real) * V == sizeof(u32x4), "a lane's staged field is 16 bytes");
            auto take = [&](const int f, real (&dst)[V]) __attribute__((always_inline)) {
                __builtin_memcpy(dst, &raw[f], sizeof(u32x4));
            };
            take(0, in.vx_v);
            take(1, in.p_up);
            take(2, in.vy_lo);
            take(3, in.vy_hi);
            take(4, in.q_lo);
            take(5, in.q_hi);
#pragma unroll
            for (int r = 0; r < NR; ++r) take(6 + r, in.psi[r]);
            auto value = [&](const int k) __attribute__((always_inline)) {
                real v;
                __builtin_memcpy(&v, &w[k * WPR], sizeof(real));
                return v;
            };
#pragma unroll
            for (int e = 0; e < V; ++e) in.vz_v[e] = value(e);
            in.vz_top = value(V);
            in.pz_lo = value(V + 1);
            in.pz_hi = value(V + 2);
        }
    };
    auto stage_issue_at = [&](const int s) __attribute__((always_inline)) {
        if (s == nx - 1) stage_issue(s, Wall{});
        else stage_issue(s, Interior{});
    };
    // Step 1 of plane s from its inputs into o (p^n of plane s in pn_c, new lower face in
    // s1_vx_lo); then carry plane s + 1's p^n and lower face.
    auto compute1 = [&](const int s, Step1In<real, V, NR> &in, PlaneState<real, V, NR> &o,
                        auto xwall_tag) __attribute__((always_inline)) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        real tape_value[V], div_value[V];
        auto run = [&](auto plain_tag) __attribute__((always_inline)) {
            advance_cells<real, V, NR, WALLS, LAYER, COMPENSATED, XWALL, decltype(plain_tag)::value,
                          PADDED>(
                a, c, s, step, pn_c, in.p_up, in.q_lo, in.q_hi, in.pz_lo, in.pz_hi, s1_vx_lo,
                in.vx_v, in.vy_lo, in.vy_hi, in.vz_v, in.vz_top, in.psi, o, tape_value, div_value);
        };
        if (plain_xy(s)) run(Wall{});
        else run(Interior{});
        if (a.wall_tape && owner && s >= i0 && s < i1)
            tape_cells<real, V, XWALL>(a, c, s, step, t2.tape_vec != 0, tape_value);
        if (DIVT && a.div_tape && owner && s >= i0 && s < i1)
            vstore<real, V>(a.div_tape + (step - a.div_tape_step0) * n_cells + L.cell(s), div_value);
        if (DIVT && owner && s >= i0 && s < i1) layer_write(s, step, o, s1_vx_lo);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            pn_c[e] = in.p_up[e];
            s1_vx_lo[e] = o.vx_hi[e];
        }
    };
    auto step1 = [&](const int s, PlaneState<real, V, NR> &o, auto xwall_tag) __attribute__((always_inline)) {
        Step1In<real, V, NR> in;
        load_direct(s, in, xwall_tag);
        compute1(s, in, o, xwall_tag);
    };
    auto step1_at = [&](const int s, PlaneState<real, V, NR> &o) __attribute__((always_inline)) {
        if (s == 0 || s == nx - 1) step1(s, o, Wall{});
        else step1(s, o, Interior{});
    };
    auto load_at = [&](const int s, Step1In<real, V, NR> &in) __attribute__((always_inline)) {
        if (s == nx - 1) load_direct(s, in, Wall{});
        else load_direct(s, in, Interior{});
    };
    auto compute1_at = [&](const int s, Step1In<real, V, NR> &in, PlaneState<real, V, NR> &o) __attribute__((always_inline)) {
        if (s == 0 || s == nx - 1) compute1(s, in, o, Wall{});
        else compute1(s, in, o, Interior{});
    };
    // Step 1 of plane s from the stage (requested one plane earlier).
    auto step1_staged_at = [&](const int s, PlaneState<real, V, NR> &o) __attribute__((always_inline)) {
        Step1In<real, V, NR> in;
        stage_read(in);
        if (s == nx - 1) compute1(s, in, o, Wall{});
        else compute1(s, in, o, Interior{});
    };

    // Second step of plane i from the first step's results: A (plane i), p_next (p^(n+1) of
    // plane i + 1), the lower face's new velocity s2_vx_lo and the neighbours in LDS.
    real s2_vx_lo[V];
    auto step2 = [&](const int i, PlaneState<real, V, NR> &A, const real (&p_next)[V],
                     auto xwall_tag) __attribute__((always_inline)) {
        constexpr bool XWALL = decltype(xwall_tag)::value;
        const int64_t cc = L.cell(i);
        real q_lo[V], q_hi[V];
        vload<real, V>(slot(i, t - G), q_lo);
        vload<real, V>(slot(i, t + G), q_hi);
        const real pz_lo = slot(i, t - 1)[V - 1];
        const real pz_hi = slot(i, t + 1)[0];
        real vz_v[V];
#pragma unroll
        for (int e = 0; e < V; ++e) vz_v[e] = A.vz[e];
        PlaneState<real, V, NR> o;
        real tape_value[V], div_value[V];
        auto run = [&](auto plain_tag) __attribute__((always_inline)) {
            advance_cells<real, V, NR, WALLS, LAYER, COMPENSATED, XWALL, decltype(plain_tag)::value,
                          PADDED>(
                a, c, i, step + 1, A.p, p_next, q_lo, q_hi, pz_lo, pz_hi, s2_vx_lo, A.vx_hi,
                A.vy_lo, A.vy_hi, vz_v, A.vz[V], A.psi, o, tape_value, div_value);
        };
        if (plain_xy(i)) run(Wall{});
        else run(Interior{});
        if (a.wall_tape) tape_cells<real, V, XWALL>(a, c, i, step + 1, t2.tape_vec != 0, tape_value);
        if (DIVT && a.div_tape)
            vstore<real, V>(a.div_tape + (step + 1 - a.div_tape_step0) * n_cells + cc, div_value);
        if (DIVT) layer_write(i, step + 1, o, s2_vx_lo);
        const int64_t cy = L.face_y(i), cz = L.face_z(i);
        if (i + 1 < i1 || i + 1 == nx) vstore<real, V>(vx_out + cc + L.plane, o.vx_hi);
        vstore<real, V>(vy_out + cy, o.vy_lo);
        if (!c.has_yhi) vstore<real, V>(vy_out + cy + zp, o.vy_hi);
        {
            real lower[V];
#pragma unroll
            for (int e = 0; e < V; ++e) lower[e] = o.vz[e];
            vstore<real, V>(vz_out + cz, lower);
        }
        if (!c.has_zhi && !(PADDED && c.partial)) vz_out[cz + V] = o.vz[V];
#pragma unroll
        for (int r = 0; r < NR; ++r) vstore<real, V>(psi_out + (int64_t)r * n_cells + cc, o.psi[r]);
        vstore<real, V>(p_out + cc, o.p);
        if (has_events(c, i)) record_cells<real, V>(a, c, i, column + 1, o.p);
#pragma unroll
        for (int e = 0; e < V; ++e) s2_vx_lo[e] = o.vx_hi[e];
    };

    // Prologue: step 1 of the halo plane i0 - 1 (its p^(n+1) only) and of plane i0.
    PlaneState<real, V, NR> A, B;
    // With the divergence tape and the sponge the prefetch would spill (three relaxation
    // states); those load each plane's inputs when they are needed. fp64 with walls, the
    // sponge and three states prefetches although it spills a little (44 bytes per lane):
    // 256^3 measured 65.5 -> 74.6 Gcell/s.
    constexpr bool PREFETCH =
        ACFDTD_FUSED2_PREFETCH && !STAGED && NR <= 3 && !(DIVT && LAYER) &&
        (sizeof(real) == 4 ? V == 4 : V == 2);
    Step1In<real, V, NR> ahead;  // PREFETCH: the next plane's step-1 inputs, in registers
    real p_prev[V];
    real v_half_lo[V];  // v^(n+1/2) on face i0
    if (valid) {
        const int s0 = i0 > 0 ? i0 - 1 : i0;
        first_face(s0);
        if (i0 > 0) {
            step1_at(i0 - 1, B);
#pragma unroll
            for (int e = 0; e < V; ++e) p_prev[e] = B.p[e];
        }
#pragma unroll
        for (int e = 0; e < V; ++e) v_half_lo[e] = s1_vx_lo[e];
        step1_at(i0, A);
        if (owner && has_events(c, i0)) record_cells<real, V>(a, c, i0, column, A.p);
        share(i0, A.p);
        if (STAGED && i0 + 1 < nx) stage_issue_at(i0 + 1);
        if (PREFETCH && i0 + 1 < nx) load_at(i0 + 1, ahead);
    }
    if (STAGED) {
        exchange_barrier();
        wait_memory();  // the first stage, requested just now: no store follows it yet
    } else {
        lds_barrier();
    }
    if (owner) {
        // v^(n+3/2) on face i0, the lower face of the first cell (the prologue of
        // fused_march_vec with the first step's results as inputs).
        const real fxf = LAYER ? uniform_load(a.face[0] + i0) : real(1);
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const real w = i0 > 0 ? v_half_lo[e] - vel * (A.p[e] - p_prev[e]) : v_half_lo[e];
            s2_vx_lo[e] = damp3<real, LAYER>(w, fxf, c.fyc, c.fzc[e]);
        }
        vstore<real, V>(vx_out + L.cell(i0), s2_vx_lo);
    }

    // One plane of the march: step 1 of plane i + 1 into `next`, step 2 of plane i from `here`
    // (the step-1 state of plane i). The march alternates the two states' roles, unrolled by
    // two, instead of copying `next` into `here` after each plane: a copy into registers that
    // a load may still target makes the wavefront wait for that load, and with it for every
    // store issued since (vector memory operations complete in issue order).
    auto plane = [&](const int i, PlaneState<real, V, NR> &here,
                     PlaneState<real, V, NR> &next) __attribute__((always_inline)) {
        const bool more = i + 1 < nx;
        if (valid && more) {
            if (STAGED) step1_staged_at(i + 1, next);
            else if (PREFETCH) compute1_at(i + 1, ahead, next);
            else step1_at(i + 1, next);
            if (owner && i + 1 < i1 && has_events(c, i + 1))
                record_cells<real, V>(a, c, i + 1, column, next.p);
            share(i + 1, next.p);
            // The next iteration's step 1 (plane i + 2), requested now: it travels during the
            // barrier and the second step. The stage is refilled only after this lane's own
            // reads of it are done.
            if (STAGED && i + 2 < nx && i + 1 < i1) {
                wait_lds();
                stage_issue_at(i + 2);
            }
            if (PREFETCH && i + 2 < nx && i + 1 < i1) load_at(i + 2, ahead);
        }
        if (STAGED) exchange_barrier();
        else lds_barrier();
        if (owner) {
            // Selected by value: a select between the two arrays' addresses would keep both
            // plane states in scratch memory.
            real p_next[V];
#pragma unroll
            for (int e = 0; e < V; ++e) p_next[e] = more ? next.p[e] : here.p[e];
            if (i == 0 || i == nx - 1) step2(i, here, p_next, Wall{});
            else step2(i, here, p_next, Interior{});
        }
    };
    for (int i = i0; i < i1; i += 2) {
        plane(i, A, B);
        if (i + 1 < i1) plane(i + 1, B, A);
    }
}

}  // namespace acfdtd
