// Dense 16-row panels of a hybrid plan on FP32 matrix cores.
//
// The reference pads column patterns up to denser kernel patterns by its ILP
// cost model (sbench/ilp_pad/nano_solver.py:76-116) and calls a tile dense
// past a density threshold (sbench/padding_reordering.py:36, 0.30).  A
// hybrid plan takes the 16-row panels at or above its threshold out of the
// register-tiled packing and stores them dense for v_mfma_f32_16x16x4_f32:
// per panel and 16-column K group g, lane l holds the four k-steps it feeds,
//   A[row = l % 16][k = 16 g + 4 j + l / 16],  j = 0..3,
// as one 16-byte word (one load per lane per K group).
//
// A workgroup (16 waves, four per SIMD) owns 4 dense panels x 32 columns of
// C.  K advances in stages of 128 (8 groups); wave w computes panel w % 4
// over groups 2 (w / 4), 2 (w / 4) + 1 of every stage (its K quarter), two
// 16x16 accumulator tiles, and the quarters are added at the end in order
// ((q0 + q1) + q2) + q3.
// The stage's B block (128 rows x 32 columns) is staged in LDS in MFMA
// B-operand order, so each lane reads the four k-steps of a group and tile as
// one 16-byte word:
//   lds[group q][tile t][lane l] = B[16 q + 4 j + l / 16][16 t + l % 16], j = 0..3.
// Stages are double-buffered: the next stage's B rows and A words are loaded
// into registers while the current stage's MFMAs run, then written to the
// other LDS buffer; one barrier per stage.  Accumulation is FP32 in the
// matrix core, k ascending in steps of 4 within each quarter.
//
// Workgroups map to (group of 4 panels, 32-column slice) so that the blocks
// the dispatcher sends to one XCD share 4 panel groups' A and 8 slices' B in
// that XCD's L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "srt_internal.h"

namespace srt {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kNT = 32;       // C columns per workgroup (2 MFMA N tiles)
constexpr int kPanels = 4;    // panels per workgroup
constexpr int kSplit = 4;     // K quarters per panel
constexpr int kWaves = kPanels * kSplit;
constexpr int kStageG = 8;    // K groups of 16 per stage
constexpr int kPartG = kStageG / kSplit;  // groups of a stage per wave
constexpr int kStageK = 16 * kStageG;
constexpr int kThreads = 64 * kWaves;
constexpr int kXcds = 8;
// B staging: a thread loads rows 16 q + 4 j + kr (j = 0..3) of one column and
// writes them as one 16-byte LDS word; (q, kr) pairs per stage: 4 kStageG.
constexpr int kBWords = kStageG * 4 * kNT / kThreads;  // 16-byte words per thread per stage (1)

struct AStage {
    f32x4 w[kPartG];
};

__global__ void __launch_bounds__(kThreads)
    dense_panels_kernel(int ndense, int G, int PG4, int NT, const f32x4 *__restrict__ apack, uint32_t a_bytes,
                        const int32_t *__restrict__ panel_row, const float *__restrict__ B, int64_t ldb,
                        uint32_t b_bytes, float *__restrict__ C, int64_t ldc, int M, int N, int K) {
    __shared__ f32x4 bt[2][kStageG][2][64];
    // XCD-aware block -> tile: blocks b, b + 8, ... run on one XCD; give them
    // consecutive logical tiles (4 panel groups x consecutive slices).
    const int T = PG4 * NT;
    const int b = blockIdx.x;
    const int L = (T % kXcds == 0) ? (b % kXcds) * (T / kXcds) + b / kXcds : b;
    const int pgb = L / (4 * NT), rem = L % (4 * NT);
    const int nt = rem / 4, pg = 4 * pgb + rem % 4;
    if (pg * kPanels >= ndense) return;  // whole workgroup idle (uniform)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int part = wave / kPanels;
    const int d = pg * kPanels + wave % kPanels;  // dense panel of this wave (may be past the list)
    const bool live = d < ndense;
    const int n0 = nt * kNT;
    const int S = (K + kStageK - 1) / kStageK;

    // Loads go through buffer resources: an offset past the resource's extent
    // reads 0, so rows past K, columns past N and groups past G need no branch.
    const __amdgpu_buffer_rsrc_t brsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(B), (short)0, (int)b_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t arsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<f32x4 *>(apack), (short)0, (int)a_bytes, 0x00020000);
    const int bc = threadIdx.x % kNT;  // staged column
    const int bn = n0 + bc;
    const int bt_t = bc / 16, bt_jj = bc % 16;
    typedef f32x4 BRegs[kBWords];
    auto load_b = [&](BRegs &breg, int s) {
#pragma unroll
        for (int e = 0; e < kBWords; ++e) {
            const int idx = threadIdx.x / kNT + (kThreads / kNT) * e;  // (q, kr) pair
            const int q = idx / 4, kr = idx % 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = s * kStageK + 16 * q + 4 * j + kr;
                const uint32_t at = ((uint32_t)k * (uint32_t)ldb + (uint32_t)bn) * 4u;
                const uint32_t off = ((k < K) & (bn < N)) ? at : 0xffffffffu;
                breg[e][j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(brsrc, off, 0, 0));
            }
        }
    };
    auto write_b = [&](const BRegs &breg, int buf) {
#pragma unroll
        for (int e = 0; e < kBWords; ++e) {
            const int idx = threadIdx.x / kNT + (kThreads / kNT) * e;
            const int q = idx / 4, kr = idx % 4;
            bt[buf][q][bt_t][kr * 16 + bt_jj] = breg[e];
        }
    };
    const uint32_t abase = ((uint32_t)(live ? d : 0) * (uint32_t)G * 64u + (uint32_t)lane) * 16u;
    auto load_a = [&](AStage &a, int s) {
#pragma unroll
        for (int q = 0; q < kPartG; ++q) {
            const int g = s * kStageG + part * kPartG + q;
            const uint32_t off = (live & (g < G)) ? abase + (uint32_t)g * 1024u : 0xffffffffu;
            a.w[q] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(arsrc, off, 0, 0));
        }
    };
    f32x4 acc[2];
    acc[0] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    acc[1] = acc[0];
    auto compute = [&](const AStage &a, int buf) {
        // all of the wave's B words of the stage first (one LDS latency per stage)
        f32x4 bw[kPartG][2];
#pragma unroll
        for (int q = 0; q < kPartG; ++q) {
            bw[q][0] = bt[buf][part * kPartG + q][0][lane];
            bw[q][1] = bt[buf][part * kPartG + q][1][lane];
        }
#pragma unroll
        for (int q = 0; q < kPartG; ++q) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w[q][j], bw[q][0][j], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w[q][j], bw[q][1][j], acc[1], 0, 0, 0);
            }
        }
        // keep that order: the LDS reads as one group, then the MFMAs
        __builtin_amdgcn_sched_group_barrier(0x0100, 2 * kPartG, 0);
        __builtin_amdgcn_sched_group_barrier(0x0008, 8 * kPartG, 0);
    };

    // Two stages per iteration so the B register sets and LDS buffers
    // alternate without copies.  Stages past the last (at most one, to make
    // the count even) are loaded, written and computed like the others: their
    // offsets are out of range and read zeros, and adding those zero products
    // leaves every accumulator bit-identical (it starts at +0, so it is never
    // -0).  The loop therefore has a single path, and each of the compiler's
    // waits is for loads issued one stage earlier.
    AStage a0, a1;
    BRegs b0, b1;
    load_b(b0, 0);
    load_a(a0, 0);
    write_b(b0, 0);
    load_b(b1, 1);
    load_a(a1, 1);
    __syncthreads();
    for (int s = 0; s < S; s += 2) {
        compute(a0, 0);
        __builtin_amdgcn_sched_barrier(0);
        write_b(b1, 1);
        load_b(b0, s + 2);
        load_a(a0, s + 2);
        __syncthreads();
        compute(a1, 1);
        __builtin_amdgcn_sched_barrier(0);
        write_b(b0, 0);
        load_b(b1, s + 3);
        load_a(a1, s + 3);
        __syncthreads();
    }
    // quarters 1..3 -> LDS (the stage buffers are free after the last
    // barrier), quarter 0 adds them in order and stores
    f32x4 *red = &bt[0][0][0][0];
    const int slot = ((part - 1) * kPanels + wave % kPanels) * 2 * 64 + lane;
    if (part > 0) {
        red[slot] = acc[0];
        red[slot + 64] = acc[1];
    }
    __syncthreads();
    if (part > 0 || !live) return;
#pragma unroll
    for (int q = 1; q < kSplit; ++q) {
        const int o = ((q - 1) * kPanels + wave % kPanels) * 2 * 64 + lane;
        acc[0] += red[o];
        acc[1] += red[o + 64];
    }
    const int row0 = panel_row[d];
    // D layout: lane l holds rows 4 * (l / 16) + r of column l % 16.
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int n = n0 + 16 * t + (lane & 15);
        if (n < N) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 4 * (lane >> 4) + r;
                if (row < M) C[(int64_t)row * ldc + n] = acc[t][r];
            }
        }
    }
}

}  // namespace

hipError_t launch_dense_panels(int ndense, int G, const void *apack, const int32_t *panel_row, const float *B,
                               int64_t ldb, float *C, int64_t ldc, int M, int N, int K, hipStream_t s) {
    if (ndense <= 0 || N <= 0) return hipSuccess;
    // byte offsets are 32-bit (buffer loads); the extent must stay below the
    // 0xffffffff offset that marks a zero read
    const int64_t a_bytes = (int64_t)ndense * G * 64 * 16, b_bytes = (int64_t)K * ldb * 4;
    if (a_bytes >= 0xffffffffLL || b_bytes >= 0xffffffffLL) return hipErrorInvalidValue;
    const int PG = (ndense + kPanels - 1) / kPanels;
    const int PG4 = (PG + 3) / 4 * 4;
    const int NT = (N + kNT - 1) / kNT;
    hipLaunchKernelGGL(dense_panels_kernel, dim3((unsigned)(PG4 * NT)), dim3(kThreads), 0, s, ndense, G, PG4, NT,
                       static_cast<const f32x4 *>(apack), (uint32_t)a_bytes, panel_row, B, ldb, (uint32_t)b_bytes, C,
                       ldc, M, N, K);
    return hipGetLastError();
}

}  // namespace srt
