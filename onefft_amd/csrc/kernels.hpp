// Host-side view of the kernels: launch parameters, the per-size kernel configuration (shared
// with the device code so the twiddle tables the host builds match what the kernels read), and
// the launch functions. No HIP device code here.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

namespace onefft {

struct cf;
struct cd;

// Largest size that runs as one LDS-resident pass per transform.
constexpr int kMaxSinglePassLog2 = 14;

// Points held per thread by the single-pass kernel of size 2^log2n.
constexpr int batch_log2p(int log2n) { return log2n <= 4 ? log2n : 4; }
// Threads per workgroup of the single-pass kernel of size 2^log2n.
constexpr int batch_wg(int log2n) {
    return log2n <= 4 ? 256 : ((1 << (log2n - 4)) > 256 ? (1 << (log2n - 4)) : 256);
}
// Transforms per workgroup.
constexpr int batch_per_wg(int log2n) { return batch_wg(log2n) / (1 << (log2n - batch_log2p(log2n))); }

struct BatchArgs {
    const cf *in;
    cf *out;
    const cf *tw;        // stage twiddles of this size (stage_tw_size(log2n, batch_log2p(log2n)) entries)
    const cd *tw64;      // FP64 bases of the wide stages (stage_tw64_size entries)
    uint64_t in_dist;    // complex values between transforms
    uint64_t out_dist;
    uint64_t batch;
    // if set (launches of one workgroup only): the workgroup's last act stores `seq` to *done,
    // after its outputs are visible system-wide, so a host thread polling that word in coherent
    // host memory sees the transform complete without waiting for the launch's completion signal
    uint32_t *done;
    uint32_t seq;
};

// Kernel variant of the single-pass transform: nontemporal global access, the LDS layout
// (fft_common.hpp: kPadded / kSwizzled / kHalves, the last from 2^kHalvesMinLog2 points), and the
// level of FP64-base twiddles (0..3, fft_common.hpp: stage_uses_tw64; with the padded and halves
// layouts), and for 32-point contiguous batches whether the pair exchange uses DPP lane swaps. The
// plan's tables serve every variant.
constexpr int kHalvesMinLog2 = 8;
struct BatchVariant {
    bool nt = true;
    int layout = 0;
    int tw64 = 1;
    bool xlane = false;  // 32 points: the transform's two threads exchange by DPP lane swaps, not LDS
};

// One launch: `batch` transforms of 2^log2n points, 0 <= log2n <= kMaxSinglePassLog2.
hipError_t launch_batch(int log2n, bool inverse, const BatchArgs &a, BatchVariant v, hipStream_t stream);

// ---- multi-pass (four-step) transforms, n > 2^kMaxSinglePassLog2 ----------------------------

constexpr int kMaxPasses = 4;
constexpr int kPassMinLog2 = 4;   // smallest sub-transform of a pass
constexpr int kPassMaxLog2 = 12;  // largest sub-transform of a pass

// Points per thread of the pass kernel for a 2^log2nf sub-transform, and the columns per
// workgroup: C adjacent columns make every global access a run of C complex values (C * 8
// bytes). The default C and the choices a plan may make: tiles of at most 8192 points with the
// padded layout (LDS of C columns <= 72 KiB), 16384 with the halves layout, whose tiles need half
// the LDS. The 2^11- and 2^12-point sub-transforms fit only 4 and 2 columns with the padded layout
// (32- and 16-byte runs): they exist for two-pass splits of sizes whose data stays in the MALL,
// where short runs cost less.
constexpr int pass_log2p(int log2nf) { return log2nf < 4 ? log2nf : 4; }
constexpr int pass_tile_max(int layout) { return layout == 2 /* fft_common.hpp: kHalves */ ? 16384 : 8192; }
constexpr bool pass_cols_ok(int log2nf, int cols, int layout = 0) {
    return (cols == 2 || cols == 4 || cols == 8 || cols == 16 || cols == 32 || cols == 64) &&
           (cols << log2nf) <= pass_tile_max(layout) && (cols << (log2nf - pass_log2p(log2nf))) <= 1024 &&
           (cols << (log2nf - pass_log2p(log2nf))) >= 64 && (cols >= 8 || log2nf >= 11);
}
// default: 16 columns (128-byte runs) where that is allowed, else the nearest allowed count
constexpr int pass_cols(int log2nf) {
    return pass_cols_ok(log2nf, 16)   ? 16
           : pass_cols_ok(log2nf, 32) ? 32
           : pass_cols_ok(log2nf, 64) ? 64
           : pass_cols_ok(log2nf, 8)  ? 8
           : pass_cols_ok(log2nf, 4)  ? 4
                                      : 2;
}
constexpr int pass_wg(int log2nf, int cols) { return cols << (log2nf - pass_log2p(log2nf)); }

struct PassArgs {
    const cf *in;
    cf *out;
    const cd *tw64;       // FP64 bases of the sub-transform's stages (fft_common.hpp: stage_has_tw64)
    const cd *tw_lo;      // w_n^i, i < 2^lo_bits (FP64)
    const cd *tw_hi;      // w_n^(i 2^lo_bits), i < n / 2^lo_bits (FP64)
    // in-place passes, the column table: w_n^(x << tw_shift) for x = j (c mod M_f) at thread
    // t = j C + c (T C entries), then x = T (c mod M_f) and x = (N_f / R_0) (c mod M_f) (C each);
    // null for the input pass
    const cd *tw_col;
    uint64_t in_dist;     // complex values between transforms of a batch
    uint64_t out_dist;
    const uint64_t *offsets;  // if set, transform b starts at offsets[b] complex values in both
                              // buffers (a ragged plan's items of one size), not at b * *_dist
    uint32_t lo_bits;
    uint32_t log2n;
    uint32_t log2m;       // low-index extent M_f of this pass (0 for the first pass)
    uint32_t tw_shift;    // log2(n / (N_f M_f)): scales the pass twiddle exponent to w_n
    // first pass only: output row of input column J; digits j_1.. of J (j_1 fastest)
    uint32_t ndig;
    uint32_t dig_log2[kMaxPasses];   // log2 N_i of each digit
    uint32_t dig_shift[kMaxPasses];  // log2 of each digit's weight in the row index
    uint32_t columns;     // n / N_f
    uint32_t strips;      // tile order: XCD-contiguous ranges run as this many strips side by side
    uint32_t col_steps;   // in-place passes: the recurrence's step bases from tw_col as well
};

// One pass of a multi-pass transform: first = reads the input (out of place), else in place on
// `out`. swap_in / swap_out exchange re and im on load / store (the inverse). cols: columns per
// workgroup (pass_cols_ok); layout: the LDS layout of the stage exchanges (kPadded or kHalves).
hipError_t launch_pass(int log2nf, int cols, int layout, bool first, bool swap_in, bool swap_out, const PassArgs &a,
                       uint64_t batch, hipStream_t stream);

// ---- ragged batches: one persistent launch over items of mixed sizes -------------------------
//
// A task covers up to S = kRaggedTaskPoints points (16 per thread of a workgroup) and is one of
//   * up to S / n whole transforms of one size n <= S (single-pass);
//   * one whole two-pass transform above S and below the split size (both passes, tile by tile:
//     a tile is S points, C = S / N_f columns of an N_f-point sub-transform), run by one
//     workgroup, whose second pass then reads its first pass's output from cache;
//   * a run of up to kRaggedTileRun tiles of the first pass of a two-pass transform of the split
//     size or larger (up to 2^20 points), so one large transform is spread over many workgroups;
//   * a run of tiles of such a transform's second pass.
// The queue holds the single-pass, whole and first-pass tasks (phase 1) before the second-pass
// runs (phase 2); a workgroup starts its first phase-2 task only once every phase-1 task's
// results are published device-wide.
constexpr int kRaggedMaxLog2 = 20;
constexpr int kRaggedTaskLog2 = 13;     // 512-thread workgroups; 8192-point tiles give the 2^20
                                        // items' passes runs of 8 values (4096: 4, 1.35x slower)
constexpr uint32_t kRaggedTaskPoints = 1u << kRaggedTaskLog2;
constexpr int kRaggedSingleMaxLog2 = kRaggedTaskLog2;
constexpr int kRaggedWG = kRaggedTaskPoints / 16;
constexpr uint32_t kRaggedTileRun = 2;   // tiles per pass task (at most 256)
constexpr int kRaggedSplitLog2 = 19;     // default split size (ONEFFT_RAGGED_SPLIT overrides)
// default smallest size that a ragged plan runs as offset-batched multi-pass groups instead of in
// the persistent kernel (the tuning table's ragged line or ONEFFT_RAGGED_BATCHED_MIN overrides)
constexpr int kRaggedBatchedMinLog2 = 16;

enum RaggedKind : uint32_t { kRaggedSingle = 0, kRaggedFirst = 1, kRaggedSecond = 2, kRaggedWhole = 3 };

// A task's first word: kind | log2n << 2 | inverse << 7 | (tiles - 1) << 8.
__host__ __device__ constexpr uint32_t ragged_word(uint32_t kind, uint32_t log2n, uint32_t inverse, uint32_t tiles = 1) {
    return kind | (log2n << 2) | (inverse << 7) | ((tiles - 1) << 8);
}

struct RaggedTask {
    uint32_t word;  // ragged_word(kind, log2n, inverse, tiles)
    uint32_t item;  // single: first index into the sorted item list; others: the item's index
    uint32_t arg;   // single: number of transforms; pass runs: the first tile; whole: 0
    uint32_t tw64;  // single: the FP64 stage bases of n, as an offset in the table blob, read per
                    // task so the compiler cannot hoist twiddle loads out of the task loop
};

// Every stage after the first of a ragged column transform is radix 16 and takes its twiddles
// from FP64 bases (fft_common.hpp: stage_uses_tw64, level 3); the offsets below index the blob of
// FP64 values.
struct RaggedSize {
    uint32_t tw64;     // single-pass sizes: offset of the FP64 stage bases
    uint32_t l1, l2;   // two-pass sizes: log2 N_1 (transformed last) and log2 N_2
    uint32_t tw64_1, tw64_2;  // FP64 stage bases of N_1 and N_2
    uint32_t lo, hi;   // offsets of the FP64 twiddle tables of w_n in the FP64 blob
    uint32_t lo_bits;
};

struct RaggedArgs {
    const cf *in;
    cf *out;
    const RaggedTask *tasks;
    const uint64_t *item_off;  // complex offset of each item, in task order
    uint32_t *head;            // [0] next task to claim, [1] phase-1 tasks published
    uint32_t *error;           // set to 1 if the bounded wait for phase 1 gives up
    const uint32_t *counts;    // when set, ntasks and nphase1 are read from counts[0], counts[1]
                               // (a queue built on the device)
    const cd *tw64;            // FP64 stage bases and two-level FP64 tables of w_n
    uint32_t ntasks;
    uint32_t nphase1;          // tasks [0, nphase1) are phase 1
    uint32_t nitems;
    uint32_t spin_limit;       // sleeps of the bounded wait for phase 1 before it gives up
    RaggedSize size[kRaggedMaxLog2 + 1];
};

// Zeroes the queue counters and the error flag (with a kernel, so the pair can be captured in a
// graph), then runs the persistent kernel on `grid` workgroups (0: as many as fit the device at
// once).
hipError_t launch_ragged(const RaggedArgs &a, uint32_t grid, hipStream_t stream);
// The persistent kernel alone, on `grid` workgroups, for a queue whose counters were zeroed and
// whose counts (a.counts) were written by earlier work on `stream`.
hipError_t launch_ragged_kernel(const RaggedArgs &a, uint32_t grid, hipStream_t stream);

// Tasks per single-pass task of 2^log2n points (whole transforms, one column each).
constexpr uint32_t ragged_per_task(int log2n) { return (uint32_t)kRaggedWG >> (log2n - batch_log2p(log2n)); }
// Workgroups of the persistent kernel that fit the device at once.
uint32_t ragged_resident_grid();

}  // namespace onefft
