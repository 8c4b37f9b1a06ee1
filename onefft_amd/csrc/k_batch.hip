// Single-pass batched transforms of up to 2^14 points: one LDS-resident workgroup pass per
// transform (several transforms per workgroup below 4096 points).
//
// This replaces the reference's per-size kernel chain for sizes that fit on chip (small kernels
// src/onefft_avx512.inc n32/n64/n128, first/middle/last passes driven by onefft_execute_passes,
// src/onefft.c:274-293): every transform is read from HBM once and written once. Each thread
// holds 16 points in registers; radix-16 (or a smaller first radix) butterflies run in VGPRs and
// the stages exchange data through padded LDS (Stockham autosort, so the output lands in natural
// order without a bit-reversal pass). The inverse is conj(F(conj x)), the same transform as the
// reference's re/im exchange around F (src/onefft.c:86-88) up to rounding: the negations fold
// into the first and last butterflies' additions, where the exchange cost register moves.
#include "k_batch.hpp"

namespace onefft {

template <int LOG2N, bool XLANE>
static void launch_small_x(bool inverse, bool nt, const BatchArgs &a, dim3 grid, uint64_t total, hipStream_t stream) {
    if (inverse) {
        if (nt) hipLaunchKernelGGL((fft_small_kernel<LOG2N, true, true, XLANE>), grid, dim3(256), 0, stream, a.in, a.out, total, a.done, a.seq);
        else hipLaunchKernelGGL((fft_small_kernel<LOG2N, true, false, XLANE>), grid, dim3(256), 0, stream, a.in, a.out, total, a.done, a.seq);
    } else {
        if (nt) hipLaunchKernelGGL((fft_small_kernel<LOG2N, false, true, XLANE>), grid, dim3(256), 0, stream, a.in, a.out, total, a.done, a.seq);
        else hipLaunchKernelGGL((fft_small_kernel<LOG2N, false, false, XLANE>), grid, dim3(256), 0, stream, a.in, a.out, total, a.done, a.seq);
    }
}

template <int LOG2N>
static hipError_t launch_small(bool inverse, BatchVariant v, const BatchArgs &a, hipStream_t stream) {
    const uint64_t total = a.batch << LOG2N;
    const uint64_t blocks = (total + kSmallBlock - 1) / kSmallBlock;
    if (blocks > 0x7fffffffu || (a.done && blocks != 1)) return hipErrorInvalidValue;
    const dim3 grid((uint32_t)blocks);
    if constexpr (LOG2N == 5) {
        if (v.xlane) {  // the DPP pair exchange exists for 32 points only
            launch_small_x<LOG2N, true>(inverse, v.nt, a, grid, total, stream);
            return hipGetLastError();
        }
    }
    launch_small_x<LOG2N, false>(inverse, v.nt, a, grid, total, stream);
    return hipGetLastError();
}

template <int LOG2N>
static hipError_t launch_batch_n(bool inverse, const BatchArgs &a, BatchVariant v, hipStream_t stream) {
    if constexpr (LOG2N <= 5) {
        // contiguous batches of up to 32 points: coalesced blocks (k_batch.hpp: fft_small_kernel)
        if (a.batch > 0 && a.in_dist == (1u << LOG2N) && a.out_dist == (1u << LOG2N) && a.batch <= (~0ull >> 5))
            return launch_small<LOG2N>(inverse, v, a, stream);
    }
    constexpr int PER = batch_per_wg(LOG2N);
    const uint64_t blocks = (a.batch + PER - 1) / PER;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffu || (a.done && blocks != 1)) return hipErrorInvalidValue;
    const dim3 grid((uint32_t)blocks);
    constexpr bool USES_LDS = stage_count(LOG2N, batch_log2p(LOG2N)) > 1;
    // FP64-base levels (fft_common.hpp: stage_uses_tw64), padded layout only; a level above the
    // last one that changes this size's kernel runs that one (k_batch_tw.hip)
    constexpr int TOP = tw64_levels(LOG2N, batch_log2p(LOG2N));
    const int level = v.tw64 < TOP ? v.tw64 : TOP;
    // the halves layout exists for the sizes whose LDS limits the workgroups per CU (kernels.hpp)
    const int layout = !USES_LDS ? kPadded : (v.layout == kHalves && LOG2N < kHalvesMinLog2) ? kPadded : v.layout;
    if (level > 0) return launch_batch_tw64(LOG2N, level, layout == kHalves, v.nt, inverse, a, grid, stream);
    if constexpr (LOG2N >= kHalvesMinLog2) {
        if (layout == kHalves) {
            if (v.nt) launch_variant<LOG2N, true, kHalves>(inverse, a, grid, stream);
            else launch_variant<LOG2N, false, kHalves>(inverse, a, grid, stream);
            return hipGetLastError();
        }
    }
    const bool swz = layout == kSwizzled;
    if (v.nt) {
        if (swz) launch_variant<LOG2N, true, kSwizzled>(inverse, a, grid, stream);
        else launch_variant<LOG2N, true, kPadded>(inverse, a, grid, stream);
    } else {
        if (swz) launch_variant<LOG2N, false, kSwizzled>(inverse, a, grid, stream);
        else launch_variant<LOG2N, false, kPadded>(inverse, a, grid, stream);
    }
    return hipGetLastError();
}

hipError_t launch_batch(int log2n, bool inverse, const BatchArgs &a, BatchVariant v, hipStream_t stream) {
    switch (log2n) {
    case 0: return launch_batch_n<0>(inverse, a, v, stream);
    case 1: return launch_batch_n<1>(inverse, a, v, stream);
    case 2: return launch_batch_n<2>(inverse, a, v, stream);
    case 3: return launch_batch_n<3>(inverse, a, v, stream);
    case 4: return launch_batch_n<4>(inverse, a, v, stream);
    case 5: return launch_batch_n<5>(inverse, a, v, stream);
    case 6: return launch_batch_n<6>(inverse, a, v, stream);
    case 7: return launch_batch_n<7>(inverse, a, v, stream);
    case 8: return launch_batch_n<8>(inverse, a, v, stream);
    case 9: return launch_batch_n<9>(inverse, a, v, stream);
    case 10: return launch_batch_n<10>(inverse, a, v, stream);
    case 11: return launch_batch_n<11>(inverse, a, v, stream);
    case 12: return launch_batch_n<12>(inverse, a, v, stream);
    case 13: return launch_batch_n<13>(inverse, a, v, stream);
    case 14: return launch_batch_n<14>(inverse, a, v, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace onefft
