// One pass of a multi-pass (four-step / Bailey) transform of n > 2^14 points: the kernel and its
// launch dispatch, compiled once per LDS layout (k_pass.hip: padded, k_pass_halves.hip: halves).
//
// n = N_1 N_2 ... N_p, input index j = j_1 + N_1 j_2 + N_1 N_2 j_3 + ..., output index
// k = sum_f M_f k_f with M_f = N_{f+1} ... N_p (M_p = 1, M_0 = n). The passes run for
// f = p, p-1, ..., 1:
//   f = p (first, out of place): for every input column J = (j_1 .. j_{p-1}) an N_p-point DFT
//       over j_p (input stride n / N_p), stored as a contiguous row at sum_{i<p} M_i j_i;
//   f < p (in place on the output): for every (O, low), low < M_f, the N_f-point DFT over j_f
//       of the values at M_{f-1} O + M_f j_f + low, each first multiplied by
//       w_{M_{f-1}}^(j_f low), stored back to the same positions with j_f replaced by k_f.
// Every pass works on C adjacent columns per workgroup, so each global access is a run of C
// (or, for the first pass's stores, N_p) consecutive complex values. Only the first pass reads
// `in`; no scratch buffer beyond `out` is needed, like the reference, which also uses its output
// as the only work space (src/onefft.c:274-293).
//
// The pass twiddles w_n^x span all n exponents. They are formed in FP64 from two small FP64
// tables (w_n^x = w_n^(x_hi 2^b) w_n^(x_lo)) and a short FP64 recurrence per thread, then rounded
// once to FP32, without reading an n-entry table from HBM. The in-place passes form their
// recurrence's bases per tile with a small column table of the plan (PassArgs::tw_col; the step
// bases too in passes whose lookups would sweep more of the low table than the L1 holds,
// instantiations with COL_STEPS). The values equal the reference's
// FP64-computed, once-rounded twiddles (src/onefft.c:112-141) except where a component that the
// reference holds as exactly 0 (a quarter turn) comes out as a residue of a few 1e-16, or where
// the exact value lies within the FP64 formation error of an FP32 rounding boundary (then one
// unit in the last place).
#pragma once

#include "fft_bodies.hpp"

namespace onefft {

template <int LOG2NF, int C, bool FIRST, bool SWAP_IN, bool SWAP_OUT, int LAYOUT, bool COL_STEPS>
__global__ __launch_bounds__(pass_wg(LOG2NF, C)) void fft_pass_kernel(PassArgs a) {
    constexpr bool USES_LDS = stage_count(LOG2NF, pass_log2p(LOG2NF)) > 1;
    // C columns of pass_span values: complex values, or floats with the halves layout
    constexpr uint32_t SPAN = pass_span(LOG2NF, C, LAYOUT);
    constexpr uint32_t VALUES = USES_LDS ? (LAYOUT == kHalves ? (C * SPAN + 1) / 2 : C * SPAN) : 1;
    __shared__ cf lds[VALUES];
    // tiles in XCD-contiguous order over the whole grid (tile-major within a transform)
    const uint32_t tiles = gridDim.x;
    const uint32_t w = xcd_strips(blockIdx.x + blockIdx.y * tiles, tiles * gridDim.y, a.strips);
    const uint32_t tile = w % tiles, which = w / tiles;
    const uint64_t in_at = a.offsets ? a.offsets[which] : (uint64_t)which * a.in_dist;
    const uint64_t out_at = a.offsets ? a.offsets[which] : (uint64_t)which * a.out_dist;
    pass_body<LOG2NF, C, FIRST, SWAP_IN, SWAP_OUT, LAYOUT, COL_STEPS>(a, a.in + in_at, a.out + out_at, tile, lds, threadIdx.x);
}

template <int LOG2NF, int C, int LAYOUT>
inline hipError_t launch_pass_nc(bool first, bool swap_in, bool swap_out, const PassArgs &a, uint64_t batch,
                                 hipStream_t stream) {
    if constexpr (!pass_cols_ok(LOG2NF, C, LAYOUT)) {
        return hipErrorInvalidValue;
    } else {
        if (a.columns % C || batch == 0 || batch > 65535) return hipErrorInvalidValue;
        if (!first && !a.tw_col) return hipErrorInvalidValue;
        dim3 grid(a.columns / C, (uint32_t)batch), block(pass_wg(LOG2NF, C));
        // in-place passes that take their step bases from the column table (PassArgs::col_steps)
        // run their own instantiation: a run-time branch between the two formations cost
        // registers and time in both
        if (first) {
            if (swap_in)
                hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, C, true, true, false, LAYOUT, false>), grid, block, 0, stream, a);
            else
                hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, C, true, false, false, LAYOUT, false>), grid, block, 0, stream, a);
        } else if (a.col_steps) {
            if (swap_out)
                hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, C, false, false, true, LAYOUT, true>), grid, block, 0, stream, a);
            else
                hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, C, false, false, false, LAYOUT, true>), grid, block, 0, stream, a);
        } else {
            if (swap_out)
                hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, C, false, false, true, LAYOUT, false>), grid, block, 0, stream, a);
            else
                hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, C, false, false, false, LAYOUT, false>), grid, block, 0, stream, a);
        }
        return hipGetLastError();
    }
}

template <int LOG2NF, int LAYOUT>
inline hipError_t launch_pass_n(int cols, bool first, bool swap_in, bool swap_out, const PassArgs &a, uint64_t batch,
                                hipStream_t stream) {
    switch (cols) {
    case 2: return launch_pass_nc<LOG2NF, 2, LAYOUT>(first, swap_in, swap_out, a, batch, stream);
    case 4: return launch_pass_nc<LOG2NF, 4, LAYOUT>(first, swap_in, swap_out, a, batch, stream);
    case 8: return launch_pass_nc<LOG2NF, 8, LAYOUT>(first, swap_in, swap_out, a, batch, stream);
    case 16: return launch_pass_nc<LOG2NF, 16, LAYOUT>(first, swap_in, swap_out, a, batch, stream);
    case 32: return launch_pass_nc<LOG2NF, 32, LAYOUT>(first, swap_in, swap_out, a, batch, stream);
    case 64: return launch_pass_nc<LOG2NF, 64, LAYOUT>(first, swap_in, swap_out, a, batch, stream);
    default: return hipErrorInvalidValue;
    }
}

template <int LAYOUT>
inline hipError_t launch_pass_layout(int log2nf, int cols, bool first, bool swap_in, bool swap_out, const PassArgs &a,
                                     uint64_t batch, hipStream_t stream) {
    // a first pass never stores the final result (there are at least two passes)
    if (first && swap_out) return hipErrorInvalidValue;
    if (!first && swap_in) return hipErrorInvalidValue;
    switch (log2nf) {
    case 4: return launch_pass_n<4, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 5: return launch_pass_n<5, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 6: return launch_pass_n<6, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 7: return launch_pass_n<7, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 8: return launch_pass_n<8, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 9: return launch_pass_n<9, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 10: return launch_pass_n<10, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 11: return launch_pass_n<11, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    case 12: return launch_pass_n<12, LAYOUT>(cols, first, swap_in, swap_out, a, batch, stream);
    default: return hipErrorInvalidValue;
    }
}

// the halves-layout pass kernels (k_pass_halves.hip)
hipError_t launch_pass_halves(int log2nf, int cols, bool first, bool swap_in, bool swap_out, const PassArgs &a,
                              uint64_t batch, hipStream_t stream);

}  // namespace onefft
