// The single-pass batch kernel (fft_batch_kernel) and its launcher, shared by k_batch.hip (the
// variants without FP64-base twiddles, and the size dispatch) and k_batch_tw.hip (the variants
// with them), which compile in parallel.
#pragma once

#include "fft_bodies.hpp"

namespace onefft {

template <int LOG2N, bool INV, bool NT, int LAYOUT, int TW64>
__global__ __launch_bounds__(batch_wg(LOG2N)) void fft_batch_kernel(const cf *__restrict__ in, cf *__restrict__ out,
                                                                      const cf *__restrict__ tw,
                                                                      const cd *__restrict__ tw64, uint64_t in_dist,
                                                                      uint64_t out_dist, uint64_t batch) {
    constexpr int N = 1 << LOG2N;
    constexpr uint32_t T = N >> batch_log2p(LOG2N);  // threads per transform
    constexpr int PER = batch_per_wg(LOG2N);         // transforms per workgroup
    constexpr bool USES_LDS = stage_count(LOG2N, batch_log2p(LOG2N)) > 1;
    constexpr uint32_t SPAN = lds_extent<LAYOUT>(N);
    static_assert(batch_wg(LOG2N) % T == 0, "workgroup must hold whole transforms");
    __shared__ cf lds[USES_LDS ? PER * SPAN : 1];

    const uint32_t t = threadIdx.x;
    if constexpr (PER == 1) {
        // one transform per workgroup, and the grid has exactly one workgroup per transform: its
        // base is uniform, so it goes into buffer descriptors
        const uint64_t which = blockIdx.x;
        const BufIO<NT> io(in + which * in_dist, out + which * out_dist);
        single_body<LOG2N, INV, LAYOUT, TW64>(io, true, tw, lds, t, tw64);
    } else {
        // the last workgroup's lanes past the batch load the batch's last transform again and
        // store nothing
        const uint32_t c = t / T;
        const uint64_t which = (uint64_t)blockIdx.x * PER + c;
        const uint64_t src = which < batch ? which : batch - 1;
        const PtrIO<NT> io{in + src * in_dist, out + which * out_dist};
        single_body<LOG2N, INV, LAYOUT, TW64>(io, which < batch, tw, lds + (USES_LDS ? c * SPAN : 0), t % T, tw64);
    }
}

template <int LOG2N, bool NT, int LAYOUT, int TW64 = 0>
inline void launch_variant(bool inverse, const BatchArgs &a, dim3 grid, hipStream_t stream) {
    if (inverse)
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, true, NT, LAYOUT, TW64>), grid, dim3(batch_wg(LOG2N)), 0, stream,
                           a.in, a.out, a.tw, a.tw64, a.in_dist, a.out_dist, a.batch);
    else
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, false, NT, LAYOUT, TW64>), grid, dim3(batch_wg(LOG2N)), 0, stream,
                           a.in, a.out, a.tw, a.tw64, a.in_dist, a.out_dist, a.batch);
}

// Launches the FP64-base variant `level` (1..3, at most the size's tw64_levels) of the padded
// kernel for 2^log2n points, log2n <= kMaxSinglePassLog2 (k_batch_tw.hip).
hipError_t launch_batch_tw64(int log2n, int level, bool nt, bool inverse, const BatchArgs &a, dim3 grid,
                             hipStream_t stream);

}  // namespace onefft
