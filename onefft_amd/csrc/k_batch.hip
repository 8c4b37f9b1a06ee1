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
#include "fft_bodies.hpp"

namespace onefft {

template <int LOG2N, bool INV, bool NT, int LAYOUT, bool TW64>
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

template <int LOG2N, bool NT, int LAYOUT, bool TW64 = false>
static void launch_variant(bool inverse, const BatchArgs &a, dim3 grid, hipStream_t stream) {
    if (inverse)
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, true, NT, LAYOUT, TW64>), grid, dim3(batch_wg(LOG2N)), 0, stream,
                           a.in, a.out, a.tw, a.tw64, a.in_dist, a.out_dist, a.batch);
    else
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, false, NT, LAYOUT, TW64>), grid, dim3(batch_wg(LOG2N)), 0, stream,
                           a.in, a.out, a.tw, a.tw64, a.in_dist, a.out_dist, a.batch);
}

template <int LOG2N>
static hipError_t launch_batch_n(bool inverse, const BatchArgs &a, BatchVariant v, hipStream_t stream) {
    constexpr int PER = batch_per_wg(LOG2N);
    const uint64_t blocks = (a.batch + PER - 1) / PER;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    const dim3 grid((uint32_t)blocks);
    constexpr bool USES_LDS = stage_count(LOG2N, batch_log2p(LOG2N)) > 1;
    constexpr bool HAS_TW64 = stage_tw64_size(LOG2N, batch_log2p(LOG2N)) > 0;
    if constexpr (HAS_TW64) {
        if (v.tw64) {  // padded layout only
            if (v.nt) launch_variant<LOG2N, true, kPadded, true>(inverse, a, grid, stream);
            else launch_variant<LOG2N, false, kPadded, true>(inverse, a, grid, stream);
            return hipGetLastError();
        }
    }
    const bool swz = USES_LDS && v.layout == kSwizzled;
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
