// The single-pass batch kernels with FP64-base twiddles (fft_common.hpp: stage_uses_tw64), padded
// LDS or, from 2^kHalvesMinLog2 points, the halves layout, in their own translation unit so they
// compile in parallel with the other variants (k_batch.hip).
#include "k_batch.hpp"

namespace onefft {

namespace {

template <int LOG2N>
hipError_t launch_tw64_n(int level, bool halves, bool nt, bool inverse, const BatchArgs &a, dim3 grid,
                         hipStream_t stream) {
    constexpr int TOP = tw64_levels(LOG2N, batch_log2p(LOG2N));
    if constexpr (TOP == 0) {
        return hipErrorInvalidValue;
    } else {
        auto run = [&](auto lv) {
            constexpr int L = decltype(lv)::value;
            if constexpr (LOG2N >= kHalvesMinLog2) {
                if (halves) {
                    if (nt) launch_variant<LOG2N, true, kHalves, L>(inverse, a, grid, stream);
                    else launch_variant<LOG2N, false, kHalves, L>(inverse, a, grid, stream);
                    return;
                }
            }
            if (nt) launch_variant<LOG2N, true, kPadded, L>(inverse, a, grid, stream);
            else launch_variant<LOG2N, false, kPadded, L>(inverse, a, grid, stream);
        };
        if (level < 1 || level > TOP) return hipErrorInvalidValue;
        if (level == 1) run(std::integral_constant<int, 1>{});
        if constexpr (TOP >= 2) {
            if (level == 2) run(std::integral_constant<int, 2>{});
        }
        if constexpr (TOP >= 3) {
            if (level == 3) run(std::integral_constant<int, 3>{});
        }
        return hipGetLastError();
    }
}

}  // namespace

hipError_t launch_batch_tw64(int log2n, int level, bool halves, bool nt, bool inverse, const BatchArgs &a, dim3 grid,
                             hipStream_t stream) {
    switch (log2n) {
    case 5: return launch_tw64_n<5>(level, halves, nt, inverse, a, grid, stream);
    case 6: return launch_tw64_n<6>(level, halves, nt, inverse, a, grid, stream);
    case 7: return launch_tw64_n<7>(level, halves, nt, inverse, a, grid, stream);
    case 8: return launch_tw64_n<8>(level, halves, nt, inverse, a, grid, stream);
    case 9: return launch_tw64_n<9>(level, halves, nt, inverse, a, grid, stream);
    case 10: return launch_tw64_n<10>(level, halves, nt, inverse, a, grid, stream);
    case 11: return launch_tw64_n<11>(level, halves, nt, inverse, a, grid, stream);
    case 12: return launch_tw64_n<12>(level, halves, nt, inverse, a, grid, stream);
    case 13: return launch_tw64_n<13>(level, halves, nt, inverse, a, grid, stream);
    case 14: return launch_tw64_n<14>(level, halves, nt, inverse, a, grid, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace onefft
