// The single-pass batch kernels with FP64-base twiddles (fft_common.hpp: stage_uses_tw64), padded
// LDS or, from 2^kHalvesMinLog2 points, the halves layout, for 2^5..2^10 points; 2^11..2^14 are in
// k_batch_tw_large.hip. Both compile in parallel with the other variants (k_batch.hip).
#include "k_batch.hpp"

namespace onefft {

hipError_t launch_batch_tw64(int log2n, int level, bool halves, bool nt, bool inverse, const BatchArgs &a, dim3 grid,
                             hipStream_t stream) {
    switch (log2n) {
    case 5: return launch_tw64_n<5>(level, halves, nt, inverse, a, grid, stream);
    case 6: return launch_tw64_n<6>(level, halves, nt, inverse, a, grid, stream);
    case 7: return launch_tw64_n<7>(level, halves, nt, inverse, a, grid, stream);
    case 8: return launch_tw64_n<8>(level, halves, nt, inverse, a, grid, stream);
    case 9: return launch_tw64_n<9>(level, halves, nt, inverse, a, grid, stream);
    case 10: return launch_tw64_n<10>(level, halves, nt, inverse, a, grid, stream);
    default: return launch_batch_tw64_large(log2n, level, halves, nt, inverse, a, grid, stream);
    }
}

}  // namespace onefft
