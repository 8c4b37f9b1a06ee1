// The single-pass batch kernels with FP64-base twiddles for 2^11..2^14 points (the largest, in a
// translation unit of their own so the build compiles them in parallel with k_batch_tw.hip).
#include "k_batch.hpp"

namespace onefft {

hipError_t launch_batch_tw64_large(int log2n, int level, bool halves, bool nt, bool inverse, const BatchArgs &a,
                                   dim3 grid, hipStream_t stream) {
    switch (log2n) {
    case 11: return launch_tw64_n<11>(level, halves, nt, inverse, a, grid, stream);
    case 12: return launch_tw64_n<12>(level, halves, nt, inverse, a, grid, stream);
    case 13: return launch_tw64_n<13>(level, halves, nt, inverse, a, grid, stream);
    case 14: return launch_tw64_n<14>(level, halves, nt, inverse, a, grid, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace onefft
