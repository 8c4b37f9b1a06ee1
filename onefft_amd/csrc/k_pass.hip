// The multi-pass kernels with padded LDS and the pass launch dispatch (kernel and index scheme:
// k_pass.hpp); the halves-layout kernels compile in parallel in k_pass_halves.hip.
#include "k_pass.hpp"

namespace onefft {

hipError_t launch_pass(int log2nf, int cols, int layout, bool first, bool swap_in, bool swap_out, const PassArgs &a,
                       uint64_t batch, hipStream_t stream) {
    if (layout == kHalves) return launch_pass_halves(log2nf, cols, first, swap_in, swap_out, a, batch, stream);
    return launch_pass_layout<kPadded>(log2nf, cols, first, swap_in, swap_out, a, batch, stream);
}

}  // namespace onefft
