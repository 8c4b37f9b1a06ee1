// The multi-pass kernels with the halves LDS layout (fft_common.hpp: kHalves; kernel and index
// scheme: k_pass.hpp), in their own translation unit so they compile in parallel with k_pass.hip.
#include "k_pass.hpp"

namespace onefft {

hipError_t launch_pass_halves(int log2nf, int cols, bool first, bool swap_in, bool swap_out, const PassArgs &a,
                              uint64_t batch, hipStream_t stream) {
    return launch_pass_layout<kHalves>(log2nf, cols, first, swap_in, swap_out, a, batch, stream);
}

}  // namespace onefft
