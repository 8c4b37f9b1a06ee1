// Measurement only (bench.py): a plain device copy with the global access pattern of the
// 4096-point single-pass kernel and none of its arithmetic. Each 256-thread workgroup moves one
// 4096-value (32 KiB) block: thread t loads values t + 256 r, r = 0..15, as 8-byte buffer loads,
// then stores them back at the same offsets in the output, with the same cache policy as the FFT
// kernel (nontemporal or not). Its rate is the bandwidth that access pattern can reach on this
// chip, the practical ceiling for the FFT kernel; the HBM peak (8 TB/s) is the theoretical one.
//
// Built into its own library (libonefft_stream.so), not into libonefft_amd.so.
#include "fft_common.hpp"

namespace onefft {

constexpr uint32_t kCopyBlock = 4096;  // values per workgroup

template <bool NT>
__global__ __launch_bounds__(256) void stream_copy_kernel(const cf *in, cf *out) {
    const uint64_t base = (uint64_t)blockIdx.x * kCopyBlock;
    const BufIO<NT> io(in + base, out + base);
    cf v[16];
    static_for<0, 16>([&](auto r) { v[r] = io.load(threadIdx.x, decltype(r)::value * 256u); });
    static_for<0, 16>([&](auto r) { io.store(threadIdx.x, decltype(r)::value * 256u, v[r]); });
}

}  // namespace onefft

// Copies `bytes` from in to out (device pointers, 8-byte aligned, bytes a multiple of 32 KiB).
// Returns 0, or -1 for an invalid size or a failed launch.
extern "C" __attribute__((visibility("default"))) int onefft_stream_copy(const void *in, void *out, size_t bytes,
                                                                        int nontemporal, hipStream_t stream) {
    using namespace onefft;
    const size_t block_bytes = kCopyBlock * sizeof(cf);
    if (bytes == 0 || bytes % block_bytes || bytes / block_bytes > 0x7fffffffu) return -1;
    if (((uintptr_t)in | (uintptr_t)out) % 8) return -1;
    const dim3 grid((uint32_t)(bytes / block_bytes));
    if (nontemporal)
        hipLaunchKernelGGL(stream_copy_kernel<true>, grid, dim3(256), 0, stream, (const cf *)in, (cf *)out);
    else
        hipLaunchKernelGGL(stream_copy_kernel<false>, grid, dim3(256), 0, stream, (const cf *)in, (cf *)out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
