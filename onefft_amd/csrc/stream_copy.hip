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

// ---- access patterns of the 2^27-point passes, without their arithmetic -----------------------
//
// Measurement only (bench.py --access-patterns): copies over two buffers of 2^27 complex values
// (1 GiB each) with the global access pattern of one pass of a 2^27-point transform whose input
// pass has 2^10-point columns (split 2^8 * 2^9 * 2^10, 16 columns per workgroup), each pattern in
// its own kernel:
//   0  the input pass's reads (16 adjacent columns x 1024 rows, rows 2^17 values apart, 1024
//      threads x 16 values), each tile stored as one contiguous block;
//   1  the same reads, each column stored as its output row the way a thread's direct stores do
//      (a wave writes 16 rows, 4 consecutive values each; rows 2^19 values apart);
//   2  the same reads and rows, stored through LDS (real parts, then imaginary parts) so that a
//      wave writes 64 consecutive values of one row;
//   3  the last in-place pass (32 adjacent columns x 256 rows, rows 2^19 values apart, 512 threads
//      x 16 values), read and written back in place;
//   4  the same tiles with rows 2^11 values (16 KiB) apart, the middle pass's stride;
//   5  the same tiles with rows 2^19 + 64 values apart (4 MiB + 512 bytes: as many rows and pages
//      per tile as pattern 3, the stride no power of two), over the 15874 tiles whose rows fit the
//      buffer.
namespace onefft {

constexpr uint32_t kPatRows = 1024, kPatCols = 1u << 17;  // the input view: 1024 rows of 2^17

// output row of input column J for the split 2^8 * 2^9 * 2^10: j1 = J mod 2^8 weighs 2^19 values,
// j2 = J / 2^8 weighs 2^10 (the pass kernels' dig_log2 / dig_shift for this split)
__device__ __forceinline__ uint64_t pattern_row(uint32_t J) { return ((uint64_t)(J & 255) << 19) + ((uint64_t)(J >> 8) << 10); }

template <int P>
__global__ __launch_bounds__(1024) void column_pattern_kernel(const cf *__restrict__ in, cf *__restrict__ out) {
    __shared__ float lds[P == 2 ? 16 * kPatRows : 1];
    const uint32_t tile = blockIdx.x, t = threadIdx.x;
    const uint32_t c = t & 15, j = t >> 4;  // column within the tile, first row
    cf v[16];
    static_for<0, 16>([&](auto k) {
        v[k] = in[(uint64_t)(j + 64 * decltype(k)::value) * kPatCols + tile * 16 + c];
    });
    if constexpr (P == 0) {
        static_for<0, 16>([&](auto k) { out[(uint64_t)tile * 16 * kPatRows + t + 1024 * decltype(k)::value] = v[k]; });
    } else if constexpr (P == 1) {
        const uint64_t row = pattern_row(tile * 16 + c);
        static_for<0, 16>([&](auto k) { out[row + j + 64 * decltype(k)::value] = v[k]; });
    } else {
        // column c's values at natural positions j + 64 k; value k of thread t is then column k,
        // position t
        float re[16];
        static_for<0, 16>([&](auto k) { lds[c * kPatRows + j + 64 * decltype(k)::value] = v[k].r; });
        __syncthreads();
        static_for<0, 16>([&](auto k) { re[k] = lds[decltype(k)::value * kPatRows + t]; });
        __syncthreads();
        static_for<0, 16>([&](auto k) { lds[c * kPatRows + j + 64 * decltype(k)::value] = v[k].i; });
        __syncthreads();
        static_for<0, 16>([&](auto k) {
            constexpr int kk = decltype(k)::value;
            out[pattern_row(tile * 16 + kk) + t] = mk(re[kk], lds[kk * kPatRows + t]);
        });
    }
}

template <uint64_t kStride>
__global__ __launch_bounds__(512) void inplace_pattern_kernel(cf *__restrict__ buf, float add) {
    const uint32_t tile = blockIdx.x, t = threadIdx.x;
    const uint64_t col = (uint64_t)tile * 32 + (t & 31);
    const uint32_t r0 = t >> 5;
    cf v[16];
    static_for<0, 16>([&](auto k) { v[k] = buf[(uint64_t)(r0 + 16 * decltype(k)::value) * kStride + col]; });
    static_for<0, 16>([&](auto k) {
        v[k].r += add;
        buf[(uint64_t)(r0 + 16 * decltype(k)::value) * kStride + col] = v[k];
    });
}

// pattern 4: tiles of 32 columns x 256 rows 2^11 values apart, the bands of 256 rows laid end to
// end (as a middle pass's tiles lie in a 2^27-point transform split 2^8 * 2^8 * 2^11)
__global__ __launch_bounds__(512) void inplace_pattern_band_kernel(cf *__restrict__ buf, float add) {
    constexpr uint64_t kStride = 1u << 11;
    const uint32_t tile = blockIdx.x, t = threadIdx.x;
    const uint64_t band = tile / 64, col = (uint64_t)(tile % 64) * 32 + (t & 31);
    const uint32_t r0 = t >> 5;
    cf *b = buf + band * 256 * kStride;
    cf v[16];
    static_for<0, 16>([&](auto k) { v[k] = b[(uint64_t)(r0 + 16 * decltype(k)::value) * kStride + col]; });
    static_for<0, 16>([&](auto k) {
        v[k].r += add;
        b[(uint64_t)(r0 + 16 * decltype(k)::value) * kStride + col] = v[k];
    });
}

}  // namespace onefft

// Runs access pattern `pattern` (0..5, above) once: `a` the input, `b` the output (patterns 3..5
// work in place on `b` and leave its values unchanged). Both hold 2^27 complex values, device
// memory, 8-byte aligned. Returns 0, or -1 for an unknown pattern or a failed launch.
extern "C" __attribute__((visibility("default"))) int onefft_access_pattern(int pattern, const void *a, void *b,
                                                                           hipStream_t stream) {
    using namespace onefft;
    if (((uintptr_t)a | (uintptr_t)b) % 8) return -1;
    const dim3 tiles(kPatCols / 16);
    switch (pattern) {
    case 0: hipLaunchKernelGGL(column_pattern_kernel<0>, tiles, dim3(1024), 0, stream, (const cf *)a, (cf *)b); break;
    case 1: hipLaunchKernelGGL(column_pattern_kernel<1>, tiles, dim3(1024), 0, stream, (const cf *)a, (cf *)b); break;
    case 2: hipLaunchKernelGGL(column_pattern_kernel<2>, tiles, dim3(1024), 0, stream, (const cf *)a, (cf *)b); break;
    case 3: hipLaunchKernelGGL(inplace_pattern_kernel<1u << 19>, dim3((1u << 19) / 32), dim3(512), 0, stream, (cf *)b, 0.0f); break;
    // 2^27 values: 2^27 / (256 rows x 32 columns) tiles; with rows 2^11 apart, tile k covers column
    // block k mod 64 of the 256-row band k / 64
    case 4: hipLaunchKernelGGL(inplace_pattern_band_kernel, dim3((1u << 27) / (256 * 32)), dim3(512), 0, stream, (cf *)b, 0.0f); break;
    // 255 (2^19 + 64) + 15874 x 32 <= 2^27
    case 5: hipLaunchKernelGGL(inplace_pattern_kernel<(1u << 19) + 64>, dim3(15874), dim3(512), 0, stream, (cf *)b, 0.0f); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
