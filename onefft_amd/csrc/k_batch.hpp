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
                                                                      uint64_t out_dist, uint64_t batch,
                                                                      uint32_t *done, uint32_t seq) {
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
    if (done) {
        // every thread's stores reach the system before the barrier, the flag after it
        __threadfence_system();
        __syncthreads();
        if (t == 0) __hip_atomic_store(done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

template <int LOG2N, bool NT, int LAYOUT, int TW64 = 0>
inline void launch_variant(bool inverse, const BatchArgs &a, dim3 grid, hipStream_t stream) {
    if (inverse)
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, true, NT, LAYOUT, TW64>), grid, dim3(batch_wg(LOG2N)), 0, stream,
                           a.in, a.out, a.tw, a.tw64, a.in_dist, a.out_dist, a.batch, a.done, a.seq);
    else
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, false, NT, LAYOUT, TW64>), grid, dim3(batch_wg(LOG2N)), 0, stream,
                           a.in, a.out, a.tw, a.tw64, a.in_dist, a.out_dist, a.batch, a.done, a.seq);
}

// Contiguous batches of transforms of n <= 32 points (in_dist = out_dist = n; the reference's small
// kernels, src/onefft_avx512.inc n32/n64/n128 and its 1, 2 and 4-point closed forms,
// src/onefft.c:158-181). Each 256-thread workgroup moves one block of 4096 values with coalesced
// 8-byte accesses (thread t: values t + 256 r), exchanges them through LDS so that each thread
// holds 16 / n whole transforms, runs the register DFTs, and exchanges back before the coalesced
// stores. For n = 32 two threads share a transform: thread j of it takes the 16-point DFT of
// x[2k + j] (the odd half then times w_32^m), and after one more exchange each combines 8 of the
// output pairs X[m], X[m + 16] = E[m] +- O[m], at the LDS positions it read them from. (The
// general kernel gives each thread 16 points of one transform: a wavefront's accesses then touch
// lines 8 n bytes apart.) Only the last workgroup's block can be partial; its lanes skip the
// values past the end.
constexpr uint32_t kSmallBlock = 4096;

// The value the neighbouring lane (lane ^ 1) holds: a DPP quad_perm [1, 0, 3, 2] move, a VALU
// operation with no LDS traffic and no barrier.
__device__ __forceinline__ float from_pair_lane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, false));
}

// XLANE (n = 32 only): the two threads of a transform exchange their halves by DPP lane swaps
// instead of through LDS.
template <int LOG2N, bool INV, bool NT, bool XLANE = false>
__global__ __launch_bounds__(256) void fft_small_kernel(const cf *__restrict__ in, cf *__restrict__ out,
                                                         uint64_t total, uint32_t *done, uint32_t seq) {
    constexpr int N = 1 << LOG2N;
    static_assert(LOG2N <= 5, "up to 32 points: one register DFT per 16 / n values, or two threads per 32");
    __shared__ cf lds[LOG2N > 0 ? lds_span(kSmallBlock) : 1];
    const uint32_t t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kSmallBlock;
    const uint64_t left = total - base;
    const uint32_t count = left < kSmallBlock ? (uint32_t)left : kSmallBlock;
    const BufIO<NT> io(in + base, out + base);
    cf v[16];
    if (count == kSmallBlock) {
        static_for<0, 16>([&](auto r) { v[r] = io.load(t, decltype(r)::value * 256u); });
    } else {
        static_for<0, 16>([&](auto r) {
            constexpr uint32_t c = decltype(r)::value * 256u;
            v[r] = t + c < count ? io.load(t, c) : mk(0.f, 0.f);
        });
    }
    if constexpr (LOG2N == 5) {
        static_for<0, 16>([&](auto r) { lds[lds_pad(t + decltype(r)::value * 256u)] = v[r]; });
        __syncthreads();
        // thread t: half j of transform c, whose values start at 32 c
        const uint32_t c32 = (t >> 1) * 32u, j = t & 1u;
        static_for<0, 16>([&](auto k) { v[k] = maybe_conj<INV>(lds[lds_pad(c32 + 2u * decltype(k)::value + j)]); });
        dft<16>(v);
        if (j) static_for<1, 16>([&](auto m) { v[m] = twc<32, decltype(m)::value>(v[m]); });
        if constexpr (XLANE) {
            // thread 0 holds E[0..15], thread 1 O[0..15]: each sends the half the other combines
            // (E[8..15] to thread 1, O[0..7] to thread 0) and forms outputs m, m + 16 for
            // m = 8 j .. 8 j + 7 from E[m] +- O[m]
            __syncthreads();  // every thread has read its inputs
            static_for<0, 8>([&](auto i) {
                constexpr int k = decltype(i)::value;
                const cf send = j ? v[k] : v[k + 8], own = j ? v[k + 8] : v[k];
                const cf got = mk(from_pair_lane(send.r), from_pair_lane(send.i));
                const cf e = j ? got : own, o = j ? own : got;
                const uint32_t m = 8u * j + k;
                lds[lds_pad(c32 + m)] = maybe_conj<INV>(add(e, o));
                lds[lds_pad(c32 + 16u + m)] = maybe_conj<INV>(sub(e, o));
            });
            __syncthreads();
            static_for<0, 16>([&](auto r) { v[r] = lds[lds_pad(t + decltype(r)::value * 256u)]; });
        } else {
        __syncthreads();  // every thread has read its inputs
        static_for<0, 16>([&](auto m) { lds[lds_pad(c32 + 16u * j + decltype(m)::value)] = v[m]; });
        __syncthreads();
        // outputs m and m + 16 for m = 8 j .. 8 j + 7, from E[m] (position m) and O[m] (m + 16)
        static_for<0, 8>([&](auto i) {
            const uint32_t m = 8u * j + decltype(i)::value;
            const cf e = lds[lds_pad(c32 + m)], o = lds[lds_pad(c32 + 16u + m)];
            lds[lds_pad(c32 + m)] = maybe_conj<INV>(add(e, o));
            lds[lds_pad(c32 + 16u + m)] = maybe_conj<INV>(sub(e, o));
        });
        __syncthreads();
        static_for<0, 16>([&](auto r) { v[r] = lds[lds_pad(t + decltype(r)::value * 256u)]; });
        }
    } else if constexpr (LOG2N > 0) {
        static_for<0, 16>([&](auto r) { lds[lds_pad(t + decltype(r)::value * 256u)] = v[r]; });
        __syncthreads();
        // thread t now works on values 16 t .. 16 t + 15 alone: 16 / n whole transforms
        static_for<0, 16>([&](auto k) { v[k] = maybe_conj<INV>(lds[lds_pad(16 * t + decltype(k)::value)]); });
        static_for<0, 16 / N>([&](auto q) { dft<N>(&v[decltype(q)::value * N]); });
        static_for<0, 16>([&](auto k) { lds[lds_pad(16 * t + decltype(k)::value)] = maybe_conj<INV>(v[k]); });
        __syncthreads();
        static_for<0, 16>([&](auto r) { v[r] = lds[lds_pad(t + decltype(r)::value * 256u)]; });
    }
    if (count == kSmallBlock) {
        static_for<0, 16>([&](auto r) { io.store(t, decltype(r)::value * 256u, v[r]); });
    } else {
        static_for<0, 16>([&](auto r) {
            constexpr uint32_t c = decltype(r)::value * 256u;
            if (t + c < count) io.store(t, c, v[r]);
        });
    }
    if (done) {
        // as in fft_batch_kernel: every thread's stores reach the system before the flag
        __threadfence_system();
        __syncthreads();
        if (t == 0) __hip_atomic_store(done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Launches the FP64-base variant `level` (1..3, at most the size's tw64_levels) of the kernel for
// 2^LOG2N points, padded LDS or, with `halves` from 2^kHalvesMinLog2 points, the halves layout.
template <int LOG2N>
inline hipError_t launch_tw64_n(int level, bool halves, bool nt, bool inverse, const BatchArgs &a, dim3 grid,
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

// 2^11..2^14 points (k_batch_tw_large.hip), called by launch_batch_tw64.
hipError_t launch_batch_tw64_large(int log2n, int level, bool halves, bool nt, bool inverse, const BatchArgs &a,
                                   dim3 grid, hipStream_t stream);

// Launches the FP64-base variant `level` (1..3, at most the size's tw64_levels) of the padded (or,
// with `halves` and log2n >= kHalvesMinLog2, the halves-layout) kernel for 2^log2n points,
// log2n <= kMaxSinglePassLog2 (k_batch_tw.hip).
hipError_t launch_batch_tw64(int log2n, int level, bool halves, bool nt, bool inverse, const BatchArgs &a, dim3 grid,
                             hipStream_t stream);

}  // namespace onefft
