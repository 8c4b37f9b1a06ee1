// Single-pass batched transforms of up to 2^14 points: one LDS-resident workgroup pass per
// transform (several transforms per workgroup below 4096 points).
//
// This replaces the reference's per-size kernel chain for sizes that fit on chip (small kernels
// src/onefft_avx512.inc n32/n64/n128, first/middle/last passes driven by onefft_execute_passes,
// src/onefft.c:274-293): every transform is read from HBM once and written once. Each thread
// holds 16 points in registers; radix-16 (or a smaller first radix) butterflies run in VGPRs and
// the stages exchange data through padded LDS (Stockham autosort, so the output lands in natural
// order without a bit-reversal pass).
#include "fft_common.hpp"
#include "kernels.hpp"

namespace onefft {

template <int LOG2N, bool INV>
__global__ __launch_bounds__(batch_wg(LOG2N)) void fft_batch_kernel(const cf *__restrict__ in, cf *__restrict__ out,
                                                                      const cf *__restrict__ tw, uint64_t in_dist,
                                                                      uint64_t out_dist, uint64_t batch) {
    constexpr int LOG2P = batch_log2p(LOG2N);
    constexpr int N = 1 << LOG2N;
    constexpr int P = 1 << LOG2P;
    constexpr int T = N / P;  // threads per transform
    constexpr int WG = batch_wg(LOG2N);
    constexpr int PER = WG / T;  // transforms per workgroup
    constexpr int S = stage_count(LOG2N, LOG2P);
    constexpr uint32_t SPAN = lds_span(N);
    static_assert(WG % T == 0, "workgroup must hold whole transforms");
    __shared__ cf lds[S > 1 ? PER * SPAN : 1];

    const uint32_t t = threadIdx.x;
    const uint32_t j = t % T;
    const uint32_t c = t / T;
    const uint64_t which = (uint64_t)blockIdx.x * PER + c;
    const bool live = which < batch;
    const cf *src = in + which * in_dist;
    cf *dst = out + which * out_dist;
    cf *L = lds + (S > 1 ? c * SPAN : 0);
    cf v[P];

    static_for<0, S>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        constexpr int LR = stage_log2r(LOG2N, LOG2P, s);
        constexpr int R = 1 << LR;
        constexpr int LNS = stage_log2ns(LOG2N, LOG2P, s);
        constexpr uint32_t NS = 1u << LNS;
        constexpr int BF = P / R;  // butterflies per thread
        constexpr uint32_t STRIDE = N / R;
        constexpr uint32_t TWO = stage_tw_offset(LOG2N, LOG2P, s);

        // inputs of this stage
        if constexpr (s == 0) {
            if (live) {
                static_for<0, BF>([&](auto q) {
                    static_for<0, R>([&](auto r) {
                        constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                        v[qq * R + rr] = maybe_swap<INV>(src[j + T * qq + rr * STRIDE]);
                    });
                });
            } else {
                static_for<0, P>([&](auto i) { v[i] = mk(0.f, 0.f); });
            }
        } else {
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R + rr] = L[lds_pad(j + T * qq + rr * STRIDE)];
                });
            });
        }

        // twiddles and butterflies
        static_for<0, BF>([&](auto q) {
            constexpr int qq = decltype(q)::value;
            if constexpr (s > 0) {
                const uint32_t k = (j + T * qq) & (NS - 1);
                static_for<1, R>([&](auto r) {
                    constexpr int rr = decltype(r)::value;
                    v[qq * R + rr] = mul(v[qq * R + rr], tw[TWO + (rr - 1) * NS + k]);
                });
            }
            dft<R>(&v[qq * R]);
        });

        // outputs of this stage
        if constexpr (s == S - 1) {
            // NS * R == N here, so output m of butterfly b goes to b + m N / R
            if (live) {
                static_for<0, BF>([&](auto q) {
                    static_for<0, R>([&](auto m) {
                        constexpr int qq = decltype(q)::value, mm = decltype(m)::value;
                        dst[j + T * qq + mm * STRIDE] = maybe_swap<INV>(v[qq * R + mm]);
                    });
                });
            }
        } else {
            if constexpr (s > 0) __syncthreads();  // every read of this stage is done
            static_for<0, BF>([&](auto q) {
                constexpr int qq = decltype(q)::value;
                const uint32_t b = j + T * qq;
                const uint32_t base = ((b >> LNS) << (LNS + LR)) + (b & (NS - 1));
                static_for<0, R>([&](auto m) {
                    constexpr int mm = decltype(m)::value;
                    L[lds_pad(base + mm * NS)] = v[qq * R + mm];
                });
            });
            __syncthreads();
        }
    });
}

template <int LOG2N>
static hipError_t launch_batch_n(bool inverse, const BatchArgs &a, hipStream_t stream) {
    constexpr int PER = batch_per_wg(LOG2N);
    const uint64_t blocks = (a.batch + PER - 1) / PER;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    dim3 grid((uint32_t)blocks), block(batch_wg(LOG2N));
    if (inverse)
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, true>), grid, block, 0, stream, a.in, a.out, a.tw, a.in_dist,
                           a.out_dist, a.batch);
    else
        hipLaunchKernelGGL((fft_batch_kernel<LOG2N, false>), grid, block, 0, stream, a.in, a.out, a.tw, a.in_dist,
                           a.out_dist, a.batch);
    return hipGetLastError();
}

hipError_t launch_batch(int log2n, bool inverse, const BatchArgs &a, hipStream_t stream) {
    switch (log2n) {
    case 0: return launch_batch_n<0>(inverse, a, stream);
    case 1: return launch_batch_n<1>(inverse, a, stream);
    case 2: return launch_batch_n<2>(inverse, a, stream);
    case 3: return launch_batch_n<3>(inverse, a, stream);
    case 4: return launch_batch_n<4>(inverse, a, stream);
    case 5: return launch_batch_n<5>(inverse, a, stream);
    case 6: return launch_batch_n<6>(inverse, a, stream);
    case 7: return launch_batch_n<7>(inverse, a, stream);
    case 8: return launch_batch_n<8>(inverse, a, stream);
    case 9: return launch_batch_n<9>(inverse, a, stream);
    case 10: return launch_batch_n<10>(inverse, a, stream);
    case 11: return launch_batch_n<11>(inverse, a, stream);
    case 12: return launch_batch_n<12>(inverse, a, stream);
    case 13: return launch_batch_n<13>(inverse, a, stream);
    case 14: return launch_batch_n<14>(inverse, a, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace onefft
