// One pass of a multi-pass (four-step / Bailey) transform of n > 2^14 points.
//
// n = N_1 N_2 ... N_p, input index j = j_1 + N_1 j_2 + N_1 N_2 j_3 + ..., output index
// k = sum_f M_f k_f with M_f = N_{f+1} ... N_p (M_p = 1, M_0 = n). The passes run for
// f = p, p-1, ..., 1:
//   f = p (first, out of place): for every input column J = (j_1 .. j_{p-1}) an N_p-point DFT
//       over j_p (input stride n / N_p), stored as a contiguous row at sum_{i<p} M_i j_i;
//   f < p (in place on the output): for every (O, low), low < M_f, the N_f-point DFT over j_f
//       of the values at M_{f-1} O + M_f j_f + low, each first multiplied by
//       w_{M_{f-1}}^(j_f low), stored back to the same positions with j_f replaced by k_f.
// Every pass works on C adjacent columns per workgroup, so each global access is a run of C
// (or, for the first pass's stores, N_p) consecutive complex values. Only the first pass reads
// `in`; no scratch buffer beyond `out` is needed, like the reference, which also uses its output
// as the only work space (src/onefft.c:274-293).
//
// The pass twiddles w_n^x span all n exponents. They are formed in FP64 from two small FP64
// tables (w_n^x = w_n^(x_hi 2^b) w_n^(x_lo)) and a short FP64 recurrence per thread, then rounded
// once to FP32: the twiddle values are those of the reference's FP64-computed, once-rounded tables
// (src/onefft.c:112-141) without reading an n-entry table from HBM.
#include "fft_common.hpp"
#include "kernels.hpp"

namespace onefft {

template <int LOG2NF, bool FIRST, bool SWAP_IN, bool SWAP_OUT>
__global__ __launch_bounds__(pass_wg(LOG2NF)) void fft_pass_kernel(PassArgs a) {
    constexpr int LOG2P = pass_log2p(LOG2NF);
    constexpr int NF = 1 << LOG2NF;
    constexpr int P = 1 << LOG2P;
    constexpr uint32_t T = NF / P;  // threads per column
    constexpr uint32_t C = pass_cols(LOG2NF);
    constexpr int S = stage_count(LOG2NF, LOG2P);
    constexpr uint32_t SPAN = lds_span(NF) | 1;  // odd: the C columns start on different banks
    __shared__ cf lds[S > 1 ? C * SPAN : 1];

    const uint32_t t = threadIdx.x;
    const uint32_t c = t % C;
    const uint32_t j = t / C;
    const uint32_t col = blockIdx.x * C + c;
    const cf *src = a.in + (uint64_t)blockIdx.y * a.in_dist;
    cf *dst = a.out + (uint64_t)blockIdx.y * a.out_dist;
    cf *L = lds + (S > 1 ? c * SPAN : 0);

    uint32_t in_base, in_step, out_base, out_step, low = 0;
    if constexpr (FIRST) {
        in_base = col;
        in_step = a.columns;
        uint32_t row = 0, rest = col;
        for (uint32_t d = 0; d < a.ndig; d++) {
            row += (rest & ((1u << a.dig_log2[d]) - 1)) << a.dig_shift[d];
            rest >>= a.dig_log2[d];
        }
        out_base = row << LOG2NF;
        out_step = 1;
    } else {
        low = col & ((1u << a.log2m) - 1);
        in_base = ((col >> a.log2m) << (a.log2m + LOG2NF)) + low;
        in_step = 1u << a.log2m;
        out_base = in_base;
        out_step = in_step;
    }

    cf v[P];
    static_for<0, S>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        constexpr int LR = stage_log2r(LOG2NF, LOG2P, s);
        constexpr int R = 1 << LR;
        constexpr int LNS = stage_log2ns(LOG2NF, LOG2P, s);
        constexpr uint32_t NS = 1u << LNS;
        constexpr int BF = P / R;
        constexpr uint32_t STRIDE = NF / R;
        constexpr uint32_t TWO = stage_tw_offset(LOG2NF, LOG2P, s);

        if constexpr (s == 0) {
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R + rr] = maybe_swap<SWAP_IN>(src[in_base + (j + T * qq + rr * STRIDE) * in_step]);
                });
            });
            if constexpr (!FIRST) {
                // element e = j + T q + STRIDE r gets w_n^((e low) << tw_shift)
                const uint32_t lo_mask = (1u << a.lo_bits) - 1;
                auto W = [&](uint32_t x) {
                    x <<= a.tw_shift;
                    return mul64(a.tw_hi[x >> a.lo_bits], a.tw_lo[x & lo_mask]);
                };
                cd wq = W(j * low);
                const cd step_q = W(T * low);
                const cd step_r = W(STRIDE * low);
                static_for<0, BF>([&](auto q) {
                    constexpr int qq = decltype(q)::value;
                    cd w = wq;
                    static_for<0, R>([&](auto r) {
                        constexpr int rr = decltype(r)::value;
                        v[qq * R + rr] = mul(v[qq * R + rr], round32(w));
                        if constexpr (rr + 1 < R) w = mul64(w, step_r);
                    });
                    if constexpr (qq + 1 < BF) wq = mul64(wq, step_q);
                });
            }
        } else {
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R + rr] = L[lds_pad(j + T * qq + rr * STRIDE)];
                });
            });
        }

        static_for<0, BF>([&](auto q) {
            constexpr int qq = decltype(q)::value;
            if constexpr (s > 0) {
                const uint32_t k = (j + T * qq) & (NS - 1);
                static_for<1, R>([&](auto r) {
                    constexpr int rr = decltype(r)::value;
                    v[qq * R + rr] = mul(v[qq * R + rr], a.tw[TWO + (rr - 1) * NS + k]);
                });
            }
            dft<R>(&v[qq * R]);
        });

        if constexpr (s == S - 1) {
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto m) {
                    constexpr int qq = decltype(q)::value, mm = decltype(m)::value;
                    dst[out_base + (j + T * qq + mm * STRIDE) * out_step] = maybe_swap<SWAP_OUT>(v[qq * R + mm]);
                });
            });
        } else {
            if constexpr (s > 0) __syncthreads();
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

template <int LOG2NF>
static hipError_t launch_pass_n(bool first, bool swap_in, bool swap_out, const PassArgs &a, uint64_t batch,
                                hipStream_t stream) {
    constexpr uint32_t C = pass_cols(LOG2NF);
    if (a.columns % C || batch == 0 || batch > 65535) return hipErrorInvalidValue;
    dim3 grid(a.columns / C, (uint32_t)batch), block(pass_wg(LOG2NF));
    if (first) {
        if (swap_in)
            hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, true, true, false>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, true, false, false>), grid, block, 0, stream, a);
    } else {
        if (swap_out)
            hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, false, false, true>), grid, block, 0, stream, a);
        else
            hipLaunchKernelGGL((fft_pass_kernel<LOG2NF, false, false, false>), grid, block, 0, stream, a);
    }
    return hipGetLastError();
}

hipError_t launch_pass(int log2nf, bool first, bool swap_in, bool swap_out, const PassArgs &a, uint64_t batch,
                       hipStream_t stream) {
    // a first pass never stores the final result (there are at least two passes)
    if (first && swap_out) return hipErrorInvalidValue;
    if (!first && swap_in) return hipErrorInvalidValue;
    switch (log2nf) {
    case 4: return launch_pass_n<4>(first, swap_in, swap_out, a, batch, stream);
    case 5: return launch_pass_n<5>(first, swap_in, swap_out, a, batch, stream);
    case 6: return launch_pass_n<6>(first, swap_in, swap_out, a, batch, stream);
    case 7: return launch_pass_n<7>(first, swap_in, swap_out, a, batch, stream);
    case 8: return launch_pass_n<8>(first, swap_in, swap_out, a, batch, stream);
    case 9: return launch_pass_n<9>(first, swap_in, swap_out, a, batch, stream);
    case 10: return launch_pass_n<10>(first, swap_in, swap_out, a, batch, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace onefft
