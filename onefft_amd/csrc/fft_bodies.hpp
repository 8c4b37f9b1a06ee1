// The two transform bodies every kernel is built from, as device functions so the standalone
// kernels (k_batch.hip, k_pass.hip) and the persistent ragged kernel (k_ragged.hip) run the same
// code:
//
//   single_body  one whole transform of up to 2^14 points in registers and LDS (Stockham stages,
//                radix 16 with a smaller first radix where needed);
//   pass_body    one tile of C adjacent columns of one pass of a multi-pass transform: an
//                N_f-point sub-transform per column, with the pass twiddles formed in FP64.
//
// Every thread of the workgroup must call a body with the same template arguments: the bodies
// contain workgroup barriers.
#pragma once

#include "fft_common.hpp"
#include "kernels.hpp"

namespace onefft {

// ---- single pass ------------------------------------------------------------------------------
//
// Thread j of the transform (0 <= j < 2^LOG2N / P) works on the transform that io addresses
// (fft_common.hpp: PtrIO, BufIO); `live` false skips the global stores but keeps the thread in
// every barrier. L is the transform's LDS region of lds_extent<LAYOUT>(2^LOG2N) values (unused
// when the transform fits one thread).

template <int LOG2N>
struct Single {
    static constexpr int LOG2P = batch_log2p(LOG2N);
    static constexpr int N = 1 << LOG2N;
    static constexpr int P = 1 << LOG2P;
    static constexpr uint32_t T = N / P;
    static constexpr int S = stage_count(LOG2N, LOG2P);
};

// stage 0's inputs: element j + T q + r N / R_0 into v[q R_0 + r]
template <int LOG2N, bool INV, class IO>
__device__ __forceinline__ void single_load(const IO &io, uint32_t j, cf (&v)[Single<LOG2N>::P]) {
    using G = Single<LOG2N>;
    constexpr int R = 1 << stage_log2r(LOG2N, G::LOG2P, 0);
    constexpr int BF = G::P / R;
    constexpr uint32_t STRIDE = G::N / R;
    static_for<0, BF>([&](auto q) {
        static_for<0, R>([&](auto r) {
            constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
            v[qq * R + rr] = maybe_conj<INV>(io.load(j, G::T * qq + rr * STRIDE));
        });
    });
}

template <int LOG2N, bool INV, int LAYOUT, int TW64, class IO>
__device__ __forceinline__ void single_finish(cf (&v)[Single<LOG2N>::P], const IO &io, bool live,
                                              const cf *__restrict__ tw, cf *L, uint32_t j,
                                              const cd *__restrict__ tw64) {
    using G = Single<LOG2N>;
    constexpr int LOG2P = G::LOG2P, N = G::N, P = G::P, S = G::S;
    constexpr uint32_t T = G::T;

    static_for<0, S>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        constexpr int LR = stage_log2r(LOG2N, LOG2P, s);
        constexpr int R = 1 << LR;
        constexpr int LNS = stage_log2ns(LOG2N, LOG2P, s);
        constexpr uint32_t NS = 1u << LNS;
        constexpr int BF = P / R;  // butterflies per thread
        constexpr uint32_t STRIDE = N / R;
        constexpr uint32_t TWO = stage_tw_offset(LOG2N, LOG2P, s);
        constexpr bool BASE64 = stage_uses_tw64(LOG2N, LOG2P, s, TW64);
        constexpr uint32_t TWO64 = stage_tw64_offset(LOG2N, LOG2P, s);

        if constexpr (s > 0 && LAYOUT != kHalves) {  // (kHalves: the previous exchange loaded v)
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R + rr] = L[lds_at<LAYOUT>(j + T * qq, rr * STRIDE)];
                });
            });
        }

        static_for<0, BF>([&](auto q) {
            constexpr int qq = decltype(q)::value;
            if constexpr (s > 0) {
                const uint32_t k = (j + T * qq) & (NS - 1);
                if constexpr (BASE64) {
                    // w^(r k) = w^k ... w^k in FP64, rounded once to FP32
                    const cd base = tw64[TWO64 + k];
                    cd w = base;
                    static_for<1, R>([&](auto r) {
                        constexpr int rr = decltype(r)::value;
                        v[qq * R + rr] = mul(v[qq * R + rr], round32(w));
                        if constexpr (rr + 1 < R) w = mul64(w, base);
                    });
                } else {
                    static_for<1, R>([&](auto r) {
                        constexpr int rr = decltype(r)::value;
                        v[qq * R + rr] = mul(v[qq * R + rr], tw[TWO + (rr - 1) * NS + k]);
                    });
                }
            }
            dft<R>(&v[qq * R]);
        });

        if constexpr (s == S - 1) {
            // NS * R == N here, so output m of butterfly b goes to b + m N / R
            if (live) {
                static_for<0, BF>([&](auto q) {
                    static_for<0, R>([&](auto m) {
                        constexpr int qq = decltype(q)::value, mm = decltype(m)::value;
                        io.store(j, T * qq + mm * STRIDE, maybe_conj<INV>(v[qq * R + mm]));
                    });
                });
            }
        } else if constexpr (LAYOUT == kHalves) {
            // the exchange to stage s + 1, real parts first, then imaginary parts, through one
            // float buffer: write (this stage's pattern), barrier, read (the next stage's), barrier
            float *Lf = reinterpret_cast<float *>(L);
            constexpr int R1 = 1 << stage_log2r(LOG2N, LOG2P, s + 1);
            constexpr int BF1 = P / R1;
            constexpr uint32_t STRIDE1 = N / R1;
            auto put = [&](auto part) {
                static_for<0, BF>([&](auto q) {
                    constexpr int qq = decltype(q)::value;
                    const uint32_t b = j + T * qq;
                    const uint32_t base = ((b >> LNS) << (LNS + LR)) + (b & (NS - 1));
                    static_for<0, R>([&](auto m) {
                        constexpr int mm = decltype(m)::value;
                        Lf[lds_at<kPadded>(base, mm * NS)] = decltype(part)::value ? v[qq * R + mm].i : v[qq * R + mm].r;
                    });
                });
            };
            float re[P];
            if constexpr (s > 0) __syncthreads();  // every read of the previous exchange is done
            put(std::false_type{});
            __syncthreads();
            static_for<0, BF1>([&](auto q) {
                static_for<0, R1>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    re[qq * R1 + rr] = Lf[lds_at<kPadded>(j + T * qq, rr * STRIDE1)];
                });
            });
            __syncthreads();
            put(std::true_type{});
            __syncthreads();
            static_for<0, BF1>([&](auto q) {
                static_for<0, R1>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R1 + rr] = mk(re[qq * R1 + rr], Lf[lds_at<kPadded>(j + T * qq, rr * STRIDE1)]);
                });
            });
        } else {
            if constexpr (s > 0) __syncthreads();  // every read of this stage is done
            static_for<0, BF>([&](auto q) {
                constexpr int qq = decltype(q)::value;
                const uint32_t b = j + T * qq;
                const uint32_t base = ((b >> LNS) << (LNS + LR)) + (b & (NS - 1));
                static_for<0, R>([&](auto m) {
                    constexpr int mm = decltype(m)::value;
                    L[lds_at<LAYOUT>(base, mm * NS)] = v[qq * R + mm];
                });
            });
            __syncthreads();
        }
    });
}

// `live` false: the loads still run (the caller points io at a valid transform), the stores do not
template <int LOG2N, bool INV, int LAYOUT, int TW64, class IO>
__device__ __forceinline__ void single_body(const IO &io, bool live, const cf *__restrict__ tw, cf *L, uint32_t j,
                                            const cd *__restrict__ tw64) {
    cf v[Single<LOG2N>::P];
    single_load<LOG2N, INV>(io, j, v);
    single_finish<LOG2N, INV, LAYOUT, TW64>(v, io, live, tw, L, j, tw64);
}

// ---- one tile of one pass of a multi-pass transform -------------------------------------------
//
// Columns tile*C .. tile*C + C-1 of the pass described by `a` (see k_pass.hip for the index
// scheme), on the transform at src/dst. Thread t of the C * 2^LOG2NF / P threads handles column
// t % C. lds holds C * pass_span(LOG2NF, C, LAYOUT) complex values (kPadded) or floats (kHalves:
// the stage exchanges move real and imaginary parts in turn, as in single_finish).
//
// The columns' LDS regions lie pass_span values apart. A wave's 64 lanes cover C columns x 64/C
// rows, and a row step moves one 4-byte bank (halves) or two (padded): the regions must start on
// banks that keep the C columns' lanes apart. With 32 banks of 4 bytes (the LDS of gfx950 as its
// bank-conflict counter sees it) that is a column step of an odd multiple of 32/C banks when
// C < 32 (an odd number of values otherwise). A bank model of every pass shape, checked against
// SQ_LDS_BANK_CONFLICT (640 and 0 cycles per wave measured and modelled for the 2^11 x 4 and
// 2^8 x 32 tiles with the old odd spans), finds this choice optimal for every allowed shape; with
// few columns it removes most conflicts (2^11 x 4: 640 -> 64 cycles per wave, 2^12 x 4: 640 -> 0).
__host__ __device__ constexpr uint32_t pass_span(int log2nf, int cols, int layout) {
    const uint32_t need = lds_span(1u << log2nf);
    const uint32_t width = layout == kHalves ? 1 : 2;  // banks per value
    const uint32_t unit = 32u / (uint32_t)cols;        // banks between columns' regions
    if (unit <= width) return need | 1;
    const uint32_t u = unit / width, r = need % (2 * u);  // want span = u (mod 2u), in values
    return need + (u >= r ? u - r : 3 * u - r);
}

template <int LOG2NF, int C, bool FIRST, bool SWAP_IN, bool SWAP_OUT, int LAYOUT = kPadded, bool COL_STEPS = false>
__device__ __forceinline__ void pass_body(const PassArgs &a, const cf *__restrict__ src, cf *__restrict__ dst,
                                          uint32_t tile, cf *lds, uint32_t t) {
    static_assert(LAYOUT == kPadded || LAYOUT == kHalves, "pass kernels use the padded or the halves layout");
    constexpr int LOG2P = pass_log2p(LOG2NF);
    constexpr int NF = 1 << LOG2NF;
    constexpr int P = 1 << LOG2P;
    constexpr uint32_t T = NF / P;  // threads per column
    constexpr int S = stage_count(LOG2NF, LOG2P);
    constexpr uint32_t SPAN = pass_span(LOG2NF, C, LAYOUT);  // the columns start on distinct banks

    // A first pass stores each column's outputs as one row, and the rows of neighbouring columns lie
    // far apart, so a wave's direct stores write C pieces of 64 / C values each. With C >= 8 (pieces
    // under 128 bytes) the final stage's outputs go through LDS instead, and each wave stores 64
    // consecutive values of one row (128, in 16-byte stores, with PAIR_STORES where the output is
    // 16-byte aligned).
    constexpr bool STAGE_OUT = FIRST && C >= 8 && S > 1;
    // staged stores of value pairs (16-byte stores): measured 1-3 % faster for the 2^11-point
    // input pass with 8 columns; the 2^9- and 2^10-point halves-layout input passes compiled with
    // them to 70-73 VGPRs (fewer workgroups per CU), so they store single values
    constexpr bool PAIR_STORES = STAGE_OUT && LAYOUT == kHalves && LOG2NF >= 11;
    const uint32_t c = t % C;
    const uint32_t j = t / C;
    const uint32_t col = tile * C + c;
    // a column's region: SPAN complex values, or with the halves layout SPAN floats
    cf *L = lds + c * SPAN;
    float *Lf = reinterpret_cast<float *>(lds) + c * SPAN;
    // the output row of input column `cl` (first pass): its digits j_1.. (j_1 fastest) weighted
    auto row_of = [&](uint32_t cl) {
        uint32_t row = 0, rest = cl;
        static_for<0, kMaxPasses - 1>([&](auto d) {  // static indexes: the arguments stay in registers
            if (decltype(d)::value < a.ndig) {
                row += (rest & ((1u << a.dig_log2[d]) - 1)) << a.dig_shift[d];
                rest >>= a.dig_log2[d];
            }
        });
        return row;
    };

    uint32_t in_base, in_step, out_base, out_step, low = 0;
    if constexpr (FIRST) {
        in_base = col;
        in_step = a.columns;
        out_base = row_of(col) << LOG2NF;
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

        if constexpr (s == 0) {
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R + rr] = maybe_swap<SWAP_IN>(src[in_base + (j + T * qq + rr * STRIDE) * in_step]);
                });
            });
            if constexpr (!FIRST) {
                // element e = j + T q + STRIDE r gets w_n^((e low) << tw_shift), formed in FP64 from
                // the bases w^(j low), w^(T low) and w^(STRIDE low). With low = low0 + (c mod M_f),
                // low0 the tile's first column mod M_f (uniform in the workgroup), w^(x low) =
                // w^(x low0) w^(x (c mod M_f)): the first factor takes at most T distinct values per
                // workgroup (nearly broadcast loads), the second comes from the pass's column table,
                // read lane by lane in order. Looked up directly, w^(j low) gathers 64 scattered
                // entries per wave. The step bases are looked up directly unless COL_STEPS (passes
                // whose lookups of x low sweep more of the low table than the L1 holds, runtime.hip)
                auto W = [&](uint32_t x) { return two_level(a.tw_hi, a.tw_lo, a.lo_bits, x << a.tw_shift); };
                const uint32_t low0 = (tile * C) & ((1u << a.log2m) - 1);
                cd wq = mul64(W(j * low0), a.tw_col[t]);
                cd step_q, step_r;
                if constexpr (COL_STEPS) {
                    step_q = mul64(W(T * low0), a.tw_col[T * C + c]);
                    step_r = mul64(W(STRIDE * low0), a.tw_col[T * C + C + c]);
                } else {
                    step_q = W(T * low);
                    step_r = W(STRIDE * low);
                }
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
        } else if constexpr (LAYOUT != kHalves) {  // (kHalves: the previous exchange loaded v)
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R + rr] = L[lds_at(j + T * qq, rr * STRIDE)];
                });
            });
        }

        static_for<0, BF>([&](auto q) {
            constexpr int qq = decltype(q)::value;
            if constexpr (s > 0) {
                const uint32_t k = (j + T * qq) & (NS - 1);
                // every stage after the first has radix 16: w^(r k) by an FP64 recurrence from the
                // base w^k, each rounded once (FP64-base level 3, fft_common.hpp: stage_uses_tw64)
                static_assert(stage_uses_tw64(LOG2NF, LOG2P, s, 3), "pass stages after the first have FP64 bases");
                const cd base = a.tw64[stage_tw64_offset(LOG2NF, LOG2P, s) + k];
                cd w = base;
                static_for<1, R>([&](auto r) {
                    constexpr int rr = decltype(r)::value;
                    v[qq * R + rr] = mul(v[qq * R + rr], round32(w));
                    if constexpr (rr + 1 < R) w = mul64(w, base);
                });
            }
            dft<R>(&v[qq * R]);
        });

        if constexpr (s == S - 1 && STAGE_OUT) {
            // output j + T q + STRIDE m of column c goes to its region's natural position; then
            // value k of thread t is tile element g = k * THREADS + t (with PAIR_STORES, pair k is
            // the elements g, g + 1 with g = 2 (k THREADS + t)): column g / NF, position g mod NF,
            // stored to that column's row
            constexpr uint32_t THREADS = C * T;
            auto put = [&](auto part) {
                static_for<0, BF>([&](auto q) {
                    static_for<0, R>([&](auto m) {
                        constexpr int qq = decltype(q)::value, mm = decltype(m)::value;
                        const uint32_t at = lds_index(j + T * qq + mm * STRIDE);
                        if constexpr (LAYOUT == kHalves)
                            Lf[at] = decltype(part)::value ? v[qq * R + mm].i : v[qq * R + mm].r;
                        else
                            L[at] = v[qq * R + mm];
                    });
                });
            };
            __syncthreads();  // every read of the previous exchange is done
            put(std::false_type{});
            __syncthreads();
            if constexpr (LAYOUT == kHalves) {
                const float *lf0 = reinterpret_cast<const float *>(lds);
                auto value = [&](uint32_t g) { return lf0[(g >> LOG2NF) * SPAN + lds_index(g & (NF - 1))]; };
                auto row = [&](uint32_t g) { return dst + ((uint64_t)row_of(tile * C + (g >> LOG2NF)) << LOG2NF) + (g & (NF - 1)); };
                float re[P];
                if constexpr (PAIR_STORES) {
                    // thread t takes the value pairs g, g + 1 with g = 2 (k THREADS + t), and each
                    // wave writes 1 KiB of one row: in 16-byte stores where the rows are 16-byte
                    // aligned (always, but for batches or ragged items whose output starts at an
                    // odd offset)
                    static_for<0, P>([&](auto k) {
                        constexpr int kk = decltype(k)::value;
                        re[kk] = value(2 * ((kk / 2) * THREADS + t) + (kk & 1));
                    });
                    __syncthreads();
                    put(std::true_type{});
                    __syncthreads();
                    auto store = [&](auto wide) {
                        static_for<0, P / 2>([&](auto k) {
                            constexpr int kk = decltype(k)::value;
                            const uint32_t g = 2 * (kk * THREADS + t);
                            const cf w0 = maybe_swap<SWAP_OUT>(mk(re[2 * kk], value(g)));
                            const cf w1 = maybe_swap<SWAP_OUT>(mk(re[2 * kk + 1], value(g + 1)));
                            cf *to = row(g);
                            if constexpr (decltype(wide)::value) {
                                *reinterpret_cast<float4 *>(to) = make_float4(w0.r, w0.i, w1.r, w1.i);
                            } else {
                                to[0] = w0;
                                to[1] = w1;
                            }
                        });
                    };
                    if (((uintptr_t)dst & 15) == 0) store(std::true_type{});
                    else store(std::false_type{});
                } else {
                    static_for<0, P>([&](auto k) { re[decltype(k)::value] = value(decltype(k)::value * THREADS + t); });
                    __syncthreads();
                    put(std::true_type{});
                    __syncthreads();
                    static_for<0, P>([&](auto k) {
                        const uint32_t g = decltype(k)::value * THREADS + t;
                        *row(g) = maybe_swap<SWAP_OUT>(mk(re[decltype(k)::value], value(g)));
                    });
                }
            } else {
                static_for<0, P>([&](auto k) {
                    const uint32_t g = decltype(k)::value * THREADS + t;
                    const cf w = lds[(g >> LOG2NF) * SPAN + lds_index(g & (NF - 1))];
                    dst[((uint64_t)row_of(tile * C + (g >> LOG2NF)) << LOG2NF) + (g & (NF - 1))] = maybe_swap<SWAP_OUT>(w);
                });
            }
        } else if constexpr (s == S - 1) {
            static_for<0, BF>([&](auto q) {
                static_for<0, R>([&](auto m) {
                    constexpr int qq = decltype(q)::value, mm = decltype(m)::value;
                    dst[out_base + (j + T * qq + mm * STRIDE) * out_step] = maybe_swap<SWAP_OUT>(v[qq * R + mm]);
                });
            });
        } else if constexpr (LAYOUT == kHalves) {
            // the exchange to stage s + 1 through one float buffer, real parts, then imaginary
            // parts (as in single_finish)
            constexpr int R1 = 1 << stage_log2r(LOG2NF, LOG2P, s + 1);
            constexpr int BF1 = P / R1;
            constexpr uint32_t STRIDE1 = NF / R1;
            auto put = [&](auto part) {
                static_for<0, BF>([&](auto q) {
                    constexpr int qq = decltype(q)::value;
                    const uint32_t b = j + T * qq;
                    const uint32_t base = ((b >> LNS) << (LNS + LR)) + (b & (NS - 1));
                    static_for<0, R>([&](auto m) {
                        constexpr int mm = decltype(m)::value;
                        Lf[lds_at(base, mm * NS)] = decltype(part)::value ? v[qq * R + mm].i : v[qq * R + mm].r;
                    });
                });
            };
            float re[P];
            if constexpr (s > 0) __syncthreads();  // every read of the previous exchange is done
            put(std::false_type{});
            __syncthreads();
            static_for<0, BF1>([&](auto q) {
                static_for<0, R1>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    re[qq * R1 + rr] = Lf[lds_at(j + T * qq, rr * STRIDE1)];
                });
            });
            __syncthreads();
            put(std::true_type{});
            __syncthreads();
            static_for<0, BF1>([&](auto q) {
                static_for<0, R1>([&](auto r) {
                    constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
                    v[qq * R1 + rr] = mk(re[qq * R1 + rr], Lf[lds_at(j + T * qq, rr * STRIDE1)]);
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
                    L[lds_at(base, mm * NS)] = v[qq * R + mm];
                });
            });
            __syncthreads();
        }
    });
}

}  // namespace onefft
