// The arithmetic behind the plan tables, free of HIP so that a CPU program can build and check it
// (the kernels include it through fft_common.hpp): the complex types, the FP64 product the kernels
// form twiddles with, the Stockham stage schedule, and the twiddle tables the host uploads.
//
// Reference: the twiddle function onefft_root_double / onefft_root (src/onefft.c:112-141).
// A CPU build defines __host__, __device__ and __forceinline__ itself (empty, empty, inline).
#pragma once

#include <stdint.h>

#include <cmath>

#ifndef ONEFFT_HD
#define ONEFFT_HD __host__ __device__ __forceinline__
#endif

namespace onefft {

// One complex value, 8-byte aligned so a load or store is one dwordx2.
struct __attribute__((aligned(8))) cf {
    float r, i;
};

// One complex value in FP64, for twiddle products that are rounded to FP32 only once.
struct __attribute__((aligned(16))) cd {
    double r, i;
};

ONEFFT_HD cf mk(float r, float i) {
    cf c;
    c.r = r;
    c.i = i;
    return c;
}
// a * b in FP64, written with explicit fused multiply-adds so that the result is the same on the
// GPU (v_fma_f64) and on any IEEE host (std::fma): every twiddle the kernels form is this product
// of table values, rounded once to FP32, and a CPU program can reproduce it bit for bit.
ONEFFT_HD cd mul64(cd a, cd b) {
    cd c;
    c.r = __builtin_fma(a.r, b.r, -(a.i * b.i));
    c.i = __builtin_fma(a.r, b.i, a.i * b.r);
    return c;
}
ONEFFT_HD cf round32(cd a) { return mk((float)a.r, (float)a.i); }
// w_n^x from the two-level FP64 tables of w_n: hi[i] = w_n^(i 2^lo_bits), lo[i] = w_n^i
ONEFFT_HD cd two_level(const cd *hi, const cd *lo, uint32_t lo_bits, uint32_t x) {
    return mul64(hi[x >> lo_bits], lo[x & ((1u << lo_bits) - 1)]);
}

// ---- Stockham stage schedule ------------------------------------------------------------------
//
// An N-point transform in which every thread holds P points runs in stages. Stage s has radix
// R_s (a divisor of P) and Ns = R_0 ... R_{s-1}; each thread runs P / R_s butterflies. Butterfly
// b reads x[b + r N / R_s], multiplies input r by w_{Ns R_s}^{r (b mod Ns)}, transforms, and
// writes output m to (b / Ns) Ns R_s + (b mod Ns) + m Ns. After the last stage the result is in
// natural order (Stockham autosort; Govindaraju et al., SC 2008). The smaller radix, if any,
// comes first.

__host__ __device__ constexpr int stage_count(int log2n, int log2p) {
    return log2n <= log2p ? 1 : log2n / log2p + (log2n % log2p ? 1 : 0);
}
__host__ __device__ constexpr int stage_log2r(int log2n, int log2p, int s) {
    return log2n <= log2p ? log2n : (log2n % log2p && s == 0) ? log2n % log2p : log2p;
}
// log2 of Ns for stage s
__host__ __device__ constexpr int stage_log2ns(int log2n, int log2p, int s) {
    int l = 0;
    for (int i = 0; i < s; i++) l += stage_log2r(log2n, log2p, i);
    return l;
}
// Offset (in complex entries) of stage s's twiddles in the per-size table, and the table size.
// Stage s stores w_{Ns R}^{r k} at [(r - 1) Ns + k] for r = 1..R-1, k = 0..Ns-1.
__host__ __device__ constexpr uint32_t stage_tw_offset(int log2n, int log2p, int s) {
    uint32_t off = 0;
    for (int i = 1; i < s; i++) {
        uint32_t r = 1u << stage_log2r(log2n, log2p, i);
        off += (r - 1) << stage_log2ns(log2n, log2p, i);
    }
    return off;
}
__host__ __device__ constexpr uint32_t stage_tw_size(int log2n, int log2p) {
    return stage_tw_offset(log2n, log2p, stage_count(log2n, log2p));
}

// Stages after the first with radix >= 4 can take one FP64 base w_{Ns R}^k per butterfly instead
// of R - 1 table twiddles and form w^(r k), r = 2..R-1, by an FP64 recurrence, each rounded once
// to FP32: 16 bytes of table per butterfly instead of 8 (R - 1), at ~6 FP64 instructions per
// twiddle. Plans hold the FP64 bases of every such stage (w_{Ns R}^k for k < Ns at
// stage_tw64_offset(.., s)); a kernel uses them for the stages with Ns >= 2^tw64_min_log2ns(level),
// the level being a tuning choice: 0 none, 1 Ns >= 64 (the stages whose FP32 tables outgrow the
// L1 cache), 2 Ns >= 16, 3 every such stage.
__host__ __device__ constexpr bool stage_has_tw64(int log2n, int log2p, int s) {
    return s > 0 && stage_log2r(log2n, log2p, s) >= 2;
}
__host__ __device__ constexpr int tw64_min_log2ns(int level) {
    return level <= 0 ? 64 : level == 1 ? 6 : level == 2 ? 4 : 0;
}
__host__ __device__ constexpr bool stage_uses_tw64(int log2n, int log2p, int s, int level) {
    return stage_has_tw64(log2n, log2p, s) && stage_log2ns(log2n, log2p, s) >= tw64_min_log2ns(level);
}
__host__ __device__ constexpr uint32_t stage_tw64_offset(int log2n, int log2p, int s) {
    uint32_t off = 0;
    for (int i = 1; i < s; i++)
        if (stage_has_tw64(log2n, log2p, i)) off += 1u << stage_log2ns(log2n, log2p, i);
    return off;
}
__host__ __device__ constexpr uint32_t stage_tw64_size(int log2n, int log2p) {
    return stage_tw64_offset(log2n, log2p, stage_count(log2n, log2p));
}
// the highest level that changes the kernel of this size (0: no stage can use FP64 bases)
__host__ __device__ constexpr int tw64_levels(int log2n, int log2p) {
    int top = 0;
    for (int level = 1; level <= 3; level++)
        for (int s = 1; s < stage_count(log2n, log2p); s++)
            if (stage_uses_tw64(log2n, log2p, s, level) && !stage_uses_tw64(log2n, log2p, s, level - 1)) top = level;
    return top;
}

// ---- twiddle tables (host) -------------------------------------------------------------------
//
// exp(-2 pi i k / 2^log2size) in FP64. The phase k / 2^log2size of a turn is held exactly as a
// 0.64 fixed-point fraction; the angle is folded into the first octant before sin/cos and the
// quadrant is applied exactly, so values are exactly symmetric and exact at multiples of an
// eighth turn. Same steps as the reference's onefft_root_double (src/onefft.c:112-135), which
// fixes the twiddle values bit for bit once rounded to FP32 (onefft_root, :137-141).
inline void root64(uint64_t k, unsigned log2size, double *re, double *im) {
    const uint64_t phase = log2size ? k << (64 - log2size) : 0;
    const unsigned quadrant = (unsigned)(phase >> 62);
    const uint64_t in_quadrant = phase & ((uint64_t(1) << 62) - 1);
    const double unit = 6.283185307179586476925286766559 / 18446744073709551616.0;  // radians per 2^-64 turn
    double c, s;
    if (in_quadrant > (uint64_t(1) << 61)) {
        const double a = (double)((uint64_t(1) << 62) - in_quadrant) * unit;
        c = std::sin(a);
        s = std::cos(a);
    } else {
        const double a = (double)in_quadrant * unit;
        c = std::cos(a);
        s = std::sin(a);
    }
    switch (quadrant) {  // (c - i s) (-i)^quadrant
    case 0: *re = c, *im = -s; break;
    case 1: *re = -s, *im = -c; break;
    case 2: *re = -c, *im = s; break;
    default: *re = s, *im = c; break;
    }
}

inline cf root32(uint64_t k, unsigned log2size) {
    double re, im;
    root64(k, log2size, &re, &im);
    return mk((float)re, (float)im);
}

// Stage twiddles of an in-LDS transform of 2^log2n points, P = 2^log2p points per thread, in
// the order the kernels read them (stage_tw_offset).
inline void fill_stage_table(int log2n, int log2p, cf *table) {
    const int stages = stage_count(log2n, log2p);
    for (int s = 1; s < stages; s++) {
        const uint32_t off = stage_tw_offset(log2n, log2p, s);
        const int lr = stage_log2r(log2n, log2p, s);
        const int lns = stage_log2ns(log2n, log2p, s);
        const uint32_t r_count = 1u << lr, ns = 1u << lns;
        for (uint32_t r = 1; r < r_count; r++)
            for (uint32_t k = 0; k < ns; k++)
                table[off + (r - 1) * ns + k] = root32((uint64_t)r * k, (unsigned)(lns + lr));
    }
}

// FP64 bases w_{Ns R}^k, k < Ns, of every stage that can use them (stage_has_tw64).
inline void fill_stage_table64(int log2n, int log2p, cd *table) {
    const int stages = stage_count(log2n, log2p);
    for (int s = 1; s < stages; s++) {
        if (!stage_has_tw64(log2n, log2p, s)) continue;
        const uint32_t off = stage_tw64_offset(log2n, log2p, s);
        const int lr = stage_log2r(log2n, log2p, s);
        const int lns = stage_log2ns(log2n, log2p, s);
        for (uint32_t k = 0; k < (1u << lns); k++)
            root64(k, (unsigned)(lns + lr), &table[off + k].r, &table[off + k].i);
    }
}

// The two-level FP64 tables of w_n, n = 2^log2n: lo[i] = w_n^i for i < 2^lo_bits and
// hi[i] = w_n^(i 2^lo_bits) for i < 2^(log2n - lo_bits), lo_bits = ceil(log2n / 2).
inline unsigned two_level_lo_bits(unsigned log2n) { return (log2n + 1) / 2; }
inline void fill_two_level(unsigned log2n, cd *lo, cd *hi) {
    const unsigned lo_bits = two_level_lo_bits(log2n);
    for (uint64_t i = 0; i < (uint64_t(1) << lo_bits); i++) root64(i, log2n, &lo[i].r, &lo[i].i);
    for (uint64_t i = 0; i < (uint64_t(1) << (log2n - lo_bits)); i++) root64(i << lo_bits, log2n, &hi[i].r, &hi[i].i);
}

}  // namespace onefft
