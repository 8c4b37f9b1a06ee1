// CPU build of the library's plan-table arithmetic (onefft_amd/csrc/plan_tables.hpp), loaded by
// tests/test_twiddles.py through ctypes. TEST INFRASTRUCTURE ONLY.
//
//   g++ -std=c++17 -O2 -ffp-contract=off -D__host__= -D__device__= -D__forceinline__=inline
//       -shared -fPIC tests/twiddle_check.cpp -o <lib>
//
// Exposes the host tables (root64/root32, the FP32 stage tables, the FP64 stage bases and
// two-level tables) and replays, value by value, how the kernels form their twiddles from those
// tables in FP64 before rounding once to FP32:
//   * stage recurrences: w^(r k) = base * base * ... (r - 1 products), base = w_{Ns R}^k
//     (fft_bodies.hpp single_finish / pass_body stages s > 0, k_ragged.hip column_fft);
//   * pass twiddles: w_n^((j + T q + STRIDE r) low << shift) = W(j low) W(T low)^q W(STRIDE low)^r
//     with W the two-level product (k_ragged.hip first_stage); the pass kernels form the base
//     W(j low), and in some passes the other two, as W(x low0) times the plan's column-table entry
//     w^(x c) (low = low0 + c, c the column within the tile; fft_bodies.hpp pass_body stage 0).
// Each formed value is compared with root32 of the same exponent, i.e. with the reference's
// onefft_root (which the test pins root32 against). Two kinds of difference are expected: a
// component that onefft_root holds as exactly 0 (its octant folding makes quarter turns exact)
// comes out of the FP64 products as a residue of order 1e-16, and an exact value that lies within
// the FP64 formation error of an FP32 rounding boundary rounds one ulp the other way.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../onefft_amd/csrc/plan_tables.hpp"

using onefft::cd;
using onefft::cf;

namespace {

// Formed twiddles against onefft_root: equal; a component onefft_root holds as exactly 0 (a
// multiple of a quarter turn) that the FP64 products leave as a tiny residue; a one-ulp difference
// (an exact value within the formation error of an FP32 rounding boundary); anything else.
struct Count {
    uint64_t checked = 0, zero_residue = 0, one_ulp = 0, other = 0;
    double max_residue = 0;
};

int32_t ulps(float a, float b) {
    int32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    // map the sign-magnitude bit patterns onto a monotonic integer line
    if (x < 0) x = INT32_MIN - x;
    if (y < 0) y = INT32_MIN - y;
    const int64_t d = (int64_t)x - (int64_t)y;
    return (int32_t)(d < 0 ? -d : d);
}

void compare(Count &c, cd formed, uint64_t k, unsigned log2size) {
    const cf got = onefft::round32(formed);
    const cf want = onefft::root32(k, log2size);
    c.checked++;
    if (got.r == want.r && got.i == want.i) return;
    bool residue = false, ulp = false, other = false;
    auto component = [&](float g, float w) {
        if (g == w) return;
        if (w == 0.0f && std::fabs(g) < 1e-12f) {
            residue = true;
            if (std::fabs((double)g) > c.max_residue) c.max_residue = std::fabs((double)g);
        } else if (ulps(g, w) == 1) {
            ulp = true;
        } else {
            other = true;
        }
    };
    component(got.r, want.r);
    component(got.i, want.i);
    if (other) c.other++;
    else if (ulp) c.one_ulp++;
    else if (residue) c.zero_residue++;
}

// Stage s of an in-LDS transform of 2^log2n points, 2^log2p per thread, with FP64 bases: every
// butterfly index k < Ns and every input r >= 1 of its radix-R butterfly.
void stage_recurrence(Count &c, int log2n, int log2p, int s) {
    const int lr = onefft::stage_log2r(log2n, log2p, s);
    const int lns = onefft::stage_log2ns(log2n, log2p, s);
    const uint32_t R = 1u << lr, NS = 1u << lns;
    for (uint32_t k = 0; k < NS; k++) {
        cd base;
        onefft::root64(k, (unsigned)(lns + lr), &base.r, &base.i);
        cd w = base;
        for (uint32_t r = 1; r < R; r++) {
            compare(c, w, (uint64_t)r * k, (unsigned)(lns + lr));
            w = onefft::mul64(w, base);
        }
    }
}

// out = {checked, zero residues, one-ulp differences, other differences, largest residue (the
// bits of a double)}
void write(const Count &c, uint64_t *out) {
    out[0] = c.checked;
    out[1] = c.zero_residue;
    out[2] = c.one_ulp;
    out[3] = c.other;
    std::memcpy(&out[4], &c.max_residue, 8);
}

}  // namespace

extern "C" {

void tw_root32(uint64_t k, unsigned log2size, float *re, float *im) {
    const cf w = onefft::root32(k, log2size);
    *re = w.r;
    *im = w.i;
}

void tw_root64(uint64_t k, unsigned log2size, double *re, double *im) { onefft::root64(k, log2size, re, im); }

uint32_t tw_stage_table_size(int log2n, int log2p) { return onefft::stage_tw_size(log2n, log2p); }
uint32_t tw_stage_table64_size(int log2n, int log2p) { return onefft::stage_tw64_size(log2n, log2p); }

// out: stage_tw_size entries (re, im)
void tw_stage_table(int log2n, int log2p, float *out) {
    onefft::fill_stage_table(log2n, log2p, reinterpret_cast<cf *>(out));
}
// out: stage_tw64_size entries (re, im)
void tw_stage_table64(int log2n, int log2p, double *out) {
    onefft::fill_stage_table64(log2n, log2p, reinterpret_cast<cd *>(out));
}
// For every entry of the FP32 stage table: the exponent and the log2 size it is w of (so the
// test can ask the reference for the same value).
void tw_stage_table_exponents(int log2n, int log2p, uint64_t *k, uint32_t *log2size) {
    const int stages = onefft::stage_count(log2n, log2p);
    for (int s = 1; s < stages; s++) {
        const uint32_t off = onefft::stage_tw_offset(log2n, log2p, s);
        const int lr = onefft::stage_log2r(log2n, log2p, s);
        const int lns = onefft::stage_log2ns(log2n, log2p, s);
        for (uint32_t r = 1; r < (1u << lr); r++)
            for (uint32_t q = 0; q < (1u << lns); q++) {
                k[off + (r - 1) * (1u << lns) + q] = (uint64_t)r * q;
                log2size[off + (r - 1) * (1u << lns) + q] = (uint32_t)(lns + lr);
            }
    }
}

unsigned tw_two_level_lo_bits(unsigned log2n) { return onefft::two_level_lo_bits(log2n); }
// lo: 2^lo_bits entries, hi: 2^(log2n - lo_bits) entries, each (re, im) in FP64
void tw_two_level(unsigned log2n, double *lo, double *hi) {
    onefft::fill_two_level(log2n, reinterpret_cast<cd *>(lo), reinterpret_cast<cd *>(hi));
}

// The single-pass kernel of 2^log2n points at FP64-base level `level`: every stage that uses
// FP64 bases (out: as write() above).
void tw_count_single(int log2n, int level, uint64_t *out) {
    Count c;
    const int lp = log2n <= 4 ? log2n : 4;  // kernels.hpp: batch_log2p
    for (int s = 1; s < onefft::stage_count(log2n, lp); s++)
        if (onefft::stage_uses_tw64(log2n, lp, s, level)) stage_recurrence(c, log2n, lp, s);
    write(c, out);
}

// One pass of a multi-pass transform of 2^log2n points: sub-transforms of 2^lf points (16 per
// thread), `cols` columns per workgroup, low index extent 2^log2m (0 for the input pass, which
// has no pass twiddles). Counts the pass twiddles for one low index in every `low_step` (all for
// 1; otherwise base + (base / low_step) mod low_step for every multiple base, so that the column
// within the tile still varies), and the stage recurrences of the sub-transform (every stage
// after the first, level 3). The bases come as the kernels form them with the plan's column table
// (runtime.hip, fft_bodies.hpp pass_body): w^(j low) = W(j low0) w^(j cm) for low = low0 + cm,
// cm the column within the tile mod 2^log2m, w^(j cm) an exact root64; the step bases w^(T low)
// and w^(STRIDE low) the same way where lo_bits > shift + 10, else as two-level lookups.
void tw_count_pass(unsigned log2n, int lf, unsigned log2m, uint32_t cols, uint32_t low_step, uint64_t *out_pass,
                   uint64_t *out_stage) {
    Count stage;
    const int lp = lf < 4 ? lf : 4;  // kernels.hpp: pass_log2p
    for (int s = 1; s < onefft::stage_count(lf, lp); s++) stage_recurrence(stage, lf, lp, s);
    write(stage, out_stage);
    Count pass;
    if (log2m > 0) {
        const unsigned lo_bits = onefft::two_level_lo_bits(log2n);
        cd *lo = new cd[size_t(1) << lo_bits];
        cd *hi = new cd[size_t(1) << (log2n - lo_bits)];
        onefft::fill_two_level(log2n, lo, hi);
        const unsigned shift = log2n - log2m - (unsigned)lf;
        const uint32_t NF = 1u << lf, P = 1u << lp, T = NF / P;
        const uint32_t R = 1u << onefft::stage_log2r(lf, lp, 0), BF = P / R, STRIDE = NF / R;
        auto W = [&](uint32_t x) { return onefft::two_level(hi, lo, lo_bits, x << shift); };
        auto exact = [&](uint64_t x) {
            cd w;
            onefft::root64((x << shift) & ((uint64_t(1) << log2n) - 1), log2n, &w.r, &w.i);
            return w;
        };
        const bool steps = lo_bits > shift + 10;  // runtime.hip: PassArgs::col_steps
        for (uint64_t base = 0; base < (uint64_t(1) << log2m); base += low_step) {
            const uint64_t low = base + (base / low_step) % low_step;
            // c mod M for the column c of the tile that holds low, and the tile's first column mod M
            const uint32_t cm = (uint32_t)low & (cols - 1), low0 = (uint32_t)low - cm;
            for (uint32_t j = 0; j < T; j++) {
                cd wq = onefft::mul64(W(j * low0), exact((uint64_t)j * cm));
                const cd step_q = steps ? onefft::mul64(W(T * low0), exact((uint64_t)T * cm)) : W(T * (uint32_t)low);
                const cd step_r =
                    steps ? onefft::mul64(W(STRIDE * low0), exact((uint64_t)STRIDE * cm)) : W(STRIDE * (uint32_t)low);
                for (uint32_t q = 0; q < BF; q++) {
                    cd w = wq;
                    for (uint32_t r = 0; r < R; r++) {
                        const uint64_t e = j + T * q + STRIDE * r;
                        compare(pass, w, (e * low) << shift, log2n);
                        if (r + 1 < R) w = onefft::mul64(w, step_r);
                    }
                    if (q + 1 < BF) wq = onefft::mul64(wq, step_q);
                }
            }
        }
        delete[] lo;
        delete[] hi;
    }
    write(pass, out_pass);
}

// The ragged kernel's two-pass item of 2^log2n points (N_1 = 2^(log2n/2) transformed last):
// the second pass's twiddles w_n^(e col), e < N_1, for every column col < n / N_1 that is a
// multiple of `col_step` (k_ragged.hip first_stage: radix R0 = 2^(l mod 4) or 16, 16 points per
// thread, W the two-level product with no shift).
void tw_count_ragged_pass(unsigned log2n, uint32_t col_step, uint64_t *out) {
    Count c;
    const unsigned l1 = log2n / 2;
    const unsigned lo_bits = onefft::two_level_lo_bits(log2n);
    cd *lo = new cd[size_t(1) << lo_bits];
    cd *hi = new cd[size_t(1) << (log2n - lo_bits)];
    onefft::fill_two_level(log2n, lo, hi);
    const uint32_t N = 1u << l1;
    const uint32_t r0_log2 = l1 <= 4 ? l1 : (l1 % 4 ? l1 % 4 : 4);
    const uint32_t R0 = 1u << r0_log2, BF = l1 <= 4 ? 1 : 16 / R0, P = R0 * BF;
    const uint32_t T = N / P, stride = N / R0;
    auto W = [&](uint32_t x) { return onefft::two_level(hi, lo, lo_bits, x); };
    for (uint64_t col = 0; col < (uint64_t(1) << (log2n - l1)); col += col_step) {
        const cd step = W(stride * (uint32_t)col);
        for (uint32_t j = 0; j < T; j++)
            for (uint32_t q = 0; q < BF; q++) {
                cd tw = W((j + T * q) * (uint32_t)col);
                for (uint32_t r = 0; r < R0; r++) {
                    compare(c, tw, (uint64_t)(j + T * q + stride * r) * col, log2n);
                    if (r + 1 < R0) tw = onefft::mul64(tw, step);
                }
            }
    }
    delete[] lo;
    delete[] hi;
    write(c, out);
}

}  // extern "C"
