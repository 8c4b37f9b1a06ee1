/*
 * CPU oracle for OneFFT's power-of-two, complex-to-complex, single-precision, out-of-place FFT.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker, never the thing measured or shipped: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it. The product
 * library (onefft_amd/) never links or calls it.
 *
 * What it restates, with the reference file each function follows (paths relative to the
 * reference repository):
 *   oracle_root_double / oracle_root   src/onefft.c:112-141  (octant-folded FP64 twiddle,
 *                                                             rounded once to FP32)
 *   oracle_bit_reverse                 src/onefft.c:143-150
 *   oracle_fft_f32                     src/onefft.c:152-230  (n = 1, 2, 4 closed forms, portable
 *                                                             radix-2 with the re/im exchange
 *                                                             for the inverse)
 *   oracle_fft_f64                     tests/test_onefft.c:30-55 (FP64 radix-2 reference), with
 *                                      the phase reduced in integers as src/onefft.c:116 does
 *   oracle_dft_f64                     include/onefft.h:30-31 (the definition, O(n^2))
 *   oracle_xorshift_fill               tests/test_onefft.c:57-61, bench/bench.c:94-98
 *   oracle_rel_rms                     tests/test_onefft.c:63-71, bench/bench.c:142-149
 * Input generators for the workloads (SplitMix64 streams) have no reference counterpart: the
 * reference fixes no generator for them, so they are defined here once and used by every test.
 *
 * How each function is pinned (checked by tests/test_oracle.py):
 *   oracle_root*      bitwise against the reference's own onefft_root, built from its sources into
 *                     oracle/_ref; exact values at multiples of an eighth turn; |w| = 1.
 *   oracle_fft_f32    bitwise against the reference built with ONEFFT_NO_SIMD (its portable
 *                     path); within tolerance of the reference's SIMD path and of FP64.
 *   oracle_fft_f64    against oracle_dft_f64 (brute force) and numpy.fft in float64; known
 *                     answers (impulse, constant, tone); Parseval; linearity.
 *   oracle_dft_f64    against numpy.fft on small sizes and against closed forms.
 *   generators        against committed golden vectors and their closed-form statistics.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* exp(-2 pi i k / 2^log2size) in FP64. The phase is k / 2^log2size of a turn, kept exactly as a
 * 0.64 fixed-point fraction; it is folded into the first octant before sin/cos so that the table
 * is exactly symmetric and exact at multiples of an eighth turn (src/onefft.c:112-135). */
void oracle_root_double(uint64_t k, unsigned log2size, double *re, double *im) {
    uint64_t turn = log2size ? k << (64 - log2size) : 0;
    unsigned quarter = (unsigned)(turn >> 62);
    uint64_t within = turn & ((UINT64_C(1) << 62) - 1);
    const double radians_per_unit = 6.283185307179586476925286766559 / 18446744073709551616.0;
    double c, s;
    if (within > (UINT64_C(1) << 61)) {
        /* second octant of the quarter: use the complementary angle */
        double a = (double)((UINT64_C(1) << 62) - within) * radians_per_unit;
        c = sin(a);
        s = cos(a);
    } else {
        double a = (double)within * radians_per_unit;
        c = cos(a);
        s = sin(a);
    }
    /* (c - i s) rotated by (-i)^quarter */
    if (quarter == 0) {
        *re = c, *im = -s;
    } else if (quarter == 1) {
        *re = -s, *im = -c;
    } else if (quarter == 2) {
        *re = -c, *im = s;
    } else {
        *re = s, *im = c;
    }
}

void oracle_root(uint64_t k, unsigned log2size, float *re, float *im) {
    double c, s;
    oracle_root_double(k, log2size, &c, &s);
    *re = (float)c;
    *im = (float)s;
}

size_t oracle_bit_reverse(size_t value, unsigned bits) {
    size_t r = 0;
    for (unsigned b = 0; b < bits; b++) {
        r = (r << 1) | (value & 1);
        value >>= 1;
    }
    return r;
}

static int oracle_log2(size_t n, unsigned *log2n) {
    if (n == 0 || (n & (n - 1)) != 0) return -1;
    unsigned l = 0;
    while (((size_t)1 << l) < n) l++;
    if (l > 30) return -1;
    *log2n = l;
    return 0;
}

/* The reference's portable single-precision transform: closed forms for n <= 4 and a radix-2
 * decimation-in-time transform with an n/2-entry twiddle table otherwise. The inverse exchanges
 * real and imaginary parts on the way in and out (src/onefft.c:86-88, 152-230). */
int oracle_fft_f32(const float *in, float *out, size_t n, int inverse) {
    unsigned log2n;
    if (oracle_log2(n, &log2n)) return -1;
    int inv = inverse ? 1 : 0;
    if (n == 1) {
        out[0] = in[0], out[1] = in[1];
        return 0;
    }
    if (n == 2) {
        float ar = in[0], ai = in[1], br = in[2], bi = in[3];
        out[0] = ar + br, out[1] = ai + bi;
        out[2] = ar - br, out[3] = ai - bi;
        return 0;
    }
    if (n == 4) {
        float sr = in[0] + in[4], si = in[1] + in[5];
        float dr0 = in[0] - in[4], di0 = in[1] - in[5];
        float tr = in[2] + in[6], ti = in[3] + in[7];
        float ur = in[2] - in[6], ui = in[3] - in[7];
        if (inv) ur = -ur, ui = -ui;
        out[0] = sr + tr, out[1] = si + ti;
        out[2] = dr0 + ui, out[3] = di0 - ur;
        out[4] = sr - tr, out[5] = si - ti;
        out[6] = dr0 - ui, out[7] = di0 + ur;
        return 0;
    }
    float *tw = (float *)malloc(n * sizeof(float)); /* n/2 complex */
    if (!tw) return -1;
    for (size_t k = 0; k < n / 2; k++) oracle_root(k, log2n, &tw[2 * k], &tw[2 * k + 1]);
    for (size_t i = 0; i < n; i++) {
        size_t j = oracle_bit_reverse(i, log2n);
        out[2 * j + inv] = in[2 * i];
        out[2 * j + 1 - inv] = in[2 * i + 1];
    }
    for (size_t half = 1; half < n; half *= 2) {
        size_t step = n / (2 * half);
        for (size_t start = 0; start < n; start += 2 * half) {
            for (size_t k = 0; k < half; k++) {
                float wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                float *a = out + 2 * (start + k);
                float *b = a + 2 * half;
                float pr = b[0] * wr - b[1] * wi;
                float pi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - pr, b[1] = a[1] - pi;
                a[0] += pr, a[1] += pi;
            }
        }
    }
    if (inv) {
        for (size_t i = 0; i < n; i++) {
            float t = out[2 * i];
            out[2 * i] = out[2 * i + 1];
            out[2 * i + 1] = t;
        }
    }
    free(tw);
    return 0;
}

/* FP64 reference: in-order radix-2 after a bit-reversal permutation, with every twiddle taken
 * from oracle_root_double (integer phase reduction), conjugated for the inverse. Same structure
 * as the reference's test oracle (tests/test_onefft.c:30-55). in and out may be the same. */
int oracle_fft_f64(const double *in, double *out, size_t n, int inverse) {
    unsigned log2n;
    if (oracle_log2(n, &log2n)) return -1;
    if (in != out) memmove(out, in, 2 * n * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        size_t j = oracle_bit_reverse(i, log2n);
        if (i < j) {
            double tr = out[2 * i], ti = out[2 * i + 1];
            out[2 * i] = out[2 * j], out[2 * i + 1] = out[2 * j + 1];
            out[2 * j] = tr, out[2 * j + 1] = ti;
        }
    }
    double *tw = (double *)malloc((n > 1 ? n : 2) * sizeof(double));
    if (!tw) return -1;
    for (size_t k = 0; k < n / 2; k++) {
        oracle_root_double(k, log2n, &tw[2 * k], &tw[2 * k + 1]);
        if (inverse) tw[2 * k + 1] = -tw[2 * k + 1];
    }
    for (size_t half = 1; half < n; half *= 2) {
        size_t step = n / (2 * half);
        for (size_t start = 0; start < n; start += 2 * half) {
            for (size_t k = 0; k < half; k++) {
                double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                double *a = out + 2 * (start + k);
                double *b = a + 2 * half;
                double pr = b[0] * wr - b[1] * wi;
                double pi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - pr, b[1] = a[1] - pi;
                a[0] += pr, a[1] += pi;
            }
        }
    }
    free(tw);
    return 0;
}

/* The definition, O(n^2), in FP64: out[k] = sum_j in[j] exp(-+2 pi i (j k mod n) / n). */
int oracle_dft_f64(const double *in, double *out, size_t n, int inverse) {
    unsigned log2n;
    if (oracle_log2(n, &log2n)) return -1;
    for (size_t k = 0; k < n; k++) {
        double sr = 0, si = 0;
        for (size_t j = 0; j < n; j++) {
            double wr, wi;
            oracle_root_double((uint64_t)((j * k) & (n - 1)), log2n, &wr, &wi);
            if (inverse) wi = -wi;
            sr += in[2 * j] * wr - in[2 * j + 1] * wi;
            si += in[2 * j] * wi + in[2 * j + 1] * wr;
        }
        out[2 * k] = sr, out[2 * k + 1] = si;
    }
    return 0;
}

/* Selected bins of the FP64 transform, O(n) each: for checking single bins of sizes whose full
 * FP64 transform is too slow for a test. */
int oracle_dft_bins_f64(const float *in, size_t n, int inverse, const uint64_t *bins, size_t count,
                        double *out) {
    unsigned log2n;
    if (oracle_log2(n, &log2n)) return -1;
    for (size_t b = 0; b < count; b++) {
        uint64_t k = bins[b];
        double sr = 0, si = 0;
        for (size_t j = 0; j < n; j++) {
            double wr, wi;
            oracle_root_double((uint64_t)(((uint64_t)j * k) & (n - 1)), log2n, &wr, &wi);
            if (inverse) wi = -wi;
            double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * wr - xi * wi;
            si += xr * wi + xi * wr;
        }
        out[2 * b] = sr, out[2 * b + 1] = si;
    }
    return 0;
}

/* The reference tests' generator: xorshift32, uniform in [-1, 1). */
void oracle_xorshift_fill(uint32_t *state, float *out, size_t count) {
    uint32_t s = *state;
    for (size_t i = 0; i < count; i++) {
        s ^= s << 13, s ^= s >> 17, s ^= s << 5;
        out[i] = (float)(int32_t)s * (1.0f / 2147483648.0f);
    }
    *state = s;
}

/* SplitMix64 stream (Steele, Lea, Flood: "Fast splittable pseudorandom number generators",
 * OOPSLA 2014). */
static uint64_t splitmix_next(uint64_t *s) {
    uint64_t z = (*s += UINT64_C(0x9E3779B97F4A7C15));
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

/* count floats uniform in [-1, 1): 24 random bits times 2^-23 minus one, exact in FP32.
 * Draw order: re then im per element, element-major. */
void oracle_uniform_fill(uint64_t seed, float *out, size_t count) {
    uint64_t s = seed;
    for (size_t i = 0; i < count; i++) {
        uint64_t u = splitmix_next(&s);
        out[i] = (float)((double)(u >> 40) * (1.0 / 8388608.0) - 1.0);
    }
}

/* count floats standard normal: Box-Muller in FP64 from two 53-bit uniforms, (0,1] and [0,1),
 * each result rounded once to FP32. Pairs fill (re, im). */
void oracle_normal_fill(uint64_t seed, float *out, size_t count) {
    uint64_t s = seed;
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t i = 0; i < count; i += 2) {
        double u1 = (double)((splitmix_next(&s) >> 11) + 1) * (1.0 / 9007199254740992.0);
        double u2 = (double)(splitmix_next(&s) >> 11) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        out[i] = (float)(r * cos(two_pi * u2));
        if (i + 1 < count) out[i + 1] = (float)(r * sin(two_pi * u2));
    }
}

/* x[j] += exp(+2 pi i k0 j / n) for j < n, phase (k0 j mod n) / n in integers, added in FP64,
 * rounded once: the unit tone of the large-transform workload. */
int oracle_add_tone(float *x, size_t n, uint64_t k0) {
    unsigned log2n;
    if (oracle_log2(n, &log2n)) return -1;
    for (size_t j = 0; j < n; j++) {
        double c, s;
        oracle_root_double((uint64_t)((k0 * (uint64_t)j) & (n - 1)), log2n, &c, &s);
        x[2 * j] = (float)((double)x[2 * j] + c);
        x[2 * j + 1] = (float)((double)x[2 * j + 1] - s);
    }
    return 0;
}

/* Relative RMS difference: sqrt(sum (got - want)^2 / sum want^2); the plain RMS if want is 0. */
double oracle_rel_rms(const float *got, const double *want, size_t count) {
    double e = 0, w = 0;
    for (size_t i = 0; i < count; i++) {
        double d = (double)got[i] - want[i];
        e += d * d;
        w += want[i] * want[i];
    }
    return w > 0 ? sqrt(e / w) : sqrt(e);
}

/* Same between two float vectors, the second taken as exact. */
double oracle_rel_rms_f32(const float *got, const float *want, size_t count) {
    double e = 0, w = 0;
    for (size_t i = 0; i < count; i++) {
        double d = (double)got[i] - (double)want[i];
        e += d * d;
        w += (double)want[i] * (double)want[i];
    }
    return w > 0 ? sqrt(e / w) : sqrt(e);
}

/* Batched helpers so Python tests can run many transforms with one call. */
int oracle_fft_f32_batch(const float *in, float *out, size_t n, size_t batch, size_t dist, int inverse) {
    for (size_t b = 0; b < batch; b++) {
        if (oracle_fft_f32(in + 2 * b * dist, out + 2 * b * dist, n, inverse)) return -1;
    }
    return 0;
}

/* FP64 transform of FP32 input, batched; out holds 2n doubles per transform (dist in complex). */
int oracle_fft_f64_from_f32_batch(const float *in, double *out, size_t n, size_t batch, size_t dist, int inverse) {
    for (size_t b = 0; b < batch; b++) {
        double *o = out + 2 * b * n;
        const float *x = in + 2 * b * dist;
        for (size_t i = 0; i < 2 * n; i++) o[i] = (double)x[i];
        if (oracle_fft_f64(o, o, n, inverse)) return -1;
    }
    return 0;
}
