// Device building blocks shared by every OneFFT gfx950 kernel: the complex type, constant
// twiddles, register DFTs of radix 2..16, the Stockham stage schedule and the LDS layout.
//
// Conventions (the reference's, include/onefft.h:26-34): data are interleaved complex float
// (re, im), forward is exp(-2 pi i jk/n), nothing is normalised. The kernels only implement the
// forward direction; the inverse exchanges re and im on the way in and out, which is the
// reference's own inverse (src/onefft.c:86-88, docs/design.md:137-144), or, in the single-pass
// kernels, conjugates on the way in and out (the same transform: swap(F(swap x)) =
// conj(F(conj x))).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "plan_tables.hpp"

namespace onefft {

ONEFFT_HD cf add(cf a, cf b) { return mk(a.r + b.r, a.i + b.i); }
ONEFFT_HD cf sub(cf a, cf b) { return mk(a.r - b.r, a.i - b.i); }
// a * w
ONEFFT_HD cf mul(cf a, cf w) { return mk(a.r * w.r - a.i * w.i, a.r * w.i + a.i * w.r); }
// a * (-i)
ONEFFT_HD cf mul_mi(cf a) { return mk(a.i, -a.r); }
// exchange re and im (the inverse transform's input and output mapping)
ONEFFT_HD cf swap_ri(cf a) { return mk(a.i, a.r); }
template <bool SWAP>
ONEFFT_HD cf maybe_swap(cf a) {
    if constexpr (SWAP) return swap_ri(a);
    else return a;
}
template <bool CONJ>
ONEFFT_HD cf maybe_conj(cf a) {
    if constexpr (CONJ) return mk(a.r, -a.i);
    else return a;
}

// ---- compile-time loops -------------------------------------------------------------------

template <int I>
using ic = std::integral_constant<int, I>;

template <int B, int E, typename F>
ONEFFT_HD void static_for(F &&f) {
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}

// ---- constant twiddles w_R^E = exp(-2 pi i E / R), R a power of two up to 32 ---------------

// cos(2 pi e / 32) for e = 0..8
constexpr double kCos32[9] = {1.0,
                              0.98078528040323044912618223613424,
                              0.92387953251128675612818318939679,
                              0.83146961230254523707878837761791,
                              0.70710678118654752440084436210485,
                              0.55557023301960222474283081394853,
                              0.38268343236508977172845998403040,
                              0.19509032201612826784828486847702,
                              0.0};

constexpr double cos32(int e) {
    e &= 31;
    return e <= 8 ? kCos32[e] : e <= 16 ? -kCos32[16 - e] : e <= 24 ? -kCos32[e - 16] : kCos32[32 - e];
}
constexpr double sin32(int e) { return cos32(e - 8); }

// a * w_R^E with every trivial or eighth-turn case written out.
template <int R, int E>
ONEFFT_HD cf twc(cf a) {
    static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "radix must be a power of two up to 32");
    constexpr int e = ((E % R) + R) % R * (32 / R);  // in 32nds of a turn
    constexpr float h = 0.70710678118654752440f;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (e == 8) {
        return mk(a.i, -a.r);  // -i
    } else if constexpr (e == 16) {
        return mk(-a.r, -a.i);
    } else if constexpr (e == 24) {
        return mk(-a.i, a.r);  // +i
    } else if constexpr (e == 4) {
        return mk((a.r + a.i) * h, (a.i - a.r) * h);  // (1 - i)/sqrt2
    } else if constexpr (e == 12) {
        return mk((a.i - a.r) * h, -(a.r + a.i) * h);  // (-1 - i)/sqrt2
    } else if constexpr (e == 20) {
        return mk(-(a.r + a.i) * h, (a.r - a.i) * h);  // (-1 + i)/sqrt2
    } else if constexpr (e == 28) {
        return mk((a.r - a.i) * h, (a.r + a.i) * h);  // (1 + i)/sqrt2
    } else {
        constexpr float c = (float)cos32(e), s = (float)sin32(e);
        // (a.r + i a.i)(c - i s)
        return mk(a.r * c + a.i * s, a.i * c - a.r * s);
    }
}

// ---- register DFTs (forward, natural order in and out) -------------------------------------

template <int R>
ONEFFT_HD void dft(cf *v);

template <>
ONEFFT_HD void dft<1>(cf *) {}

template <>
ONEFFT_HD void dft<2>(cf *v) {
    cf a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
}

template <>
ONEFFT_HD void dft<4>(cf *v) {
    cf s0 = add(v[0], v[2]), d0 = sub(v[0], v[2]);
    cf s1 = add(v[1], v[3]), d1 = mul_mi(sub(v[1], v[3]));
    v[0] = add(s0, s1);
    v[2] = sub(s0, s1);
    v[1] = add(d0, d1);
    v[3] = sub(d0, d1);
}

// R = A * B by one decimation-in-time split: x[a + A b] -> A DFTs of size B over b, twiddles
// w_R^(a k), then B DFTs of size A over a; output index k + B * m.
template <int R>
ONEFFT_HD void dft(cf *v) {
    constexpr int A = (R == 8) ? 2 : 4;
    constexpr int B = R / A;
    cf y[A][B];
    static_for<0, A>([&](auto a) {
        static_for<0, B>([&](auto b) { y[a][b] = v[a + A * b]; });
        dft<B>(y[a]);
        static_for<1, B>([&](auto k) { y[a][k] = twc<R, decltype(a)::value * decltype(k)::value>(y[a][k]); });
    });
    static_for<0, B>([&](auto k) {
        cf z[A];
        static_for<0, A>([&](auto a) { z[a] = y[a][k]; });
        dft<A>(z);
        static_for<0, A>([&](auto m) { v[k + B * m] = z[m]; });
    });
}

// ---- LDS layout ---------------------------------------------------------------------------
//
// Three layouts of a transform's values in LDS, chosen per kernel variant:
//   kPadded   one complex of padding after every 16: stage accesses that stride by powers of two
//             spread over the banks, at 1/16 extra LDS;
//   kSwizzled no padding; the low 4 index bits are XORed with bits 4..7 and 8..11, a bijection
//             that spreads the same accesses over the banks and keeps a 4096-point transform at
//             exactly 32 KiB (one more workgroup per CU);
//   kHalves   the padded index over floats: a stage exchange moves the real parts, then the
//             imaginary parts, through one buffer of n + n/16 floats, half the padded layout's
//             LDS (twice the workgroups per CU) for twice the barriers.
// lds_extent is in complex values (kHalves: enough of them to hold the floats).
enum LdsLayout : int { kPadded = 0, kSwizzled = 1, kHalves = 2 };

template <int LAYOUT = kPadded>
__host__ __device__ constexpr uint32_t lds_index(uint32_t i) {
    if constexpr (LAYOUT == kSwizzled) return i ^ (((i >> 4) ^ (i >> 8)) & 15u);
    else return i + (i >> 4);
}
template <int LAYOUT = kPadded>
__host__ __device__ constexpr uint32_t lds_extent(uint32_t n) {
    return LAYOUT == kSwizzled ? n : LAYOUT == kHalves ? (n + (n >> 4) + 1) / 2 : n + (n >> 4);
}
// lds_index(b + c) with the part that depends on b computed once: for a base b computed at run
// time and a compile-time step c, the padded index is lds_index(b) + c + (c >> 4) whenever the low
// four bits of b and c do not carry ((b & 15) + (c & 15) < 16), and the c part folds into the
// LDS instruction's offset. Every Stockham access here satisfies it: a read takes b < stride and c
// a multiple of the stride (a power of two); a write takes b = q Ns R + k, k < Ns, and c a multiple
// of Ns, and Ns R >= 16 or every sum stays below Ns R. The swizzled layout does not fold.
template <int LAYOUT = kPadded>
__host__ __device__ constexpr uint32_t lds_at(uint32_t b, uint32_t c) {
    if constexpr (LAYOUT == kSwizzled) return lds_index<kSwizzled>(b + c);
    else return lds_index<kPadded>(b) + c + (c >> 4);  // kPadded, and kHalves in float units
}
// the padded layout, used by the kernels without a layout choice
__host__ __device__ constexpr uint32_t lds_pad(uint32_t i) { return lds_index<kPadded>(i); }
__host__ __device__ constexpr uint32_t lds_span(uint32_t n) { return lds_extent<kPadded>(n); }

// ---- workgroup order --------------------------------------------------------------------------
//
// The GPU hands the workgroups of a launch to its 8 XCDs in turn (workgroup i to XCD i mod 8), and
// each XCD has its own L2. Renumbering workgroup b of `count` so that XCD x runs one contiguous
// range of the work makes neighbouring tiles, which share cache lines, DRAM pages and page-table
// entries, meet in the same L2. A bijection on [0, count) for any count; on a device with another
// number of XCDs it only changes the order.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t count) {
    const uint32_t q = count / kXcds, r = count % kXcds, x = b % kXcds, i = b / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}
// The same, with each XCD's range cut into `strips` equal parts that the XCD runs side by side
// (its k-th workgroup takes the (k / strips)-th tile of strip k mod strips), so that the tiles in
// flight on one XCD lie in several parts of the range; strips 1, or a range it does not divide,
// gives xcd_contiguous.
__device__ __forceinline__ uint32_t xcd_strips(uint32_t b, uint32_t count, uint32_t strips) {
    const uint32_t q = count / kXcds, r = count % kXcds, x = b % kXcds;
    uint32_t i = b / kXcds;
    const uint32_t len = q + (x < r ? 1 : 0);
    if (strips > 1 && len % strips == 0) i = (i % strips) * (len / strips) + i / strips;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// ---- global access --------------------------------------------------------------------------
//
// Streamed data (read once, written once) can use nontemporal loads and stores, which keep it
// from displacing the twiddle tables in the caches.
typedef float f2v __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ cf gload(const cf *p) {
    if constexpr (NT) {
        f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v *>(p));
        return mk(v.x, v.y);
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ void gstore(cf *p, cf c) {
    if constexpr (NT) {
        f2v v = {c.r, c.i};
        __builtin_nontemporal_store(v, reinterpret_cast<f2v *>(p));
    } else {
        *p = c;
    }
}


// One transform's global data, addressed as base + (lane + c) complex values with a per-lane
// part `lane` and a compile-time part `c`.
//
// PtrIO: per-lane 64-bit addresses (transforms of different lanes may lie anywhere).
// BufIO: the transform's base, uniform over the workgroup, in a buffer descriptor held in scalar
//        registers; the byte offset splits into one per-lane register (lane * 8) and a scalar
//        offset (c * 8), so an access costs no vector address arithmetic. Offsets are 32-bit:
//        the transform must lie within 4 GiB of its base, which holds for one transform.
template <bool NT>
struct PtrIO {
    const cf *src;
    cf *dst;
    __device__ __forceinline__ cf load(uint32_t lane, uint32_t c) const { return gload<NT>(src + lane + c); }
    __device__ __forceinline__ void store(uint32_t lane, uint32_t c, cf v) const { gstore<NT>(dst + lane + c, v); }
};

// word 3 of a gfx9-family buffer descriptor: 32-bit data format, no swizzle, no stride
constexpr int kBufferWord3 = 0x00020000;

template <bool NT>
struct BufIO {
    static constexpr int kAux = NT ? 2 : 0;  // the nt cache-policy bit
    __amdgpu_buffer_rsrc_t src, dst;
    __device__ __forceinline__ BufIO(const cf *s, cf *d)
        : src(__builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(s), 0, -1, kBufferWord3)),
          dst(__builtin_amdgcn_make_buffer_rsrc(d, 0, -1, kBufferWord3)) {}
    // the builtins move the 8 bytes as a GCC-style vector of two 32-bit words
    typedef unsigned Word __attribute__((vector_size(8)));
    __device__ __forceinline__ cf load(uint32_t lane, uint32_t c) const {
        const f2v v = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(src, (int)(lane * 8u),
                                                                                   (int)(c * 8u), kAux));
        return mk(v.x, v.y);
    }
    __device__ __forceinline__ void store(uint32_t lane, uint32_t c, cf v) const {
        const f2v w = {v.r, v.i};
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(Word, w), dst, (int)(lane * 8u), (int)(c * 8u), kAux);
    }
};

}  // namespace onefft
