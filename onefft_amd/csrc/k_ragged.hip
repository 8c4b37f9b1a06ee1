// Ragged batches in one launch: a persistent kernel whose workgroups claim tasks from a device
// queue until it is empty.
//
// The reference transforms one array per call (include/onefft.h:61-77); a caller with many
// arrays of mixed sizes calls it once per array. Here every array of a batch is an item
// (offset, log2 n, direction). A task is up to S / n whole single-pass items of one size n <= S
// (S = kRaggedTaskPoints), one whole two-pass item below the split size (both passes by one
// workgroup, whose barriers order them), or a run of tiles (C = S / N_f columns each) of one pass
// of a larger two-pass item, so a large item is spread over many workgroups (kernels.hpp).
//
// The queue runs in two phases: single-pass, whole and first-pass tasks, then second-pass runs,
// which read what the first passes wrote. Device-scope release and acquire on this chip write
// back and invalidate a whole XCD's L2, so they happen once per workgroup, not per task: when a
// workgroup claims past phase 1 it publishes its phase-1 results (every wave writes back, then one
// atomic adds the number of phase-1 tasks it ran), and before its first phase-2 task it waits
// until all phase-1 tasks are published, then drops its stale cached lines. Only tasks already
// claimed, hence held by running workgroups that will finish them without waiting, are waited for,
// so the launch needs no co-residency of its workgroups. The wait is bounded: past the limit it
// sets an error flag the host reports, and goes on. The host puts the largest items first within
// each phase, so the long runs start early and the queue ends with short ones.
#include "fft_bodies.hpp"

namespace onefft {

namespace {

// LDS of one task: the largest of the single-pass layouts (kRaggedTaskPoints points) and the pass
// layouts (C = kRaggedTaskPoints / N_f columns of pass_span(N_f, C) values).
constexpr uint32_t ragged_lds_values() {
    uint32_t m = kRaggedTaskPoints / 16 * lds_span(16);
    for (int l = 6; l <= 10; l++) {
        uint32_t v = (kRaggedTaskPoints >> l) * pass_span(l, (int)(kRaggedTaskPoints >> l), kPadded);
        if (v > m) m = v;
    }
    return m;
}
constexpr uint32_t kLds = ragged_lds_values();

// One column transform of N = 2^L points (L <= kRaggedTaskLog2) with every stage's shape chosen
// at run time:
// a first stage of radix R0 (N itself below 16 points, else 2^(L mod 4), or 16), then radix-16
// stages through LDS. The kernel thus has a few short code paths instead of one fully unrolled
// body per size, which keeps its register use near that of the standalone kernels.
//
// Element e of the column is src[e * in_step], output k goes to dst[k * out_step]. With `twiddle`
// the loaded element e is multiplied by w_n^((e low) << shift), formed in FP64 (second passes).
struct Column {
    const cf *src;
    cf *dst;
    uint32_t in_step, out_step;
    bool live, swap_in, swap_out, twiddle;
    const cd *lo, *hi;
    uint32_t lo_bits, low, shift;
};

__device__ __forceinline__ cf swap_if(cf v, bool s) { return s ? swap_ri(v) : v; }

// The first stage: BF = P / R0 butterflies of radix R0 per thread (BF = 1 when the column has at
// most 16 points, else 16 / R0). All P loads are issued before any arithmetic, so a thread waits
// for global memory once per column rather than once per butterfly.
template <int R0, int BF>
__device__ __forceinline__ void first_stage(const Column &io, uint32_t N, uint32_t j, bool only, cf *L,
                                            cf (&v)[16]) {
    constexpr uint32_t P = R0 * BF;
    const uint32_t T = N / P, stride = N / R0;
    static_for<0, BF>([&](auto q) {
        static_for<0, R0>([&](auto r) {
            constexpr int qq = decltype(q)::value, rr = decltype(r)::value;
            v[qq * R0 + rr] = io.live ? swap_if(io.src[(j + T * qq + rr * stride) * io.in_step], io.swap_in)
                                      : mk(0.f, 0.f);
        });
    });
    if (io.twiddle) {
        auto W = [&](uint32_t x) { return two_level(io.hi, io.lo, io.lo_bits, x << io.shift); };
        const cd step = W(stride * io.low);
        static_for<0, BF>([&](auto q) {
            constexpr int qq = decltype(q)::value;
            cd tw = W((j + T * qq) * io.low);
            static_for<0, R0>([&](auto r) {
                v[qq * R0 + decltype(r)::value] = mul(v[qq * R0 + decltype(r)::value], round32(tw));
                if constexpr (decltype(r)::value + 1 < R0) tw = mul64(tw, step);
            });
        });
    }
    static_for<0, BF>([&](auto q) {
        constexpr int qq = decltype(q)::value;
        dft<R0>(v + qq * R0);
        const uint32_t b = j + T * qq;
        static_for<0, R0>([&](auto m) {
            constexpr int mm = decltype(m)::value;
            if (only) {
                if (io.live) io.dst[(b + mm * stride) * io.out_step] = swap_if(v[qq * R0 + mm], io.swap_out);
            } else {
                L[lds_pad(b * R0 + mm)] = v[qq * R0 + mm];
            }
        });
    });
}

__device__ void column_fft(const Column &io, uint32_t log2n, const cd *tw64, cf *L, uint32_t j) {
    const uint32_t N = 1u << log2n;
    const uint32_t r0_log2 = log2n <= 4 ? log2n : (log2n % 4 ? log2n % 4 : 4);
    const bool only = log2n <= 4;  // one stage: R0 = N points per thread
    cf v[16];
    switch (r0_log2) {
    case 0: first_stage<1, 1>(io, N, j, only, L, v); break;
    case 1: only ? first_stage<2, 1>(io, N, j, only, L, v) : first_stage<2, 8>(io, N, j, only, L, v); break;
    case 2: only ? first_stage<4, 1>(io, N, j, only, L, v) : first_stage<4, 4>(io, N, j, only, L, v); break;
    case 3: only ? first_stage<8, 1>(io, N, j, only, L, v) : first_stage<8, 2>(io, N, j, only, L, v); break;
    default: first_stage<16, 1>(io, N, j, only, L, v); break;
    }
    if (only) return;
    __syncthreads();
    // radix-16 stages: butterfly j reads j + r N/16, twiddles w_{16 NS}^(r (j mod NS))
    const uint32_t stride = N / 16;
    uint32_t log2ns = r0_log2, off64 = 0;
    for (;;) {
        const uint32_t NS = 1u << log2ns;
        const bool last = log2ns + 4 == log2n;
        static_for<0, 16>([&](auto r) { v[r] = L[lds_pad(j + decltype(r)::value * stride)]; });
        const uint32_t k = j & (NS - 1);
        // w^(r k) by an FP64 recurrence from the stage's base w^k, each rounded once (FP64-base
        // level 3, fft_common.hpp: stage_uses_tw64)
        const cd w1 = tw64[off64 + k];
        cd w = w1;
        static_for<1, 16>([&](auto r) {
            v[r] = mul(v[r], round32(w));
            if constexpr (decltype(r)::value + 1 < 16) w = mul64(w, w1);
        });
        dft<16>(v);
        if (last) {
            if (io.live) static_for<0, 16>([&](auto m) {
                io.dst[(j + decltype(m)::value * stride) * io.out_step] = swap_if(v[m], io.swap_out);
            });
            break;
        }
        __syncthreads();
        const uint32_t base = ((j >> log2ns) << (log2ns + 4)) + k;
        static_for<0, 16>([&](auto m) { L[lds_pad(base + decltype(m)::value * NS)] = v[m]; });
        __syncthreads();
        off64 += NS;
        log2ns += 4;
    }
}

// up to kRaggedTaskPoints / n whole transforms of one size n <= kRaggedTaskPoints, one column per
// transform
__device__ void single_task(const RaggedArgs &a, const RaggedTask &tk, uint32_t log2n, bool inv, cf *lds,
                            uint32_t t) {
    const uint32_t N = 1u << log2n;
    const uint32_t T = N < 16 ? 1 : N / 16;  // threads per transform
    const uint32_t c = t / T;
    Column io;
    io.live = c < tk.arg;
    const uint64_t off = io.live ? a.item_off[tk.item + c] : 0;
    io.src = a.in + off;
    io.dst = a.out + off;
    io.in_step = io.out_step = 1;
    io.swap_in = io.swap_out = inv;
    io.twiddle = false;
    column_fft(io, log2n, a.tw64 + tk.tw64, lds + c * lds_span(N), t % T);
}

// one tile (C = kRaggedTaskPoints / N_f columns) of one pass of a two-pass item (index scheme of
// k_pass.hip)
__device__ void pass_tile(const RaggedArgs &a, const RaggedSize &z, const cf *in, cf *out, uint32_t tw64, bool first,
                          uint32_t tile, bool inv, cf *lds, uint32_t t) {
    const uint32_t lf = first ? z.l2 : z.l1;     // this pass's sub-transform
    const uint32_t C = kRaggedTaskPoints >> lf;  // columns per tile
    const uint32_t c = t % C, j = t / C;
    const uint32_t col = tile * C + c;
    Column io;
    io.live = true;
    io.twiddle = !first;
    io.lo = a.tw64 + z.lo;
    io.hi = a.tw64 + z.hi;
    io.lo_bits = z.lo_bits;
    io.shift = 0;
    if (first) {
        // factor 2 (N_2 points) first: column j_1 of the input (stride N_1) -> row at N_2 j_1
        io.src = in + col;
        io.in_step = 1u << z.l1;
        io.dst = out + ((uint64_t)col << z.l2);
        io.out_step = 1;
        io.swap_in = inv;
        io.swap_out = false;
        io.low = 0;
    } else {
        // factor 1 (N_1 points) in place: column k_2 at stride N_2, twiddles w_n^(j_1 k_2)
        io.src = out + col;
        io.dst = out + col;
        io.in_step = io.out_step = 1u << z.l2;
        io.swap_in = false;
        io.swap_out = inv;
        io.low = col;
    }
    column_fft(io, lf, a.tw64 + tw64, lds + c * pass_span((int)lf, (int)C, kPadded), j);
}

// a run of tiles of one pass of a two-pass item
__device__ void pass_tiles(const RaggedArgs &a, const RaggedTask &tk, uint32_t log2n, bool inv, bool first, cf *lds,
                           uint32_t t) {
    const RaggedSize &z = a.size[log2n];
    const uint64_t off = a.item_off[tk.item];
    const uint32_t run = ((tk.word >> 8) & 255u) + 1;
    // The table offset goes through an empty asm statement each tile, so the compiler treats it
    // as a new value and reloads the twiddle bases per tile instead of holding every tile's in
    // registers across the loop (which would cost occupancy).
    for (uint32_t i = 0; i < run; i++) {
        uint32_t tw64 = first ? z.tw64_2 : z.tw64_1;
        asm volatile("" : "+v"(tw64));
        pass_tile(a, z, a.in + off, a.out + off, tw64, first, tk.arg + i, inv, lds, t);
        __syncthreads();  // LDS reuse
    }
}

}  // namespace

__global__ __launch_bounds__(kRaggedWG) void fft_ragged_kernel(RaggedArgs a) {
    __shared__ cf lds[kLds];
    __shared__ uint32_t claimed;
    const uint32_t t = threadIdx.x;
    uint32_t ntasks = a.ntasks, nphase1 = a.nphase1;
    if (a.counts) {  // a queue built on the device: its counts are in memory (uniform loads)
        ntasks = __builtin_amdgcn_readfirstlane(a.counts[0]);
        nphase1 = __builtin_amdgcn_readfirstlane(a.counts[1]);
    }
    uint32_t mine = 0;       // phase-1 tasks this workgroup ran
    bool published = false;  // its phase-1 results are published
    for (;;) {
        if (t == 0) claimed = __hip_atomic_fetch_add(a.head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t id = __builtin_amdgcn_readfirstlane(claimed);  // uniform: task fields go to SGPRs
        __syncthreads();  // everyone has read `claimed` before it is overwritten
        if (id >= nphase1 && !published) {
            // phase 1 is claimed out: every wave writes its stores back, then one atomic publishes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (t == 0) __hip_atomic_fetch_add(a.head + 1, mine, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            published = true;
            if (id < ntasks) {
                // a second-pass task: wait until every phase-1 task is published
                if (t == 0) {
                    for (uint32_t spins = 0;
                         __hip_atomic_load(a.head + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nphase1;
                         spins++) {
                        if (spins >= a.spin_limit) {
                            __hip_atomic_store(a.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(8);
                    }
                    // the thread that read the count acquires the publishers' releases ...
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
                __syncthreads();
                // ... and, after the barrier, every wave drops its stale cached lines
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        if (id >= ntasks) break;
        const RaggedTask tk = a.tasks[id];
        const uint32_t kind = tk.word & 3u, log2n = (tk.word >> 2) & 31u;
        const bool inv = (tk.word >> 7) & 1u;
        if (kind == kRaggedSingle) {
            single_task(a, tk, log2n, inv, lds, t);
        } else if (kind == kRaggedWhole) {
            pass_tiles(a, tk, log2n, inv, true, lds, t);
            // the first pass's global stores, made by this workgroup, are visible to it here
            pass_tiles(a, tk, log2n, inv, false, lds, t);
        } else {
            pass_tiles(a, tk, log2n, inv, kind == kRaggedFirst, lds, t);
        }
        if (id < nphase1) mine++;
        __syncthreads();  // LDS is free for the next task
    }
}

__global__ __launch_bounds__(64) void fft_ragged_reset(uint32_t *head, uint32_t *error) {
    if (threadIdx.x < 2) head[threadIdx.x] = 0;
    if (threadIdx.x == 2) *error = 0;
}

uint32_t ragged_resident_grid() {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fft_ragged_kernel, kRaggedWG, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1024;
    }
    if (per < 1) per = 1;
    return (uint32_t)(cus * per);
}

hipError_t launch_ragged(const RaggedArgs &a, uint32_t grid, hipStream_t stream) {
    hipLaunchKernelGGL(fft_ragged_reset, dim3(1), dim3(64), 0, stream, a.head, a.error);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.ntasks == 0) return hipSuccess;
    if (grid == 0) grid = ragged_resident_grid();
    if (grid > a.ntasks) grid = a.ntasks;
    return launch_ragged_kernel(a, grid, stream);
}

hipError_t launch_ragged_kernel(const RaggedArgs &a, uint32_t grid, hipStream_t stream) {
    if (grid == 0) grid = ragged_resident_grid();
    hipLaunchKernelGGL(fft_ragged_kernel, dim3(grid), dim3(kRaggedWG), 0, stream, a);
    return hipGetLastError();
}

}  // namespace onefft
