// Ragged executions that allocate nothing: onefft_ragged_execute_items. No counterpart in the
// reference (whose calls transform one array each, include/onefft.h:61-77); it is for callers that
// cannot allocate device memory, such as code under HIP graph capture and tight serving loops.
//
// The item descriptors are in device memory. A planner kernel (one workgroup) validates them and
// builds, in a file-scope device array (`scratch`, kScratchBytes per GPU, held until the process
// ends), what onefft_ragged_plan_create builds on the host: the item offsets in task order, the
// task queue (phase 1: whole two-pass items, first-pass tile runs, single-pass groups; phase 2:
// second-pass tile runs; largest sizes first) and its zeroed counters. The persistent ragged
// kernel (k_ragged.hip) then runs that queue, reading the task counts from `scratch`. The FP64
// twiddle tables of every size 2^0..2^20 are uploaded into `scratch` once per device, on the first
// call there, which must therefore not be under graph capture; every later call is two kernel
// launches and no other HIP call that allocates or copies.
//
// Every call on a GPU shares that GPU's `scratch`: calls that overlap on one GPU (on different
// streams, or from different threads without ordering) race and must not be issued. A call's
// results match onefft_ragged_execute's on the same items bit for bit (each item's arithmetic does
// not depend on how items are grouped into tasks).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "../../include/onefft.h"
#include "fft_common.hpp"
#include "kernels.hpp"
#include "plan_tables.hpp"
#include "runtime.hpp"

namespace onefft {
namespace {

constexpr uint32_t kStaticMaxItems = 1u << 17;  // item descriptors per call
constexpr uint32_t kStaticMaxTasks = 1u << 18;  // tasks per call (the planner's status 3 above)
constexpr int kMaxDevices = 16;
constexpr int kSizes = kRaggedMaxLog2 + 1;

// Offsets (in FP64 complex values) of every size's tables in the table blob: the stage bases of
// the single-pass sizes and of the two-pass sizes' sub-transforms (2^0..2^13), and the two-level
// tables of w_n for the two-pass sizes (2^14..2^20).
struct TableLayout {
    uint32_t stage64[kSizes] = {};
    uint32_t lo[kSizes] = {}, hi[kSizes] = {};
    uint32_t total = 0;
};
constexpr TableLayout table_layout() {
    TableLayout t;
    uint32_t at = 1;  // entry 0 unused, so no size has offset 0 by accident
    for (int l = 0; l <= kRaggedSingleMaxLog2; l++) {
        t.stage64[l] = at;
        at += stage_tw64_size(l, batch_log2p(l)) + 1;
    }
    for (int l = kRaggedSingleMaxLog2 + 1; l <= kRaggedMaxLog2; l++) {
        const uint32_t lo_bits = (uint32_t)(l + 1) / 2;
        t.lo[l] = at;
        at += 1u << lo_bits;
        t.hi[l] = at;
        at += 1u << (l - lo_bits);
    }
    t.total = at;
    return t;
}
constexpr TableLayout kTables = table_layout();

// `scratch` layout, in bytes: queue header (u32 words: [0] next task, [1] phase-1 tasks published,
// [4] task count, [5] phase-1 task count), tasks, item offsets, FP64 tables.
constexpr size_t kOffTasks = 256;
constexpr size_t kOffItems = kOffTasks + size_t(kStaticMaxTasks) * sizeof(RaggedTask);
constexpr size_t kOffTables = kOffItems + size_t(kStaticMaxItems) * sizeof(uint64_t);
constexpr size_t kScratchBytes = (kOffTables + size_t(kTables.total) * sizeof(cd) + 255) / 256 * 256;
static_assert(kOffItems % 256 == 0 && kOffTables % 256 == 0, "256-byte aligned pieces");

// The scratch buffer itself. In host code this name is only a handle for hipGetSymbolAddress: the
// device address of each GPU's copy is looked up once per device (device_scratch) and only that
// address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceState {
    std::atomic<bool> ready{false};
    unsigned char *ptr = nullptr;  // this device's `scratch`
    uint32_t grid = 0;             // resident workgroups of the persistent kernel
};
DeviceState g_state[kMaxDevices];
std::mutex g_setup;
cd g_table_image[kTables.total];  // host image of the tables (static storage: no heap)
bool g_image_built = false;

// This device's scratch address, set up on the first call on the device (nullptr on failure).
unsigned char *device_scratch(int device, uint32_t *grid) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the allocation-free path's limit (16)");
        return nullptr;
    }
    DeviceState &st = g_state[device];
    if (!st.ready.load(std::memory_order_acquire)) {
        std::lock_guard<std::mutex> lk(g_setup);
        if (!st.ready.load(std::memory_order_relaxed)) {
            if (!g_image_built) {
                for (int l = 0; l <= kRaggedSingleMaxLog2; l++)
                    fill_stage_table64(l, batch_log2p(l), g_table_image + kTables.stage64[l]);
                for (int l = kRaggedSingleMaxLog2 + 1; l <= kRaggedMaxLog2; l++)
                    fill_two_level((unsigned)l, g_table_image + kTables.lo[l], g_table_image + kTables.hi[l]);
                g_image_built = true;
            }
            void *p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                (void)hipGetLastError();
                set_error("scratch symbol lookup", e);
                return nullptr;
            }
            unsigned char *base = static_cast<unsigned char *>(p);
            e = hipMemcpy(base + kOffTables, g_table_image, sizeof(g_table_image), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                set_error("scratch table upload", e);
                return nullptr;
            }
            st.ptr = base;
            st.grid = ragged_resident_grid();
            st.ready.store(true, std::memory_order_release);
        }
    }
    *grid = st.grid;
    return st.ptr;
}

struct PlannerArgs {
    const onefft_ragged_item *items;
    uint32_t count;
    uint64_t extent;
    RaggedTask *tasks;
    uint64_t *item_off;
    uint32_t *header;
    uint32_t *status;
    uint32_t split, tile_run;
    uint32_t single_tw64[kRaggedSingleMaxLog2 + 1];  // stage-base offsets of the single-pass sizes
};

// Bins of items by (log2 n, direction), in task order: larger sizes first, forward before inverse.
constexpr int kBins = 2 * kSizes;
__device__ __forceinline__ int bin_of(uint32_t log2n, uint32_t inverse) {
    return (kRaggedMaxLog2 - (int)log2n) * 2 + (int)inverse;
}

// One workgroup: counts the items per bin, lays out the item order and the task queue (as
// onefft_ragged_plan_create does on the host), then each item writes its offset and its tasks.
// Status: 0 planned, 2 an item with log2 n > 20, a direction other than 0/1 or a range beyond
// `extent`, 3 more than kStaticMaxTasks tasks. On failure the queue is empty.
__global__ __launch_bounds__(1024) void fft_ragged_planner(PlannerArgs p) {
    __shared__ uint32_t cnt[kBins], fill[kBins], start[kBins], task0[kBins], runs_before[kBins];
    __shared__ uint32_t bad, ntasks, nphase1, whole_base, first_base, nsplit_items;
    const uint32_t t = threadIdx.x;
    if (t < kBins) {
        cnt[t] = 0;
        fill[t] = 0;
    }
    if (t == 0) bad = 0;
    __syncthreads();
    for (uint32_t i = t; i < p.count; i += blockDim.x) {
        const onefft_ragged_item it = p.items[i];
        const bool ok = it.log2n <= (uint32_t)kRaggedMaxLog2 && it.inverse <= 1 && it.offset <= p.extent &&
                        (uint64_t(1) << (it.log2n & 31)) <= p.extent - it.offset;
        if (ok) atomicAdd(&cnt[bin_of(it.log2n, it.inverse)], 1u);
        else atomicOr(&bad, 1u);
    }
    __syncthreads();
    if (t == 0) {
        // item positions: the bins in order; tasks: whole items, first-pass runs, single groups
        // (phase 1), then second-pass runs (phase 2)
        uint32_t pos = 0, whole = 0, runs = 0, split_items = 0;
        for (int b = 0; b < kBins; b++) {
            const uint32_t l = (uint32_t)(kRaggedMaxLog2 - b / 2);
            start[b] = pos;
            pos += cnt[b];
            if (l > (uint32_t)kRaggedSingleMaxLog2) {
                if (l >= p.split) {
                    const uint32_t tiles = 1u << (l - kRaggedTaskLog2);
                    runs_before[b] = runs;
                    runs += cnt[b] * ((tiles + p.tile_run - 1) / p.tile_run);
                    split_items += cnt[b];
                } else {
                    whole += cnt[b];
                }
            }
        }
        uint32_t singles = 0;
        for (int b = 0; b < kBins; b++) {
            const uint32_t l = (uint32_t)(kRaggedMaxLog2 - b / 2);
            if (l <= (uint32_t)kRaggedSingleMaxLog2) {
                task0[b] = whole + runs + singles;
                const uint32_t cap = ragged_per_task((int)l);
                singles += (cnt[b] + cap - 1) / cap;
            }
        }
        const uint64_t total = (uint64_t)whole + 2ull * runs + singles;
        uint32_t status = 0;
        if (bad) status = 2;
        else if (total > kStaticMaxTasks) status = 3;
        whole_base = 0;
        first_base = whole;
        nsplit_items = split_items;
        nphase1 = status ? 0 : whole + runs + singles;
        ntasks = status ? 0 : (uint32_t)total;
        p.header[0] = 0;
        p.header[1] = 0;
        p.header[4] = ntasks;
        p.header[5] = nphase1;
        *p.status = status;
    }
    __syncthreads();
    if (ntasks == 0) return;
    for (uint32_t i = t; i < p.count; i += blockDim.x) {
        const onefft_ragged_item it = p.items[i];
        const int b = bin_of(it.log2n, it.inverse);
        const uint32_t r = atomicAdd(&fill[b], 1u);
        const uint32_t k = start[b] + r;  // the item's position in task order
        p.item_off[k] = it.offset;
        const uint32_t l = it.log2n;
        if (l > (uint32_t)kRaggedSingleMaxLog2) {
            const uint32_t tiles = 1u << (l - kRaggedTaskLog2);
            if (l >= p.split) {
                const uint32_t runs = (tiles + p.tile_run - 1) / p.tile_run;
                const uint32_t base = runs_before[b] + r * runs;
                for (uint32_t j = 0; j < runs; j++) {
                    const uint32_t t0 = j * p.tile_run;
                    const uint32_t run = tiles - t0 < p.tile_run ? tiles - t0 : p.tile_run;
                    p.tasks[first_base + base + j] = {ragged_word(kRaggedFirst, l, it.inverse, run), k, t0, 0};
                    p.tasks[nphase1 + base + j] = {ragged_word(kRaggedSecond, l, it.inverse, run), k, t0, 0};
                }
            } else {
                // whole items follow the split items in the item order
                p.tasks[whole_base + k - nsplit_items] = {ragged_word(kRaggedWhole, l, it.inverse, tiles), k, 0, 0};
            }
        } else {
            const uint32_t cap = ragged_per_task((int)l);
            if (r % cap == 0) {
                const uint32_t left = cnt[b] - r;
                p.tasks[task0[b] + r / cap] = {ragged_word(kRaggedSingle, l, it.inverse), k, left < cap ? left : cap,
                                               p.single_tw64[l]};
            }
        }
    }
}

}  // namespace

}  // namespace onefft

using namespace onefft;

extern "C" int onefft_ragged_execute_items(const onefft_ragged_item *d_items, size_t count, size_t extent,
                                           const float *d_in, float *d_out, uint32_t *d_status,
                                           struct ihipStream_t *stream) {
    if (count > kStaticMaxItems) {
        set_error("onefft_ragged_execute_items: more than 131072 items");
        return -1;
    }
    if (!d_status || (count && !d_items) || !d_in || !d_out) {
        set_error("null item list, buffer or status");
        return -1;
    }
    if ((((uintptr_t)d_in | (uintptr_t)d_out | (uintptr_t)d_items) & 7) || ((uintptr_t)d_status & 3)) {
        set_error("device buffers must be 8-byte aligned (the status 4-byte aligned)");
        return -1;
    }
    if (extent > (size_t(1) << 40)) {
        set_error("extent beyond 2^40 complex values");
        return -1;
    }
    const unsigned __int128 bytes = (unsigned __int128)extent * sizeof(cf);
    if (!disjoint(d_in, bytes, d_out, bytes)) {
        set_error("input and output overlap");
        return -1;
    }
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        (void)hipGetLastError();
        set_error("no HIP device");
        return -1;
    }
    uint32_t grid = 0;
    unsigned char *base = device_scratch(device, &grid);
    if (!base) return -1;
    const RaggedChoice choice = ragged_choice();

    PlannerArgs p{};
    p.items = d_items;
    p.count = (uint32_t)count;
    p.extent = extent;
    p.tasks = reinterpret_cast<RaggedTask *>(base + kOffTasks);
    p.item_off = reinterpret_cast<uint64_t *>(base + kOffItems);
    p.header = reinterpret_cast<uint32_t *>(base);
    p.status = d_status;
    p.split = choice.split;
    p.tile_run = choice.tile_run;
    for (int l = 0; l <= kRaggedSingleMaxLog2; l++) p.single_tw64[l] = kTables.stage64[l];

    RaggedArgs a{};
    a.in = reinterpret_cast<const cf *>(d_in);
    a.out = reinterpret_cast<cf *>(d_out);
    a.tasks = p.tasks;
    a.item_off = p.item_off;
    a.head = p.header;
    a.error = d_status;  // the bounded wait's failure: status 1
    a.counts = p.header + 4;
    a.tw64 = reinterpret_cast<const cd *>(base + kOffTables);
    a.ntasks = 0;
    a.nphase1 = 0;
    a.nitems = (uint32_t)count;
    a.spin_limit = choice.spin_limit;
    for (int l = 0; l <= kRaggedMaxLog2; l++) {
        RaggedSize &z = a.size[l];
        if (l <= kRaggedSingleMaxLog2) {
            z.tw64 = kTables.stage64[l];
        } else {
            z.l1 = (uint32_t)l / 2;
            z.l2 = (uint32_t)l - z.l1;
            z.tw64_1 = kTables.stage64[z.l1];
            z.tw64_2 = kTables.stage64[z.l2];
            z.lo = kTables.lo[l];
            z.hi = kTables.hi[l];
            z.lo_bits = two_level_lo_bits((unsigned)l);
        }
    }
    hipLaunchKernelGGL(fft_ragged_planner, dim3(1), dim3(1024), 0, stream, p);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_ragged_kernel(a, grid, stream);
    if (e != hipSuccess) {
        set_error("ragged launch", e);
        return -1;
    }
    return 0;
}
