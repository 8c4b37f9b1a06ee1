// Internal interfaces between the runtime's translation units (not part of the public API).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "fft_common.hpp"
#include "kernels.hpp"

struct onefft_plan;
struct onefft_ragged_plan;

namespace onefft {

// n -> log2 n for a power of two from 1 to 2^30, else false.
bool size_log2(size_t n, unsigned *log2n);

// Record the message returned by onefft_last_error() on this thread.
void set_error(const char *what, hipError_t e);
void set_error(const char *what);

// Launch `batch` transforms on device memory, asynchronously on `stream`. With d_offsets (device
// memory, `batch` entries; multi-pass plans only), transform b is at d_offsets[b] complex values
// in both buffers and the distances are ignored.
// With `done` (device address of a word in coherent host memory; one single-pass transform only),
// the kernel stores `seq` there once its output is visible to the host (kernels.hpp: BatchArgs).
int execute_device(const onefft_plan *plan, const cf *d_in, cf *d_out, uint64_t batch, uint64_t in_dist,
                   uint64_t out_dist, bool inverse, hipStream_t stream, const uint64_t *d_offsets = nullptr,
                   uint32_t *done = nullptr, uint32_t seq = 0);

// onefft_plan_create without its side effect on the staging pool (staging = false): the plans
// that ragged plans hold for their batched groups. Freed by onefft_plan_destroy.
onefft_plan *plan_build(size_t n, bool staging);

// Sizes the staging pool of `device` (staging.hip): creates its slots (ONEFFT_STAGING_SLOTS, each
// with a stream and events) and, with `buffers`, gives every slot what a host-pointer call of n
// points needs, so that such calls allocate nothing. false on failure.
bool staging_prepare(int device, size_t n, bool buffers);
// Frees every staging slot of every device (onefft_cache_clear).
void staging_release_all();

// The library's allocations (alloc.hip): hipMalloc, hipHostMalloc, a non-blocking stream and a
// timing-free event, each of which the ONEFFT_FAIL_ALLOC switch can make fail.
hipError_t dev_malloc(void **p, size_t bytes);
hipError_t host_malloc(void **p, size_t bytes, unsigned flags);
hipError_t stream_create(hipStream_t *s);
hipError_t event_create(hipEvent_t *e);

// The synchronous drop-in call: host or device pointers of any alignment.
int execute_host_call(const onefft_plan *plan, const float *in, float *out, bool inverse);

// The cache's plan for size n (created on first use), or nullptr.
const onefft_plan *cached_plan(size_t n);

// One size's entry of the tuning table (tuning.hip), per working-set regime: 0 when the batch's
// input plus output fit the 256 MiB Infinity Cache (MALL), 1 when they do not.
constexpr int kTuneMaxLog2 = 30;
constexpr uint64_t kMallBytes = uint64_t(256) << 20;
inline int tune_regime(uint64_t n, uint64_t batch) { return 16 * n * batch > kMallBytes ? 1 : 0; }
struct TuneEntry {
    enum Kind { kNone, kSingle, kMulti } kind = kNone;
    BatchVariant variant;        // kSingle
    int passes = 0;              // kMulti: the split, log2 N_1 first, and columns per pass factor
    int split[kMaxPasses] = {};
    int cols[kMaxPasses] = {};
    int pass_layout = 0;         // kMulti: LDS layout of every pass (kPadded or kHalves)
};
const TuneEntry &tune_lookup(unsigned log2n, int regime);

// Makes `device` current for the guard's lifetime and restores the previous device after; `ok`
// is false if the device could not be selected.
struct DeviceGuard {
    int previous = -1;
    bool ok = true;
    explicit DeviceGuard(int device);
    ~DeviceGuard();
    DeviceGuard(const DeviceGuard &) = delete;
    DeviceGuard &operator=(const DeviceGuard &) = delete;
};

// The ragged executions' choices: the smallest size run as offset-batched multi-pass groups
// (log2; kRaggedMaxLog2 + 1 for none), the split size (log2; two-pass items of the persistent
// kernel from it up are spread over tile runs, smaller ones run whole), tiles per pass task, the
// order of the groups and the persistent kernel (RaggedOrder), the persistent kernel's
// workgroups, and the bounded wait's sleep count. The allocation-free entry point uses split, tile run and wait only (its persistent
// kernel runs every item).
enum RaggedOrder : uint32_t {
    kGroupsFirst = 0,      // the groups, then the persistent kernel, on the caller's stream
    kPersistentFirst = 1,  // the persistent kernel, then the groups
    kConcurrent = 2,       // the persistent kernel on the plan's own stream beside the groups
    kSplitGroups = 3,      // the groups over two streams (the caller's and the plan's), balanced by
                           // bytes, then the persistent kernel after both
    kSplitAll = 4,         // as kSplitGroups, with the persistent kernel after the groups of the
                           // stream that has fewer bytes
    kSplitLead = 5,        // as kSplitGroups, with the persistent kernel before the groups of the
                           // plan's stream
};
constexpr uint32_t kRaggedOrders = 6;
struct RaggedChoice {
    uint32_t batched_min, split, tile_run, order, spin_limit;
    uint32_t grid;  // persistent workgroups (0: as many as are resident at once)
};
RaggedChoice ragged_choice();

// The tuning table's ragged line (tuning.hip), if it has one.
struct RaggedTune {
    bool set = false;
    uint32_t batched_min = 0, split = 0, tile_run = 0, order = 0, grid = 0;
};
const RaggedTune &ragged_tune();

// The device a plan or a ragged plan belongs to.
int plan_device(const onefft_plan *plan);
int ragged_plan_device(const onefft_ragged_plan *plan);

// Checks the device-pointer arguments of a batched call: non-null, 8-byte aligned, distances of
// at least the size when batch > 1, and input and output byte ranges that do not overlap (the
// passes of a multi-pass transform read input that other workgroups would already have
// overwritten). Records the reason and returns false otherwise.
bool device_args_ok(const onefft_plan *plan, const void *d_in, const void *d_out, size_t batch, size_t in_dist,
                    size_t out_dist);

// True if the byte ranges [a, a + alen) and [b, b + blen) are disjoint.
inline bool disjoint(const void *a, unsigned __int128 alen, const void *b, unsigned __int128 blen) {
    const unsigned __int128 x = (uintptr_t)a, y = (uintptr_t)b;
    return alen == 0 || blen == 0 || x + alen <= y || y + blen <= x;
}

}  // namespace onefft
