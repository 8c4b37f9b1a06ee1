// Ragged-batch plans: the host builds the task queue and the tables once. An execution is, for
// the items of the batched size and up (2^15..2^20 points; the threshold is tuned), one group per
// size and direction run by the multi-pass kernels at the items' offsets with that size's tuned
// split; for the rest, one reset kernel plus one persistent kernel (k_ragged.hip) over a device
// task queue.
//
// No counterpart in the reference, whose API transforms one array per call
// (include/onefft.h:61-77); onefft_ragged_execute_per_item runs the same items that way (one
// plan call per item on the device) for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/onefft.h"
#include "fft_common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using onefft::cd;
using onefft::cf;

// The items of one size and direction that run as one offset-batched multi-pass execution.
struct RaggedGroup {
    const onefft_plan *plan;   // the size's plan (one of sized_plans)
    bool inverse;
    uint32_t count;
    const uint64_t *offsets;   // device memory: the items' offsets, `count` entries
    int lane;                  // split orders: 0 the caller's stream, 1 the plan's stream
};

struct onefft_ragged_plan {
    int device = 0;
    size_t count = 0;
    uint64_t extent = 0;  // complex values the buffers must hold
    void *dmem = nullptr;
    onefft::RaggedArgs args{};  // the persistent kernel's items (args.ntasks == 0: none)
    uint32_t grid = 0;
    uint32_t order = 0;              // onefft::RaggedOrder
    hipStream_t side = nullptr;      // kConcurrent: the persistent kernel's stream; split orders:
                                     // lane 1's stream
    int persistent_lane = 0;         // kSplitAll: the lane whose groups the persistent kernel follows
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<RaggedGroup> groups;          // largest size first
    std::vector<onefft_plan *> sized_plans;   // owned: one per group size
    std::vector<onefft_ragged_item> items;    // as given, for the per-item comparison
};

namespace onefft {

RaggedChoice ragged_choice() {
    RaggedChoice c;
    const RaggedTune &t = ragged_tune();
    // items of 2^batched_min points and more run as offset-batched multi-pass groups
    // (kRaggedMaxLog2 + 1: none); the tuning table's choice, ONEFFT_RAGGED_BATCHED_MIN overrides it
    c.batched_min = t.set ? t.batched_min : (uint32_t)kRaggedBatchedMinLog2;
    if (const char *env = std::getenv("ONEFFT_RAGGED_BATCHED_MIN")) {
        const int v = std::atoi(env);
        if (v > kMaxSinglePassLog2 && v <= kRaggedMaxLog2 + 1) c.batched_min = (uint32_t)v;
    }
    // two-pass items below the split size run whole on one workgroup; larger ones are split into
    // runs of tiles over both phases (ONEFFT_RAGGED_SPLIT: the split size's log2)
    c.split = t.set ? t.split : (uint32_t)kRaggedSplitLog2;
    if (const char *env = std::getenv("ONEFFT_RAGGED_SPLIT")) {
        const int v = std::atoi(env);
        if (v > kRaggedSingleMaxLog2 && v <= kRaggedMaxLog2 + 1) c.split = (uint32_t)v;
    }
    // ONEFFT_RAGGED_TILE_RUN: tiles per pass task, 1..256
    c.tile_run = t.set ? t.tile_run : kRaggedTileRun;
    if (const char *env = std::getenv("ONEFFT_RAGGED_TILE_RUN")) {
        const int v = std::atoi(env);
        if (v >= 1 && v <= 256) c.tile_run = (uint32_t)v;
    }
    // ONEFFT_RAGGED_ORDER: 0 groups first, 1 persistent kernel first, 2 side by side, 3..5 the
    // groups split over two streams (runtime.hpp: RaggedOrder)
    c.order = t.set ? t.order : (uint32_t)kGroupsFirst;
    if (const char *env = std::getenv("ONEFFT_RAGGED_ORDER")) {
        const int v = std::atoi(env);
        if (v >= 0 && v < (int)kRaggedOrders) c.order = (uint32_t)v;
    }
    // the persistent kernel's workgroups: the table's choice (0: as many as are resident at once);
    // ONEFFT_RAGGED_GRID overrides it
    c.grid = t.set ? t.grid : 0;
    if (const char *env = std::getenv("ONEFFT_RAGGED_GRID")) {
        const long g = std::atol(env);
        if (g >= 0 && g <= (1L << 20)) c.grid = (uint32_t)g;
    }
    // the bounded wait for phase 1: ~3 s of 0.2 us sleeps (ONEFFT_RAGGED_SPIN_LIMIT overrides it,
    // so tests can make the wait give up and check that the failure is reported)
    c.spin_limit = 1u << 24;
    if (const char *env = std::getenv("ONEFFT_RAGGED_SPIN_LIMIT")) c.spin_limit = (uint32_t)std::strtoul(env, nullptr, 10);
    return c;
}

}  // namespace onefft

using namespace onefft;

extern "C" onefft_ragged_plan *onefft_ragged_plan_create(const onefft_ragged_item *items, size_t count) {
    if (!items && count) {
        set_error("null item list");
        return nullptr;
    }
    if (count > 0xffffffffu) {
        set_error("too many items");
        return nullptr;
    }
    // validate: sizes, directions, and no two items overlap
    std::vector<size_t> by_offset(count);
    for (size_t i = 0; i < count; i++) {
        if (items[i].log2n > (uint32_t)kRaggedMaxLog2 || items[i].inverse > 1) {
            set_error("ragged item size above 2^20 or direction not 0/1");
            return nullptr;
        }
        by_offset[i] = i;
    }
    std::sort(by_offset.begin(), by_offset.end(),
              [&](size_t a, size_t b) { return items[a].offset < items[b].offset; });
    uint64_t extent = 0;
    for (size_t k = 0; k < count; k++) {
        const onefft_ragged_item &it = items[by_offset[k]];
        if (it.offset < extent) {
            set_error("ragged items overlap");
            return nullptr;
        }
        extent = it.offset + (uint64_t(1) << it.log2n);
    }

    onefft_ragged_plan *plan = new (std::nothrow) onefft_ragged_plan;
    if (!plan) return nullptr;
    if (hipGetDevice(&plan->device) != hipSuccess) {
        set_error("no HIP device");
        delete plan;
        return nullptr;
    }
    plan->count = count;
    plan->extent = extent;
    plan->items.assign(items, items + count);
    const RaggedChoice choice = ragged_choice();
    const uint32_t split = choice.split, tile_run = choice.tile_run;
    auto batched = [&](const onefft_ragged_item &it) { return it.log2n >= choice.batched_min; };

    // the batched groups: one per size and direction, largest size first
    std::vector<std::vector<uint64_t>> group_off;
    std::vector<std::pair<uint32_t, bool>> group_key;
    for (int l = kRaggedMaxLog2; l > kMaxSinglePassLog2; l--) {
        if ((uint32_t)l < choice.batched_min) break;
        for (int inv = 0; inv < 2; inv++) {
            std::vector<uint64_t> offs;
            for (size_t i = 0; i < count; i++)
                if (items[i].log2n == (uint32_t)l && items[i].inverse == (uint32_t)inv) offs.push_back(items[i].offset);
            if (offs.empty()) continue;
            group_off.push_back(std::move(offs));
            group_key.push_back({(uint32_t)l, inv != 0});
        }
    }

    // the persistent kernel's items: tables, all FP64: the stage bases of every size used as a
    // single pass or as a sub-transform, and the two-level table of w_n of every two-pass size
    std::vector<uint32_t> order;
    for (size_t i = 0; i < count; i++)
        if (!batched(items[i])) order.push_back((uint32_t)i);
    const size_t nper = order.size();
    bool used[kRaggedMaxLog2 + 1] = {};
    for (uint32_t i : order) used[items[i].log2n] = true;
    std::vector<cd> tw64(1);
    int32_t stage64_at[kRaggedMaxLog2 + 1];
    for (auto &v : stage64_at) v = -1;
    auto stage64_table = [&](int log2) -> uint32_t {
        if (stage64_at[log2] < 0) {
            const int lp = batch_log2p(log2);
            const uint32_t entries = stage_tw64_size(log2, lp);
            stage64_at[log2] = (int32_t)tw64.size();
            tw64.resize(tw64.size() + entries + 1);
            if (entries) fill_stage_table64(log2, lp, tw64.data() + stage64_at[log2]);
        }
        return (uint32_t)stage64_at[log2];
    };
    RaggedArgs &a = plan->args;
    std::memset(&a, 0, sizeof(a));
    for (int l = 0; l <= kRaggedMaxLog2; l++) {
        if (!used[l]) continue;
        RaggedSize &z = a.size[l];
        if (l <= kRaggedSingleMaxLog2) {
            z.tw64 = stage64_table(l);
        } else {
            z.l1 = (uint32_t)l / 2;
            z.l2 = (uint32_t)l - z.l1;
            z.tw64_1 = stage64_table((int)z.l1);
            z.tw64_2 = stage64_table((int)z.l2);
            z.lo_bits = two_level_lo_bits((unsigned)l);
            z.lo = (uint32_t)tw64.size();
            z.hi = z.lo + (1u << z.lo_bits);
            tw64.resize(z.hi + (size_t(1) << (l - z.lo_bits)));
            fill_two_level((unsigned)l, tw64.data() + z.lo, tw64.data() + z.hi);
        }
    }

    // items in task order: two-pass items (largest first), then single-pass items grouped by size
    // (largest first) and direction
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        if (items[x].log2n != items[y].log2n) return items[x].log2n > items[y].log2n;
        return items[x].inverse < items[y].inverse;
    });
    size_t nbig = 0;
    while (nbig < nper && items[order[nbig]].log2n > (uint32_t)kRaggedSingleMaxLog2) nbig++;
    std::vector<uint64_t> item_off(nper ? nper : 1);
    for (size_t k = 0; k < nper; k++) item_off[k] = items[order[k]].offset;

    std::vector<RaggedTask> tasks;
    auto pass_tasks = [&](uint32_t kind) {
        for (size_t k = 0; k < nbig; k++) {
            const onefft_ragged_item &it = items[order[k]];
            if (it.log2n < split) continue;
            const uint32_t tiles = 1u << (it.log2n - kRaggedTaskLog2);  // tiles per pass
            for (uint32_t t0 = 0; t0 < tiles; t0 += tile_run) {
                const uint32_t run = tiles - t0 < tile_run ? tiles - t0 : tile_run;
                tasks.push_back({ragged_word(kind, it.log2n, it.inverse, run), (uint32_t)k, t0, 0});
            }
        }
    };
    // phase 1: the whole two-pass items (the longest tasks, largest first), the first passes of
    // the split items, then the single-pass items grouped by size (largest first) and direction
    for (size_t k = 0; k < nbig; k++) {
        const onefft_ragged_item &it = items[order[k]];
        if (it.log2n < split)
            tasks.push_back(
                {ragged_word(kRaggedWhole, it.log2n, it.inverse, 1u << (it.log2n - kRaggedTaskLog2)), (uint32_t)k, 0, 0});
    }
    pass_tasks(kRaggedFirst);
    for (size_t k = nbig; k < nper;) {
        const onefft_ragged_item &it = items[order[k]];
        size_t e = k;
        const uint32_t cap = ragged_per_task((int)it.log2n);
        while (e < nper && e - k < cap && items[order[e]].log2n == it.log2n && items[order[e]].inverse == it.inverse)
            e++;
        tasks.push_back({ragged_word(kRaggedSingle, it.log2n, it.inverse), (uint32_t)k, (uint32_t)(e - k), a.size[it.log2n].tw64});
        k = e;
    }
    const size_t nphase1 = tasks.size();
    // phase 2: the second passes
    pass_tasks(kRaggedSecond);
    if (tasks.size() > 0xffffffffu) {
        set_error("too many tasks");
        delete plan;
        return nullptr;
    }

    // one device allocation: tasks | item offsets | queue counters | tw64 | group offsets
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t o_tasks = 0;
    const size_t o_off = up(o_tasks + tasks.size() * sizeof(RaggedTask));
    const size_t o_head = up(o_off + item_off.size() * sizeof(uint64_t));
    const size_t o_tw64 = up(o_head + 4 * sizeof(uint32_t));
    size_t end = up(o_tw64 + tw64.size() * sizeof(cd));
    std::vector<size_t> o_group(group_off.size());
    for (size_t g = 0; g < group_off.size(); g++) {
        o_group[g] = end;
        end = up(end + group_off[g].size() * sizeof(uint64_t));
    }
    const size_t bytes = end;
    std::vector<unsigned char> image(bytes, 0);
    std::memcpy(image.data() + o_tasks, tasks.data(), tasks.size() * sizeof(RaggedTask));
    std::memcpy(image.data() + o_off, item_off.data(), item_off.size() * sizeof(uint64_t));
    std::memcpy(image.data() + o_tw64, tw64.data(), tw64.size() * sizeof(cd));
    for (size_t g = 0; g < group_off.size(); g++)
        std::memcpy(image.data() + o_group[g], group_off[g].data(), group_off[g].size() * sizeof(uint64_t));
    if (dev_malloc(&plan->dmem, bytes) != hipSuccess) {
        set_error("ragged plan allocation failed");
        delete plan;
        return nullptr;
    }
    if (hipMemcpy(plan->dmem, image.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        set_error("ragged plan upload failed");
        onefft_ragged_plan_destroy(plan);
        return nullptr;
    }
    unsigned char *d = static_cast<unsigned char *>(plan->dmem);
    a.tasks = reinterpret_cast<const RaggedTask *>(d + o_tasks);
    a.item_off = reinterpret_cast<const uint64_t *>(d + o_off);
    a.head = reinterpret_cast<uint32_t *>(d + o_head);
    a.error = a.head + 2;
    a.counts = nullptr;
    a.tw64 = reinterpret_cast<const cd *>(d + o_tw64);
    a.ntasks = (uint32_t)tasks.size();
    a.nphase1 = (uint32_t)nphase1;
    a.nitems = (uint32_t)nper;
    a.spin_limit = choice.spin_limit;
    // the persistent kernel waits only for tasks that running workgroups have claimed, so any
    // grid up to the resident count works; fewer workgroups publish and wait at the phase change
    // fewer times (each publication writes back its XCD's L2)
    plan->grid = ragged_resident_grid();
    if (choice.grid) plan->grid = std::min(plan->grid, choice.grid);
    plan->order = choice.order;
    // side by side needs both parts, a stream for the persistent kernel and two events; the split
    // orders need two groups or more
    if (plan->order >= kSplitGroups && group_off.size() < 2) plan->order = kGroupsFirst;
    if (plan->order == kConcurrent && !(a.ntasks && !group_off.empty())) plan->order = kGroupsFirst;
    if (plan->order == kConcurrent || plan->order >= kSplitGroups) {
        if (stream_create(&plan->side) != hipSuccess || event_create(&plan->fork) != hipSuccess ||
            event_create(&plan->join) != hipSuccess) {
            set_error("ragged plan stream or event creation failed");
            onefft_ragged_plan_destroy(plan);
            return nullptr;
        }
    }

    // the groups' plans (tables only: no staging), one per size
    for (size_t g = 0; g < group_off.size(); g++) {
        const uint32_t l = group_key[g].first;
        onefft_plan *sp = nullptr;
        for (onefft_plan *q : plan->sized_plans)
            if (onefft_plan_size(q) == (size_t(1) << l)) sp = q;
        if (!sp) {
            sp = plan_build(size_t(1) << l, false);
            if (!sp) {
                onefft_ragged_plan_destroy(plan);
                return nullptr;
            }
            plan->sized_plans.push_back(sp);
        }
        plan->groups.push_back({sp, group_key[g].second, (uint32_t)group_off[g].size(),
                                reinterpret_cast<const uint64_t *>(d + o_group[g]), 0});
    }
    // split orders: each group (largest first) to the lane with fewer bytes so far, so that one
    // lane's launches fill the other's drain between kernels
    if (plan->order >= kSplitGroups) {
        uint64_t bytes[2] = {0, 0};
        for (RaggedGroup &g : plan->groups) {
            g.lane = bytes[1] < bytes[0] ? 1 : 0;
            bytes[g.lane] += (uint64_t)g.count * onefft_plan_size(g.plan);
        }
        plan->persistent_lane = bytes[1] < bytes[0] ? 1 : 0;
    }
    return plan;
}

extern "C" void onefft_ragged_plan_destroy(onefft_ragged_plan *plan) {
    if (!plan) return;
    for (onefft_plan *p : plan->sized_plans) onefft_plan_destroy(p);
    if (plan->dmem || plan->side || plan->fork || plan->join) {
        DeviceGuard g(plan->device);
        if (plan->side) (void)hipStreamDestroy(plan->side);
        if (plan->fork) (void)hipEventDestroy(plan->fork);
        if (plan->join) (void)hipEventDestroy(plan->join);
        if (plan->dmem) (void)hipFree(plan->dmem);
    }
    delete plan;
}

extern "C" size_t onefft_ragged_extent(const onefft_ragged_plan *plan) { return plan ? (size_t)plan->extent : 0; }

// the persistent kernel's tasks (the batched groups' items are not counted)
extern "C" size_t onefft_ragged_task_count(const onefft_ragged_plan *plan) { return plan ? plan->args.ntasks : 0; }

int onefft::ragged_plan_device(const onefft_ragged_plan *plan) { return plan->device; }

// Both buffers hold `extent` complex values; the two-pass items' first passes read the input
// while other workgroups write the output, so the two must not overlap.
static bool ragged_args_ok(const onefft_ragged_plan *plan, const float *d_in, float *d_out) {
    if (!plan || (plan->count && (!d_in || !d_out))) {
        set_error("null plan or buffer");
        return false;
    }
    if (((uintptr_t)d_in | (uintptr_t)d_out) & 7) {
        set_error("device buffers must be 8-byte aligned");
        return false;
    }
    const unsigned __int128 bytes = (unsigned __int128)plan->extent * sizeof(cf);
    if (!disjoint(d_in, bytes, d_out, bytes)) {
        set_error("input and output overlap");
        return false;
    }
    return true;
}

extern "C" int onefft_ragged_execute(const onefft_ragged_plan *plan, const float *d_in, float *d_out,
                                     struct ihipStream_t *stream) {
    if (!ragged_args_ok(plan, d_in, d_out)) return -1;
    DeviceGuard g(plan->device);
    if (!g.ok) {
        set_error("cannot select the plan's device");
        return -1;
    }
    const cf *in = reinterpret_cast<const cf *>(d_in);
    cf *out = reinterpret_cast<cf *>(d_out);
    // the batched groups, largest first: all (lane < 0), or those of one lane
    auto groups = [&](hipStream_t s, int lane = -1) -> int {
        for (const RaggedGroup &g : plan->groups)
            if (lane < 0 || g.lane == lane)
                if (execute_device(g.plan, in, out, g.count, 0, 0, g.inverse, s, g.offsets) != 0) return -1;
        return 0;
    };
    // the persistent kernel over the other items
    auto persistent = [&](hipStream_t s) -> int {
        if (plan->args.ntasks == 0) return 0;
        RaggedArgs a = plan->args;
        a.in = in;
        a.out = out;
        hipError_t e = launch_ragged(a, plan->grid, s);
        if (e != hipSuccess) {
            set_error("ragged launch", e);
            return -1;
        }
        return 0;
    };
    if (plan->side) {
        // fork: the side stream starts after what the caller queued; join: the caller's stream
        // continues after the side stream's kernels (this also holds under stream capture)
        hipError_t e = hipEventRecord(plan->fork, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(plan->side, plan->fork, 0);
        if (e != hipSuccess) {
            set_error("ragged fork", e);
            return -1;
        }
        int rc;
        if (plan->order == kConcurrent) {
            rc = groups(stream) | persistent(plan->side);
        } else if (plan->order == kSplitLead) {
            rc = persistent(plan->side) | groups(stream, 0) | groups(plan->side, 1);
        } else {
            rc = groups(stream, 0) | groups(plan->side, 1);
            if (plan->order == kSplitAll) rc |= persistent(plan->persistent_lane ? plan->side : stream);
        }
        e = hipEventRecord(plan->join, plan->side);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, plan->join, 0);
        if (e != hipSuccess) {
            set_error("ragged join", e);
            return -1;
        }
        if (plan->order == kSplitGroups) rc |= persistent(stream);
        return rc;
    }
    if (plan->order == kPersistentFirst) return persistent(stream) != 0 ? -1 : groups(stream);
    return groups(stream) != 0 ? -1 : persistent(stream);
}

extern "C" int onefft_ragged_execute_per_item(const onefft_ragged_plan *plan, const float *d_in, float *d_out,
                                              struct ihipStream_t *stream) {
    if (!ragged_args_ok(plan, d_in, d_out)) return -1;
    DeviceGuard g(plan->device);
    if (!g.ok) {
        set_error("cannot select the plan's device");
        return -1;
    }
    const cf *in = reinterpret_cast<const cf *>(d_in);
    cf *out = reinterpret_cast<cf *>(d_out);
    for (const onefft_ragged_item &it : plan->items) {
        const size_t n = (size_t)1 << it.log2n;
        const onefft_plan *p = cached_plan(n);
        if (!p || execute_device(p, in + it.offset, out + it.offset, 1, n, n, it.inverse != 0, stream) != 0)
            return -1;
    }
    return 0;
}

extern "C" int onefft_ragged_wait(const onefft_ragged_plan *plan, struct ihipStream_t *stream) {
    if (!plan) {
        set_error("null plan");
        return -1;
    }
    DeviceGuard g(plan->device);
    if (!g.ok) {
        set_error("cannot select the plan's device");
        return -1;
    }
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        set_error("ragged execution", e);
        return -1;
    }
    // the kernel's bounded wait for phase 1 sets the error flag if it ever gives up
    uint32_t flag = 0;
    e = hipMemcpy(&flag, plan->args.error, sizeof(flag), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("ragged status read", e);
        return -1;
    }
    if (flag) {
        set_error("ragged execution: the wait for the first passes timed out");
        return -1;
    }
    return 0;
}
