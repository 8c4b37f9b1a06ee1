// Host runtime behind include/onefft.h: argument checks, plans with device-resident twiddle
// tables, the write-once plan cache, the host-pointer staging path and the device-pointer,
// batched entry points.
//
// Reference counterparts (src/onefft.c): plan creation and validation :1111-1136, the cache
// :1144-1189, the public transforms :1138-1142 and :1169-1181, the twiddle function :112-141.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/onefft.h"
#include "fft_common.hpp"
#include "kernels.hpp"
#include "plan_tables.hpp"
#include "runtime.hpp"

using onefft::BatchVariant;
using onefft::cd;
using onefft::cf;

namespace {
constexpr unsigned kMaxLog2 = 30;  // the reference's limit (src/onefft.c:1117-1121)
// Plans up to this size also size the staging pool's buffers for host-pointer calls of their size
// (per slot at most 2 x 32 MiB of device memory and a 16 MiB pinned ring), so those calls allocate
// nothing on any thread; larger plans, often used with device pointers only, create the pool's
// slots (streams) but leave the buffers to the first host-pointer call or onefft_cache_prepare.
constexpr unsigned kPlanStagingMaxLog2 = 22;
}  // namespace

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Pass split of a multi-pass size: log2 N_1 .. log2 N_p (N_1 is transformed last), the columns
// per workgroup of each factor's pass, and the passes' LDS layout. The environment variable ONEFFT_PASS_SPLIT
// ("a,b[,c[,d]]", log2 N_1 first) overrides for sizes whose factors it matches, to sweep splits
// without rebuilding (the reference's analogue is its ONEFFT_X86_TUNING build switch,
// src/onefft.c:652-697); else the tuning table's entry; else an even split into passes of at
// most 2^9 points, the larger factors first in execution order.
int choose_cols(int log2nf);
int choose_split_factors(unsigned log2n, int regime, int *factors);

int choose_split(unsigned log2n, int regime, int *factors, int *cols, int *layout) {
    const int p = choose_split_factors(log2n, regime, factors);
    const onefft::TuneEntry &t = onefft::tune_lookup(log2n, regime);
    bool from_table = t.kind == onefft::TuneEntry::kMulti && t.passes == p;
    for (int i = 0; i < p && from_table; i++) from_table = t.split[i] == factors[i];
    for (int i = 0; i < p; i++) cols[i] = from_table ? t.cols[i] : choose_cols(factors[i]);
    *layout = from_table ? t.pass_layout : onefft::kPadded;
    // ONEFFT_PASS_LAYOUT: 0 (padded) or 2 (halves) for every pass, for measurement
    if (const char *env = std::getenv("ONEFFT_PASS_LAYOUT")) {
        const int v = std::atoi(env);
        if (v == onefft::kPadded || v == onefft::kHalves) *layout = v;
    }
    // ONEFFT_PASS_COLS: one count for every pass, or a list "c_1,..,c_p" (N_1 first); counts a
    // pass does not allow are ignored
    if (const char *env = std::getenv("ONEFFT_PASS_COLS")) {
        int c[onefft::kMaxPasses] = {}, k = 0;
        for (const char *q = env; *q && k < onefft::kMaxPasses;) {
            char *end;
            const long v = std::strtol(q, &end, 10);
            if (end == q) break;
            c[k++] = (int)v;
            q = *end == ',' ? end + 1 : end;
        }
        for (int i = 0; i < p; i++) {
            const int want = k == p ? c[i] : c[0];
            if (k && onefft::pass_cols_ok(factors[i], want, *layout)) cols[i] = want;
        }
    }
    // a table's counts for the halves layout may not fit the padded one that ONEFFT_PASS_LAYOUT chose
    for (int i = 0; i < p; i++)
        if (!onefft::pass_cols_ok(factors[i], cols[i], *layout)) cols[i] = choose_cols(factors[i]);
    return p;
}

int choose_split_factors(unsigned log2n, int regime, int *factors) {
    if (const char *env = std::getenv("ONEFFT_PASS_SPLIT")) {
        int f[onefft::kMaxPasses], p = 0, sum = 0;
        const char *s = env;
        while (*s && p < onefft::kMaxPasses) {
            char *end;
            long v = std::strtol(s, &end, 10);
            if (end == s) break;
            f[p++] = (int)v;
            sum += (int)v;
            s = *end == ',' ? end + 1 : end;
        }
        bool ok = p >= 2 && sum == (int)log2n;
        for (int i = 0; i < p && ok; i++) ok = f[i] >= onefft::kPassMinLog2 && f[i] <= onefft::kPassMaxLog2;
        if (ok) {
            for (int i = 0; i < p; i++) factors[i] = f[i];
            return p;
        }
    }
    const onefft::TuneEntry &t = onefft::tune_lookup(log2n, regime);
    if (t.kind == onefft::TuneEntry::kMulti) {
        for (int i = 0; i < t.passes; i++) factors[i] = t.split[i];
        return t.passes;
    }
    int p = (int)((log2n + 8) / 9);
    if (p < 2) p = 2;
    for (int i = 0; i < p; i++) factors[i] = (int)log2n / p + (i >= p - (int)log2n % p ? 1 : 0);
    return p;
}

// Single-pass kernel variant: the tuning table's choice, else the BatchVariant defaults
// (kernels.hpp). ONEFFT_BATCH_VARIANT="nt,layout[,tw64[,xlane]]" (nt 0/1, layout 0..2, tw64 the
// FP64-base level 0..3, 1 when omitted, xlane 0/1 the DPP pair exchange of 32-point batches)
// overrides the table, for measurement.
onefft::BatchVariant choose_variant(unsigned log2n, int regime) {
    onefft::BatchVariant v;
    const onefft::TuneEntry &t = onefft::tune_lookup(log2n, regime);
    if (t.kind == onefft::TuneEntry::kSingle) v = t.variant;
    if (const char *env = std::getenv("ONEFFT_BATCH_VARIANT")) {
        int nt = 1, layout = 0, tw64 = 1, xlane = 0;
        if (std::sscanf(env, "%d,%d,%d,%d", &nt, &layout, &tw64, &xlane) >= 2) {
            v.nt = nt != 0;
            v.layout = layout < 0 ? 0 : layout > 2 ? 2 : layout;
            v.tw64 = tw64 < 0 ? 0 : tw64 > 3 ? 3 : tw64;
            v.xlane = xlane != 0;
        }
    }
    return v;
}

// Columns per workgroup of a pass over 2^log2nf-point sub-transforms: the kernels' default.
int choose_cols(int log2nf) { return onefft::pass_cols(log2nf); }

}  // namespace

// ---- plans ----------------------------------------------------------------------------------

struct PassPlan {
    int log2nf;
    int cols;  // columns per workgroup
    onefft::PassArgs args;  // pointers to data filled in per call
};

// The passes of a multi-pass transform, in execution order.
struct PassConfig {
    int passes = 0;
    int layout = onefft::kPadded;  // LDS layout of every pass
    PassPlan pass[onefft::kMaxPasses];
};

// Choices are made per working-set regime (runtime.hpp: tune_regime), index 0 or 1.
struct onefft_plan {
    size_t n = 0;
    unsigned log2n = 0;
    int device = 0;
    void *dmem = nullptr;  // the one device allocation holding every table
    bool multi = false;    // false: one LDS pass
    const cf *tw_single = nullptr;
    const cd *tw64_single = nullptr;
    BatchVariant variant[2];  // single-pass kernel variant per regime
    PassConfig cfg[2];        // multi-pass split per regime
};

namespace onefft {

static thread_local std::string t_error;

void set_error(const char *what, hipError_t e) { t_error = std::string(what) + ": " + hipGetErrorString(e); }
void set_error(const char *what) { t_error = what; }

bool size_log2(size_t n, unsigned *log2n) {
    if (n == 0 || (n & (n - 1)) != 0) return false;
    unsigned l = 0;
    while (((size_t)1 << l) < n) l++;
    if (l > kMaxLog2) return false;
    *log2n = l;
    return true;
}

DeviceGuard::DeviceGuard(int device) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) {
        ok = false;
        return;
    }
    if (cur != device) {
        if (hipSetDevice(device) != hipSuccess) ok = false;
        else previous = cur;
    }
}

DeviceGuard::~DeviceGuard() {
    if (previous >= 0) (void)hipSetDevice(previous);
}

// Launches the transform of `batch` arrays on `stream`; data pointers are device memory.
int execute_device(const onefft_plan *plan, const cf *d_in, cf *d_out, uint64_t batch, uint64_t in_dist,
                   uint64_t out_dist, bool inverse, hipStream_t stream, const uint64_t *d_offsets, uint32_t *done,
                   uint32_t seq) {
    if (batch == 0) return 0;
    const int regime = tune_regime(plan->n, batch);
    if (d_offsets && !plan->multi) {
        set_error("transforms at listed offsets need a multi-pass plan");
        return -1;
    }
    if (done && (plan->multi || batch != 1)) {
        set_error("a completion word needs one single-pass transform");
        return -1;
    }
    if (!plan->multi) {
        BatchArgs a{d_in, d_out, plan->tw_single, plan->tw64_single, in_dist, out_dist, batch, done, seq};
        hipError_t e = launch_batch((int)plan->log2n, inverse, a, plan->variant[regime], stream);
        if (e != hipSuccess) {
            set_error("single-pass launch", e);
            return -1;
        }
        return 0;
    }
    const PassConfig &cfg = plan->cfg[regime];
    for (uint64_t done = 0; done < batch;) {
        const uint64_t chunk = batch - done < 65535 ? batch - done : 65535;
        for (int i = 0; i < cfg.passes; i++) {
            PassArgs a = cfg.pass[i].args;
            const bool first = i == 0, last = i == cfg.passes - 1;
            if (d_offsets) {
                a.in = first ? d_in : d_out;
                a.out = d_out;
                a.offsets = d_offsets + done;
            } else {
                a.in = first ? d_in + done * in_dist : d_out + done * out_dist;
                a.out = d_out + done * out_dist;
                a.in_dist = first ? in_dist : out_dist;
                a.out_dist = out_dist;
            }
            hipError_t e = launch_pass(cfg.pass[i].log2nf, cfg.pass[i].cols, cfg.layout, first, first && inverse,
                                       last && inverse, a, chunk, stream);
            if (e != hipSuccess) {
                set_error("pass launch", e);
                return -1;
            }
        }
        done += chunk;
    }
    return 0;
}

}  // namespace onefft

using namespace onefft;

onefft_plan *onefft::plan_build(size_t n, bool staging) {
    unsigned log2n;
    if (!size_log2(n, &log2n)) return nullptr;
    onefft_plan *plan = new (std::nothrow) onefft_plan;
    if (!plan) return nullptr;
    plan->n = n;
    plan->log2n = log2n;
    plan->variant[0] = choose_variant(log2n, 0);
    plan->variant[1] = choose_variant(log2n, 1);
    if (hipGetDevice(&plan->device) != hipSuccess) {
        set_error("no HIP device");
        delete plan;
        return nullptr;
    }

    // host image of the device tables, then one allocation and one copy
    std::vector<unsigned char> image;
    auto reserve = [&](size_t bytes) {
        size_t off = align_up(image.size(), 256);
        image.resize(off + bytes);
        return off;
    };
    size_t off_single = 0, off_single64 = 0, off_stage64[2][kMaxPasses] = {};
    size_t off_lo = 0, off_hi = 0, off_col[2][kMaxPasses];
    for (auto &per : off_col)
        for (size_t &o : per) o = SIZE_MAX;  // no column table
    if (log2n <= (unsigned)kMaxSinglePassLog2) {
        const int lp = batch_log2p((int)log2n);
        const uint32_t entries = stage_tw_size((int)log2n, lp);
        off_single = reserve(entries * sizeof(cf) + sizeof(cf));
        fill_stage_table((int)log2n, lp, reinterpret_cast<cf *>(image.data() + off_single));
        off_single64 = reserve(stage_tw64_size((int)log2n, lp) * sizeof(cd) + sizeof(cd));
        fill_stage_table64((int)log2n, lp, reinterpret_cast<cd *>(image.data() + off_single64));
    } else {
        plan->multi = true;
        // two-level FP64 table of w_n, shared by both configurations
        const unsigned lo_bits = two_level_lo_bits(log2n);
        const size_t lo_count = (size_t)1 << lo_bits, hi_count = (size_t)1 << (log2n - lo_bits);
        off_lo = reserve(lo_count * sizeof(cd));
        off_hi = reserve(hi_count * sizeof(cd));
        fill_two_level(log2n, reinterpret_cast<cd *>(image.data() + off_lo), reinterpret_cast<cd *>(image.data() + off_hi));
        for (int r = 0; r < 2; r++) {
            int factors[kMaxPasses], cols[kMaxPasses], layout = kPadded;
            const int p = choose_split(log2n, r, factors, cols, &layout);
            PassConfig &cfg = plan->cfg[r];
            cfg.passes = p;
            cfg.layout = layout;
            // execution order: factor f = p, p-1, ..., 1 (factors[f-1])
            for (int i = 0; i < p; i++) {
                const int f = p - i;  // 1-based factor index
                const int lf = factors[f - 1];
                const int lp = pass_log2p(lf);
                off_stage64[r][i] = reserve(stage_tw64_size(lf, lp) * sizeof(cd) + sizeof(cd));
                fill_stage_table64(lf, lp, reinterpret_cast<cd *>(image.data() + off_stage64[r][i]));
                PassArgs &a = cfg.pass[i].args;
                std::memset(&a, 0, sizeof(a));
                cfg.pass[i].log2nf = lf;
                cfg.pass[i].cols = cols[f - 1];
                unsigned log2m = 0;
                for (int l = f + 1; l <= p; l++) log2m += (unsigned)factors[l - 1];
                a.lo_bits = lo_bits;
                a.log2n = log2n;
                a.log2m = log2m;
                a.tw_shift = log2n - log2m - (unsigned)lf;
                if (log2m > 0) {
                    // the in-place pass's column table (kernels.hpp: PassArgs::tw_col); its step
                    // bases come from it too where the lookups of x low would sweep more than
                    // 2^10 entries (16 KiB) of the low level, w_n^(x << tw_shift) for
                    // x < 2^(lo_bits - tw_shift), and so miss the L1; below that they hit the L1
                    // and the table's two more products cost more than they save
                    const uint32_t C = (uint32_t)cols[f - 1], T = 1u << (lf - lp);
                    const uint32_t stride = 1u << (lf - stage_log2r(lf, lp, 0));
                    const uint32_t mask = (1u << log2m) - 1;  // c mod M_f: below C when C > M_f
                    off_col[r][i] = reserve((T * C + 2 * C) * sizeof(cd));
                    cd *col = reinterpret_cast<cd *>(image.data() + off_col[r][i]);
                    auto root = [&](uint64_t x, cd *w) { root64((x << a.tw_shift) & (n - 1), log2n, &w->r, &w->i); };
                    for (uint32_t t = 0; t < T * C; t++) root((uint64_t)(t / C) * ((t % C) & mask), &col[t]);
                    for (uint32_t c = 0; c < C; c++) {
                        root((uint64_t)T * (c & mask), &col[T * C + c]);
                        root((uint64_t)stride * (c & mask), &col[T * C + C + c]);
                    }
                    a.col_steps = lo_bits > a.tw_shift + 10 ? 1 : 0;
                }
                a.columns = (uint32_t)(n >> lf);
                // ONEFFT_PASS_STRIPS: strips of the input pass's tile order (for measurement)
                a.strips = 1;
                if (const char *env = std::getenv("ONEFFT_PASS_STRIPS"))
                    if (i == 0) a.strips = (uint32_t)std::strtoul(env, nullptr, 10);
                if (i == 0) {
                    a.ndig = (uint32_t)(p - 1);
                    for (int d = 0; d < p - 1; d++) {  // digit j_{d+1}
                        a.dig_log2[d] = (uint32_t)factors[d];
                        uint32_t shift = 0;
                        for (int l = d + 2; l <= p - 1; l++) shift += (uint32_t)factors[l - 1];
                        a.dig_shift[d] = shift;
                    }
                }
            }
        }
    }

    if (dev_malloc(&plan->dmem, image.size()) != hipSuccess) {
        set_error("plan table allocation failed");
        delete plan;
        return nullptr;
    }
    if (hipMemcpy(plan->dmem, image.data(), image.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        set_error("plan table upload failed");
        (void)hipFree(plan->dmem);
        delete plan;
        return nullptr;
    }
    auto dev = [&](size_t off) { return reinterpret_cast<unsigned char *>(plan->dmem) + off; };
    if (!plan->multi) {
        plan->tw_single = reinterpret_cast<const cf *>(dev(off_single));
        plan->tw64_single = reinterpret_cast<const cd *>(dev(off_single64));
    } else {
        for (int r = 0; r < 2; r++) {
            for (int i = 0; i < plan->cfg[r].passes; i++) {
                PassArgs &a = plan->cfg[r].pass[i].args;
                a.tw64 = reinterpret_cast<const cd *>(dev(off_stage64[r][i]));
                a.tw_lo = reinterpret_cast<const cd *>(dev(off_lo));
                a.tw_hi = reinterpret_cast<const cd *>(dev(off_hi));
                a.tw_col = off_col[r][i] == SIZE_MAX ? nullptr : reinterpret_cast<const cd *>(dev(off_col[r][i]));
            }
        }
    }
    if (staging && !staging_prepare(plan->device, n, log2n <= kPlanStagingMaxLog2)) {
        (void)hipFree(plan->dmem);
        delete plan;
        return nullptr;
    }
    return plan;
}

extern "C" onefft_plan *onefft_plan_create(size_t n) { return plan_build(n, true); }

extern "C" void onefft_plan_destroy(onefft_plan *plan) {
    if (!plan) return;
    if (plan->dmem) {
        DeviceGuard g(plan->device);
        (void)hipFree(plan->dmem);
    }
    delete plan;
}

extern "C" size_t onefft_plan_size(const onefft_plan *plan) { return plan ? plan->n : 0; }

extern "C" const char *onefft_implementation(void) { return "hip-gfx950"; }

extern "C" const char *onefft_last_error(void) { return t_error.c_str(); }

static int run_sync(const onefft_plan *plan, const float *in, float *out, bool inverse) {
    if (!plan || !in || !out) {
        set_error("null plan or buffer");
        return -1;
    }
    DeviceGuard g(plan->device);
    if (!g.ok) {
        set_error("cannot select the plan's device");
        return -1;
    }
    return execute_host_call(plan, in, out, inverse);
}

extern "C" void onefft_plan_forward(const onefft_plan *plan, const float *in, float *out) {
    (void)run_sync(plan, in, out, false);
}
extern "C" void onefft_plan_inverse(const onefft_plan *plan, const float *in, float *out) {
    (void)run_sync(plan, in, out, true);
}

int onefft::plan_device(const onefft_plan *plan) { return plan->device; }

bool onefft::device_args_ok(const onefft_plan *plan, const void *d_in, const void *d_out, size_t batch,
                            size_t in_dist, size_t out_dist) {
    if (!plan || !d_in || !d_out) {
        set_error("null plan or buffer");
        return false;
    }
    if (((uintptr_t)d_in | (uintptr_t)d_out) & 7) {
        set_error("device buffers must be 8-byte aligned");
        return false;
    }
    if (batch > 1 && (in_dist < plan->n || out_dist < plan->n)) {
        set_error("batch distance smaller than the transform size");
        return false;
    }
    if (batch == 0) return true;
    // bytes from the first transform's start to the last one's end
    const auto span = [&](size_t dist) { return ((unsigned __int128)(batch - 1) * dist + plan->n) * sizeof(cf); };
    if (!disjoint(d_in, span(in_dist), d_out, span(out_dist))) {
        set_error("input and output overlap");
        return false;
    }
    return true;
}

static int run_device(const onefft_plan *plan, const float *d_in, float *d_out, size_t batch, size_t in_dist,
                      size_t out_dist, bool inverse, hipStream_t stream) {
    if (!device_args_ok(plan, d_in, d_out, batch, in_dist, out_dist)) return -1;
    DeviceGuard g(plan->device);
    if (!g.ok) {
        set_error("cannot select the plan's device");
        return -1;
    }
    return execute_device(plan, reinterpret_cast<const cf *>(d_in), reinterpret_cast<cf *>(d_out), batch, in_dist,
                          out_dist, inverse, stream);
}

extern "C" int onefft_plan_forward_device(const onefft_plan *plan, const float *d_in, float *d_out,
                                          struct ihipStream_t *stream) {
    return run_device(plan, d_in, d_out, 1, plan ? plan->n : 0, plan ? plan->n : 0, false, stream);
}
extern "C" int onefft_plan_inverse_device(const onefft_plan *plan, const float *d_in, float *d_out,
                                          struct ihipStream_t *stream) {
    return run_device(plan, d_in, d_out, 1, plan ? plan->n : 0, plan ? plan->n : 0, true, stream);
}
extern "C" int onefft_plan_forward_batch(const onefft_plan *plan, const float *d_in, float *d_out, size_t batch,
                                         size_t in_dist, size_t out_dist, struct ihipStream_t *stream) {
    return run_device(plan, d_in, d_out, batch, in_dist, out_dist, false, stream);
}
extern "C" int onefft_plan_inverse_batch(const onefft_plan *plan, const float *d_in, float *d_out, size_t batch,
                                         size_t in_dist, size_t out_dist, struct ihipStream_t *stream) {
    return run_device(plan, d_in, d_out, batch, in_dist, out_dist, true, stream);
}

// ---- the plan cache (src/onefft.c:1144-1189) -------------------------------------------------
//
// One slot per log2 n, written once with a compare-and-swap and only read afterwards: a warm call
// is one acquire load. Two threads that create the same plan at once both build it; the loser
// destroys its copy.

static std::atomic<onefft_plan *> g_cache[64];

const onefft_plan *onefft::cached_plan(size_t n) {
    unsigned log2n;
    if (!size_log2(n, &log2n)) return nullptr;
    std::atomic<onefft_plan *> &slot = g_cache[log2n];
    onefft_plan *plan = slot.load(std::memory_order_acquire);
    if (plan) return plan;
    plan = onefft_plan_create(n);
    if (!plan) return nullptr;
    onefft_plan *expected = nullptr;
    if (!slot.compare_exchange_strong(expected, plan, std::memory_order_acq_rel, std::memory_order_acquire)) {
        onefft_plan_destroy(plan);
        plan = expected;
    }
    return plan;
}

extern "C" int onefft_forward(const float *in, float *out, size_t n) {
    const onefft_plan *plan = cached_plan(n);
    if (!plan) return -1;
    return run_sync(plan, in, out, false);
}

extern "C" int onefft_inverse(const float *in, float *out, size_t n) {
    const onefft_plan *plan = cached_plan(n);
    if (!plan) return -1;
    return run_sync(plan, in, out, true);
}

extern "C" int onefft_cache_prepare(size_t n) {
    const onefft_plan *plan = cached_plan(n);
    if (!plan) return -1;
    DeviceGuard g(plan->device);
    if (!g.ok) {
        set_error("cannot select the plan's device");
        return -1;
    }
    return staging_prepare(plan->device, n, true) ? 0 : -1;
}

extern "C" void onefft_cache_clear(void) {
    for (auto &slot : g_cache) onefft_plan_destroy(slot.exchange(nullptr, std::memory_order_acq_rel));
    staging_release_all();
}
