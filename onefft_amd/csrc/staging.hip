// The synchronous drop-in path: onefft_forward/inverse and onefft_plan_forward/inverse with host
// or device pointers of any alignment (the reference accepts any alignment,
// include/onefft.h:61-64; its tests misalign by 4 and 12 bytes, tests/test_onefft.c:118-119).
//
// Staging lives in a process-wide pool per device of slots, each one caller's worth: a stream, a
// pair of device staging buffers, a ring of pinned host chunks and a coherent pinned buffer. A
// call claims a free slot with one atomic exchange (trying first the slot its thread used last)
// and returns it when done, so calls never wait for each other and plans stay read-only
// (reference include/onefft.h:102-104). onefft_plan_create and onefft_cache_prepare size the
// pool (ONEFFT_STAGING_SLOTS slots, default 4) for their size, after which calls of that size
// allocate nothing on any thread while at most that many run at once (the reference's promise,
// README.md:35-41, include/onefft.h:66-71). Only a call that finds every slot busy, or a slot
// smaller than it needs, allocates (a new slot, or a larger buffer). Slots are released by
// onefft_cache_clear, never at thread or process exit.
//
// Per argument:
//   * device memory, 8-byte aligned: transformed where it lies;
//   * device memory, misaligned: copied device to device into the staging buffer and back;
//   * host memory, pageable or pinned, transforms of one LDS pass (n <= 2^14, 128 KiB): copied by
//     the CPU into the slot's coherent pinned buffer, which the kernel reads and writes directly
//     over the host link (no DMA transfers, whose set-up latency dominates at these sizes);
//   * pinned host memory, larger: copied by DMA straight to and from the staging buffers;
//   * ordinary (pageable) host memory up to 4 MiB: HIP's own pageable copies (lower set-up latency
//     than the ring below at these sizes);
//   * ordinary (pageable) host memory, larger: a pipeline through the pinned ring. The input is cut into
//     chunks; the CPU copies chunk i into a free pinned slot while the DMA engine moves chunk
//     i - 1 to the device, and the output comes back the same way in reverse, so the CPU copy
//     and the link run at the same time; the CPU copies are split over a small thread pool
//     (CopyPool). (HIP's own pageable copies run at ~30 GB/s host to device and ~18 GB/s back on
//     MI355X hosts; pinned DMA at ~57 GB/s each way; one copying thread ~28 GB/s.)
// A call with a device-memory argument first waits for all work queued so far on the null stream,
// like a synchronous hipMemcpy, so a producer kernel on the caller's default stream finishes
// before the transform reads.
// The output is written only after the transform has completed without error.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>

#include "../../include/onefft.h"
#include "copy_pool.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace onefft {
namespace {

constexpr size_t kChunk = size_t(4) << 20;  // bytes per pinned ring chunk
constexpr int kRing = 4;                    // chunks in a ring: 16 MiB pinned per slot
constexpr size_t kDirectBytes = size_t(8) << kMaxSinglePassLog2;  // one LDS-pass transform
constexpr size_t kRingMinBytes = kChunk;                          // smaller: HIP's pageable copies
constexpr int kMaxSlots = 64;
constexpr int kMaxDevices = 16;
constexpr int kDefaultSlots = 4;
constexpr size_t kWordBytes = 256;  // the direct buffer's completion word, on its own cache line
// how long a direct call polls its completion word before it waits on the stream instead (the
// GPU may be busy with other work); the result is the same either way. ONEFFT_POLL_US overrides it
// (0: wait on the stream at once), so tests can take the fallback.
constexpr std::chrono::microseconds kPollFor{200};
std::chrono::microseconds poll_for() {
    if (const char *env = std::getenv("ONEFFT_POLL_US")) {
        const long v = std::atol(env);
        if (v >= 0 && v <= 10000000) return std::chrono::microseconds(v);
    }
    return kPollFor;
}

// what a call needs besides the stream
enum Need : unsigned { kNeedBuf = 1, kNeedRing = 2, kNeedDirect = 4 };

unsigned needs_for_size(size_t bytes) {
    return kNeedBuf | (bytes >= kRingMinBytes ? kNeedRing : 0u) | (bytes <= kDirectBytes ? kNeedDirect : 0u);
}

// One caller's staging, held by one call at a time (busy).
struct Slot {
    std::atomic<bool> busy{false};
    hipStream_t stream = nullptr;
    hipEvent_t ev[kRing] = {};
    void *buf = nullptr;  // 2 * cap bytes: input then output
    size_t cap = 0;
    char *pinned = nullptr;      // kRing * kChunk bytes of pinned host memory
    char *direct = nullptr;      // 2 * kDirectBytes of coherent pinned host memory: input, output,
                                 // then the completion word
    char *direct_dev = nullptr;  // the same memory as the device addresses it
    uint32_t seq = 0;            // the last value the kernel was asked to store in the word

    // the stream and events; all or nothing
    bool create() {
        if (stream_create(&stream) != hipSuccess) {
            set_error("staging stream creation failed");
            return false;
        }
        bool ok = true;
        for (int i = 0; ok && i < kRing; i++) ok = event_create(&ev[i]) == hipSuccess;
        if (!ok) {
            set_error("staging event creation failed");
            release();
        }
        return ok;
    }

    bool serves(size_t bytes, unsigned need) const {
        return (!(need & kNeedBuf) || cap >= bytes) && (!(need & kNeedRing) || pinned) &&
               (!(need & kNeedDirect) || direct);
    }

    // grows the slot to serve `need` for transforms of `bytes` (allocates); a failure leaves the
    // slot as it was
    bool ensure(size_t bytes, unsigned need) {
        if ((need & kNeedRing) && !pinned) {
            void *p = nullptr;
            if (host_malloc(&p, kRing * kChunk, hipHostMallocDefault) != hipSuccess) {
                set_error("pinned staging allocation failed");
                return false;
            }
            pinned = static_cast<char *>(p);
        }
        if ((need & kNeedDirect) && !direct) {
            void *p = nullptr, *d = nullptr;
            if (host_malloc(&p, 2 * kDirectBytes + kWordBytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
                set_error("pinned staging allocation failed");
                return false;
            }
            if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
                (void)hipGetLastError();
                (void)hipHostFree(p);
                set_error("pinned staging mapping failed");
                return false;
            }
            direct = static_cast<char *>(p);
            direct_dev = static_cast<char *>(d);
            __atomic_store_n(reinterpret_cast<uint32_t *>(direct + 2 * kDirectBytes), 0u, __ATOMIC_RELEASE);
            seq = 0;
        }
        if ((need & kNeedBuf) && cap < bytes) {
            void *p = nullptr;
            if (dev_malloc(&p, 2 * bytes) != hipSuccess) {
                set_error("staging allocation failed");
                return false;
            }
            if (buf) {
                (void)hipStreamSynchronize(stream);
                (void)hipFree(buf);
            }
            buf = p;
            cap = bytes;
        }
        return true;
    }

    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (buf) (void)hipFree(buf);
        if (pinned) (void)hipHostFree(pinned);
        if (direct) (void)hipHostFree(direct);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        for (auto &e : ev) e = nullptr;
        buf = nullptr;
        cap = 0;
        pinned = direct = direct_dev = nullptr;
    }
};

// The slots of one device. Slots [0, count) exist; they are created under `grow` and published by
// the release store of count, then only claimed and returned.
struct Pool {
    std::atomic<int> count{0};
    Slot slots[kMaxSlots];
    std::mutex grow;  // slot creation, onefft_cache_prepare and onefft_cache_clear
};

// Never destroyed: the HIP runtime may already be gone when static destructors run.
Pool *g_pools[kMaxDevices];
std::mutex g_pools_mutex;

Pool *pool(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    // pools are created once and never freed; the mutex covers the creation only
    std::lock_guard<std::mutex> lk(g_pools_mutex);
    Pool *p = __atomic_load_n(&g_pools[device], __ATOMIC_RELAXED);
    if (!p) {
        p = new (std::nothrow) Pool;
        __atomic_store_n(&g_pools[device], p, __ATOMIC_RELEASE);
    }
    return p;
}

Pool *pool_fast(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    Pool *p = __atomic_load_n(&g_pools[device], __ATOMIC_ACQUIRE);
    return p ? p : pool(device);
}

int slots_wanted() {
    const char *env = std::getenv("ONEFFT_STAGING_SLOTS");
    const int v = env ? std::atoi(env) : kDefaultSlots;
    return v < 1 ? 1 : v > kMaxSlots ? kMaxSlots : v;
}

thread_local int t_hint = 0;

// A new slot, claimed by the caller; under p.grow.
Slot *add_slot(Pool &p, bool claimed) {
    const int m = p.count.load(std::memory_order_relaxed);
    if (m >= kMaxSlots) {
        set_error("every staging slot is busy");
        return nullptr;
    }
    Slot &s = p.slots[m];
    if (!s.create()) return nullptr;
    s.busy.store(claimed, std::memory_order_relaxed);
    p.count.store(m + 1, std::memory_order_release);
    return &s;
}

// A free slot of `device`, claimed by the caller (nullptr on failure). Lock-free unless every
// slot is busy.
Slot *claim(int device) {
    Pool *p = pool_fast(device);
    if (!p) {
        set_error("no staging pool for this device");
        return nullptr;
    }
    for (int attempt = 0;; attempt++) {
        const int n = p->count.load(std::memory_order_acquire);
        for (int i = 0; i < n; i++) {
            const int idx = (t_hint + i) % n;
            Slot &s = p->slots[idx];
            if (!s.busy.load(std::memory_order_relaxed) && !s.busy.exchange(true, std::memory_order_acquire)) {
                t_hint = idx;
                return &s;
            }
        }
        if (n < kMaxSlots) {
            std::lock_guard<std::mutex> lk(p->grow);
            if (p->count.load(std::memory_order_relaxed) != n) continue;  // another thread added one
            Slot *s = add_slot(*p, true);
            if (s) t_hint = n;
            return s;
        }
        // kMaxSlots calls at once: wait for one to finish
        if (attempt > 1000000) {
            set_error("every staging slot is busy");
            return nullptr;
        }
        std::this_thread::yield();
    }
}

struct Claim {
    Slot *s;
    explicit Claim(Slot *slot) : s(slot) {}
    ~Claim() {
        if (s) s->busy.store(false, std::memory_order_release);
    }
    Claim(const Claim &) = delete;
    Claim &operator=(const Claim &) = delete;
};

enum class Mem { kPageable, kPinned, kDevice };

Mem memory_kind(const void *p) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();  // plain host memory is not an error for us
        return Mem::kPageable;
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) return Mem::kDevice;
    if (attr.type == hipMemoryTypeHost) return Mem::kPinned;
    return Mem::kPageable;
}

// pageable host -> device, through the pinned ring; asynchronous on s.stream when it returns
bool stage_in(Slot &s, void *d_dst, const void *src, size_t bytes) {
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    for (size_t i = 0; i < chunks; i++) {
        const int slot = int(i % kRing);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        if (i >= size_t(kRing)) {
            const hipError_t e = hipEventSynchronize(s.ev[slot]);
            if (e != hipSuccess) {
                set_error("input copy", e);
                return false;
            }
        }
        char *pin = s.pinned + slot * kChunk;
        CopyPool::get().copy(pin, static_cast<const char *>(src) + off, len);
        hipError_t e = hipMemcpyAsync(static_cast<char *>(d_dst) + off, pin, len, hipMemcpyHostToDevice, s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.ev[slot], s.stream);
        if (e != hipSuccess) {
            set_error("input copy", e);
            return false;
        }
    }
    return true;
}

// device -> pageable host, through the pinned ring, after everything queued on s.stream;
// synchronous. Nothing is written to dst unless the queued work completed without error.
bool stage_out(Slot &s, void *dst, const void *d_src, size_t bytes) {
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    auto enqueue = [&](size_t i) {
        const int slot = int(i % kRing);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        hipError_t e = hipMemcpyAsync(s.pinned + slot * kChunk, static_cast<const char *>(d_src) + off, len,
                                      hipMemcpyDeviceToHost, s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.ev[slot], s.stream);
        return e;
    };
    for (size_t i = 0; i < chunks && i < size_t(kRing); i++) {
        const hipError_t e = enqueue(i);
        if (e != hipSuccess) {
            set_error("output copy", e);
            (void)hipStreamSynchronize(s.stream);
            return false;
        }
    }
    for (size_t i = 0; i < chunks; i++) {
        const int slot = int(i % kRing);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        const hipError_t e = hipEventSynchronize(s.ev[slot]);
        if (e != hipSuccess) {
            set_error(i == 0 ? "transform" : "output copy", e);
            (void)hipStreamSynchronize(s.stream);
            return false;
        }
        CopyPool::get().copy(static_cast<char *>(dst) + off, s.pinned + slot * kChunk, len);
        if (i + kRing < chunks) {
            const hipError_t e2 = enqueue(i + kRing);
            if (e2 != hipSuccess) {
                set_error("output copy", e2);
                (void)hipStreamSynchronize(s.stream);
                return false;
            }
        }
    }
    return true;
}

// Polls the completion word until it holds `seq` (true) or poll_for() has passed (false).
bool poll_word(const uint32_t *word, uint32_t seq) {
    const auto until = std::chrono::steady_clock::now() + poll_for();
    for (;;) {
        for (int i = 0; i < 64; i++) {
            if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) return true;
            __builtin_ia32_pause();
        }
        if (std::chrono::steady_clock::now() > until) return false;
    }
}

// One single-pass transform on s.stream whose kernel then stores a new sequence number to the
// slot's completion word; returns once the word holds it, or, past kPollFor, once the stream has
// finished (polling returns ~4 us sooner than waiting for the launch's completion signal).
bool run_with_word(Slot &s, const onefft_plan *plan, const cf *d_in, cf *d_out, bool inverse) {
    const size_t n = onefft_plan_size(plan);
    uint32_t *word = reinterpret_cast<uint32_t *>(s.direct + 2 * kDirectBytes);
    uint32_t *word_dev = reinterpret_cast<uint32_t *>(s.direct_dev + 2 * kDirectBytes);
    const uint32_t seq = ++s.seq == 0 ? ++s.seq : s.seq;  // never 0, the word's initial value
    if (execute_device(plan, d_in, d_out, 1, n, n, inverse, s.stream, nullptr, word_dev, seq) != 0) {
        (void)hipStreamSynchronize(s.stream);
        return false;
    }
    if (poll_word(word, seq)) return true;
    const hipError_t e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) {
        set_error("transform", e);
        return false;
    }
    if (__atomic_load_n(word, __ATOMIC_ACQUIRE) != seq) {
        set_error("transform completed without its completion word");
        return false;
    }
    return true;
}

int transform(Slot &s, const onefft_plan *plan, const float *in, float *out, bool inverse) {
    const size_t n = onefft_plan_size(plan);
    const size_t bytes = n * 2 * sizeof(float);
    const Mem in_kind = memory_kind(in), out_kind = memory_kind(out);
    // device buffers are used where they lie when aligned; an input that overlaps the output
    // (which the API does not allow) is still copied first, so the result stays correct
    const bool in_direct = in_kind == Mem::kDevice && ((uintptr_t)in & 7) == 0 && disjoint(in, bytes, out, bytes);
    const bool out_direct = out_kind == Mem::kDevice && ((uintptr_t)out & 7) == 0;
    // transforms of one LDS pass: the kernel stores the completion word in the coherent buffer;
    // host memory, pageable or pinned, goes through that buffer too (a CPU copy of up to 128 KiB
    // costs less than setting up two DMA transfers)
    const bool small = bytes <= kDirectBytes;
    const bool small_host = small && in_kind != Mem::kDevice && out_kind != Mem::kDevice;
    const bool small_device = small && in_direct && out_direct;
    unsigned need = 0;
    if (small_host || small_device) need = kNeedDirect;
    else if (!(in_direct && out_direct))
        need = kNeedBuf | ((in_kind == Mem::kPageable || out_kind == Mem::kPageable) && bytes >= kRingMinBytes ? kNeedRing : 0u);
    if (!s.serves(bytes, need) && !s.ensure(bytes, need)) return -1;

    if (in_kind == Mem::kDevice || out_kind == Mem::kDevice) {
        // order after the work already queued on the null stream (a synchronous copy's rule): the
        // host waits for it, which costs nothing when it is idle (recording an event on the null
        // stream for this stream to wait on cost ~16 us per call)
        const hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) {
            set_error("ordering after the null stream", e);
            return -1;
        }
    }
    if (small_host) {
        // the kernel reads the input from and writes the output to coherent host memory
        std::memcpy(s.direct, in, bytes);
        const cf *h_in = reinterpret_cast<const cf *>(s.direct_dev);
        cf *h_out = reinterpret_cast<cf *>(s.direct_dev + kDirectBytes);
        if (!run_with_word(s, plan, h_in, h_out, inverse)) return -1;
        std::memcpy(out, s.direct + kDirectBytes, bytes);
        return 0;
    }
    if (small_device)
        return run_with_word(s, plan, reinterpret_cast<const cf *>(in), reinterpret_cast<cf *>(out), inverse) ? 0 : -1;
    const cf *d_in = reinterpret_cast<const cf *>(in);
    cf *d_out = reinterpret_cast<cf *>(out);
    if (!in_direct) {
        d_in = reinterpret_cast<const cf *>(s.buf);
        if (in_kind == Mem::kPageable && bytes >= kRingMinBytes) {
            if (!stage_in(s, s.buf, in, bytes)) {
                (void)hipStreamSynchronize(s.stream);
                return -1;
            }
        } else {
            const hipError_t e = hipMemcpyAsync(s.buf, in, bytes, hipMemcpyDefault, s.stream);
            if (e != hipSuccess) {
                set_error("input copy", e);
                (void)hipStreamSynchronize(s.stream);
                return -1;
            }
        }
    }
    if (!out_direct) d_out = reinterpret_cast<cf *>(static_cast<char *>(s.buf) + s.cap);
    if (execute_device(plan, d_in, d_out, 1, n, n, inverse, s.stream) != 0) {
        (void)hipStreamSynchronize(s.stream);
        return -1;
    }
    if (!out_direct && out_kind == Mem::kPageable && bytes >= kRingMinBytes)
        return stage_out(s, out, d_out, bytes) ? 0 : -1;
    hipError_t e = hipSuccess;
    if (!out_direct) {
        // pinned or small pageable host output, or misaligned device output: one copy after the
        // transform has completed
        e = hipStreamSynchronize(s.stream);
        if (e != hipSuccess) {
            set_error("transform", e);
            return -1;
        }
        e = hipMemcpyAsync(out, d_out, bytes, hipMemcpyDefault, s.stream);
        if (e != hipSuccess) {
            set_error("output copy", e);
            (void)hipStreamSynchronize(s.stream);
            return -1;
        }
    }
    e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) {
        set_error("transform", e);
        return -1;
    }
    return 0;
}

}  // namespace

bool staging_prepare(int device, size_t n, bool buffers) {
    Pool *p = pool(device);
    if (!p) {
        set_error("no staging pool for this device");
        return false;
    }
    std::lock_guard<std::mutex> lk(p->grow);
    const int want = slots_wanted();
    while (p->count.load(std::memory_order_relaxed) < want)
        if (!add_slot(*p, false)) return false;
    if (!buffers) return true;
    const size_t bytes = n * 2 * sizeof(float);
    const unsigned need = needs_for_size(bytes);
    if (need & kNeedRing) (void)CopyPool::get();  // its helper threads start now, not in a call
    const int count = p->count.load(std::memory_order_relaxed);
    for (int i = 0; i < count; i++) {
        Slot &s = p->slots[i];
        // take the slot from whichever call holds it (prepare is not a real-time call)
        while (s.busy.exchange(true, std::memory_order_acquire)) std::this_thread::yield();
        const bool ok = s.serves(bytes, need) || s.ensure(bytes, need);
        s.busy.store(false, std::memory_order_release);
        if (!ok) return false;
    }
    return true;
}

void staging_release_all() {
    std::lock_guard<std::mutex> lk(g_pools_mutex);
    for (int d = 0; d < kMaxDevices; d++) {
        Pool *p = __atomic_load_n(&g_pools[d], __ATOMIC_RELAXED);
        if (!p) continue;
        std::lock_guard<std::mutex> lg(p->grow);
        const int count = p->count.load(std::memory_order_relaxed);
        if (!count) continue;
        DeviceGuard g(d);
        for (int i = 0; i < count; i++) p->slots[i].release();
        p->count.store(0, std::memory_order_release);
    }
}

int execute_host_call(const onefft_plan *plan, const float *in, float *out, bool inverse) {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        set_error("no HIP device");
        return -1;
    }
    Claim c(claim(device));
    if (!c.s) return -1;
    return transform(*c.s, plan, in, out, inverse);
}

}  // namespace onefft
