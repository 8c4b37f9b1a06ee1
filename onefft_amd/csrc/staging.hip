// The synchronous drop-in path: onefft_forward/inverse and onefft_plan_forward/inverse with host
// or device pointers of any alignment (the reference accepts any alignment,
// include/onefft.h:61-64; its tests misalign by 4 and 12 bytes, tests/test_onefft.c:118-119).
//
// Each calling thread owns a stream, a pair of device staging buffers (grown to the largest size
// it has transformed) and a ring of pinned host chunks, all kept until the thread exits. Plans
// stay read-only, so threads never contend (reference include/onefft.h:102-104). Per argument:
//   * device memory, 8-byte aligned: transformed where it lies;
//   * device memory, misaligned: copied device to device into the staging buffer and back;
//   * pinned host memory: copied by DMA straight to and from the staging buffers;
//   * ordinary (pageable) host memory, transforms of one LDS pass (n <= 2^14, 128 KiB): copied by
//     the CPU into a per-thread coherent pinned buffer that the kernel reads and writes directly
//     over the host link (no DMA transfers, whose set-up latency dominates at these sizes);
//   * ordinary (pageable) host memory up to 4 MiB: HIP's own pageable copies (lower set-up latency
//     than the ring below at these sizes);
//   * ordinary (pageable) host memory, larger: a pipeline through the pinned ring. The input is cut into
//     chunks; the CPU copies chunk i into a free pinned slot while the DMA engine moves chunk
//     i - 1 to the device, and the output comes back the same way in reverse, so the CPU copy
//     and the link run at the same time; the CPU copies are split over a small thread pool
//     (CopyPool). (HIP's own pageable copies run at ~30 GB/s host to device and ~18 GB/s back on
//     MI355X hosts; pinned DMA at ~57 GB/s each way; one copying thread ~28 GB/s.)
// The output is written only after the transform has completed without error.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/onefft.h"
#include "copy_pool.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace onefft {
namespace {

constexpr size_t kChunk = size_t(4) << 20;  // bytes per pinned slot
constexpr int kSlots = 4;                   // slots in the ring: 16 MiB pinned per thread
constexpr size_t kDirectBytes = size_t(8) << kMaxSinglePassLog2;  // one LDS-pass transform
constexpr size_t kRingMinBytes = kChunk;                          // smaller: HIP's pageable copies


struct Staging {
    int device = -1;
    hipStream_t stream = nullptr;
    void *buf = nullptr;  // 2 * cap bytes: input then output
    size_t cap = 0;
    char *pinned = nullptr;  // kSlots * kChunk bytes of pinned host memory
    hipEvent_t ev[kSlots] = {};
    char *direct = nullptr;      // 2 * kDirectBytes of coherent pinned host memory: input, output
    char *direct_dev = nullptr;  // the same memory as the device addresses it

    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (buf) (void)hipFree(buf);
        if (pinned) (void)hipHostFree(pinned);
        if (direct) (void)hipHostFree(direct);
        direct = direct_dev = nullptr;
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : ev) e = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        buf = nullptr;
        pinned = nullptr;
        stream = nullptr;
        cap = 0;
        device = -1;
    }
    ~Staging() { release(); }
};

thread_local Staging t_stage;

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

// the pinned ring and its events; all or nothing (s.pinned is set last)
bool ring_reserve(Staging &s) {
    if (s.pinned) return true;
    for (auto &e : s.ev) {
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            e = nullptr;
            set_error("staging event creation failed");
            return false;
        }
    }
    char *p = nullptr;
    if (hipHostMalloc(reinterpret_cast<void **>(&p), kSlots * kChunk, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        set_error("pinned staging allocation failed");
        return false;
    }
    s.pinned = p;
    return true;
}

// the coherent pinned buffer pair the single-pass kernels read and write in place
bool direct_reserve(Staging &s) {
    if (s.direct) return true;
    char *p = nullptr;
    void *d = nullptr;
    if (hipHostMalloc(reinterpret_cast<void **>(&p), 2 * kDirectBytes,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        (void)hipGetLastError();
        set_error("pinned staging allocation failed");
        return false;
    }
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(p);
        set_error("pinned staging mapping failed");
        return false;
    }
    s.direct = p;
    s.direct_dev = static_cast<char *>(d);
    return true;
}

// pageable host -> device, through the pinned ring; asynchronous on s.stream when it returns
bool stage_in(Staging &s, void *d_dst, const void *src, size_t bytes) {
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    for (size_t i = 0; i < chunks; i++) {
        const int slot = int(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        if (i >= size_t(kSlots)) {
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
bool stage_out(Staging &s, void *dst, const void *d_src, size_t bytes) {
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    auto enqueue = [&](size_t i) {
        const int slot = int(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        hipError_t e = hipMemcpyAsync(s.pinned + slot * kChunk, static_cast<const char *>(d_src) + off, len,
                                      hipMemcpyDeviceToHost, s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.ev[slot], s.stream);
        return e;
    };
    for (size_t i = 0; i < chunks && i < size_t(kSlots); i++) {
        const hipError_t e = enqueue(i);
        if (e != hipSuccess) {
            set_error("output copy", e);
            (void)hipStreamSynchronize(s.stream);
            return false;
        }
    }
    for (size_t i = 0; i < chunks; i++) {
        const int slot = int(i % kSlots);
        const size_t off = i * kChunk, len = bytes - off < kChunk ? bytes - off : kChunk;
        const hipError_t e = hipEventSynchronize(s.ev[slot]);
        if (e != hipSuccess) {
            set_error(i == 0 ? "transform" : "output copy", e);
            (void)hipStreamSynchronize(s.stream);
            return false;
        }
        CopyPool::get().copy(static_cast<char *>(dst) + off, s.pinned + slot * kChunk, len);
        if (i + kSlots < chunks) {
            const hipError_t e2 = enqueue(i + kSlots);
            if (e2 != hipSuccess) {
                set_error("output copy", e2);
                (void)hipStreamSynchronize(s.stream);
                return false;
            }
        }
    }
    return true;
}

}  // namespace

hipStream_t thread_stream(int device) {
    Staging &s = t_stage;
    if (s.device != device) {
        s.release();
        s.device = device;
    }
    if (!s.stream && hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        s.stream = nullptr;
        set_error("stream creation failed");
    }
    return s.stream;
}

bool staging_reserve(int device, size_t n) {
    if (!thread_stream(device)) return false;
    Staging &s = t_stage;
    if (n && !ring_reserve(s)) return false;
    if (n && n * 2 * sizeof(float) <= kDirectBytes && !direct_reserve(s)) return false;
    const size_t bytes = n * 2 * sizeof(float);
    if (s.cap >= bytes) return true;
    if (s.buf) {
        (void)hipStreamSynchronize(s.stream);
        (void)hipFree(s.buf);
        s.buf = nullptr;
        s.cap = 0;
    }
    if (hipMalloc(&s.buf, 2 * bytes) != hipSuccess) {
        (void)hipGetLastError();
        s.buf = nullptr;
        set_error("staging allocation failed");
        return false;
    }
    s.cap = bytes;
    return true;
}

int execute_host_call(const onefft_plan *plan, const float *in, float *out, bool inverse) {
    const size_t n = onefft_plan_size(plan);
    const size_t bytes = n * 2 * sizeof(float);
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        set_error("no HIP device");
        return -1;
    }
    const Mem in_kind = memory_kind(in), out_kind = memory_kind(out);
    // device buffers are used where they lie when aligned; an input that overlaps the output
    // (which the API does not allow) is still copied first, so the result stays correct
    const bool in_direct = in_kind == Mem::kDevice && ((uintptr_t)in & 7) == 0 && disjoint(in, bytes, out, bytes);
    const bool out_direct = out_kind == Mem::kDevice && ((uintptr_t)out & 7) == 0;
    if (!staging_reserve(device, (in_direct && out_direct) ? 0 : n)) return -1;
    Staging &s = t_stage;
    if (in_kind == Mem::kPageable && out_kind == Mem::kPageable && bytes <= kDirectBytes) {
        std::memcpy(s.direct, in, bytes);
        const cf *h_in = reinterpret_cast<const cf *>(s.direct_dev);
        cf *h_out = reinterpret_cast<cf *>(s.direct_dev + kDirectBytes);
        if (execute_device(plan, h_in, h_out, 1, n, n, inverse, s.stream) != 0) {
            (void)hipStreamSynchronize(s.stream);
            return -1;
        }
        const hipError_t e = hipStreamSynchronize(s.stream);
        if (e != hipSuccess) {
            set_error("transform", e);
            return -1;
        }
        std::memcpy(out, s.direct + kDirectBytes, bytes);
        return 0;
    }
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

}  // namespace onefft
