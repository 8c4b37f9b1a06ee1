// The synchronous drop-in path: onefft_forward/inverse and onefft_plan_forward/inverse with host
// or device pointers of any alignment (the reference accepts any alignment,
// include/onefft.h:61-64; its tests misalign by 4 and 12 bytes, tests/test_onefft.c:118-119).
//
// Each calling thread owns a stream and a pair of device staging buffers, grown to the largest
// size it has transformed and kept until the thread exits. Host data (and device data that is not
// 8-byte aligned) is copied into the staging buffers, transformed there, and copied back; device
// data that is aligned is transformed where it lies. Plans stay read-only, so threads never
// contend (reference include/onefft.h:102-104).
#include <hip/hip_runtime.h>

#include "../../include/onefft.h"
#include "runtime.hpp"

namespace onefft {
namespace {

struct Staging {
    int device = -1;
    hipStream_t stream = nullptr;
    void *buf = nullptr;  // 2 * cap bytes: input then output
    size_t cap = 0;

    void release() {
        if (buf) (void)hipFree(buf);
        if (stream) (void)hipStreamDestroy(stream);
        buf = nullptr;
        stream = nullptr;
        cap = 0;
        device = -1;
    }
    ~Staging() { release(); }
};

thread_local Staging t_stage;

bool is_device_pointer(const void *p) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();  // plain host memory is not an error for us
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
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
    const bool in_direct = is_device_pointer(in) && ((uintptr_t)in & 7) == 0;
    const bool out_direct = is_device_pointer(out) && ((uintptr_t)out & 7) == 0;
    if (!staging_reserve(device, (in_direct && out_direct) ? 0 : n)) return -1;
    Staging &s = t_stage;
    const cf *d_in = reinterpret_cast<const cf *>(in);
    cf *d_out = reinterpret_cast<cf *>(out);
    hipError_t e;
    if (!in_direct) {
        d_in = reinterpret_cast<const cf *>(s.buf);
        e = hipMemcpyAsync(s.buf, in, bytes, hipMemcpyDefault, s.stream);
        if (e != hipSuccess) {
            set_error("input copy", e);
            return -1;
        }
    }
    if (!out_direct) d_out = reinterpret_cast<cf *>(static_cast<char *>(s.buf) + s.cap);
    if (execute_device(plan, d_in, d_out, 1, n, n, inverse, s.stream) != 0) {
        (void)hipStreamSynchronize(s.stream);
        return -1;
    }
    if (!out_direct) {
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
