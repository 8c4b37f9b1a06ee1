// The library's allocations: every device allocation, pinned host allocation, stream and event it
// creates goes through these wrappers, so that one switch can make a chosen one fail and tests can
// drive every error path (the reference's contract on allocation failure: -1 or NULL with `out`
// untouched, include/onefft.h:73-75,94-97; src/onefft.c:1122-1127).
//
// ONEFFT_FAIL_ALLOC (read at every wrapped call):
//   "k"    (k >= 1) the k-th wrapped call counted from when the variable last changed fails;
//   "0"    every wrapped call fails;
//   unset  none fails.
// Anything after the number (e.g. "3/b") only makes a new value, which restarts the count.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "runtime.hpp"

namespace onefft {
namespace {

std::mutex g_fail_mutex;
std::string g_fail_value;
unsigned long g_fail_count = 0;

bool fail_now() {
    const char *env = std::getenv("ONEFFT_FAIL_ALLOC");
    if (!env || !*env) return false;
    std::lock_guard<std::mutex> lk(g_fail_mutex);
    if (g_fail_value != env) {
        g_fail_value = env;
        g_fail_count = 0;
    }
    const unsigned long k = std::strtoul(env, nullptr, 10);
    ++g_fail_count;
    return k == 0 || g_fail_count == k;
}

}  // namespace

hipError_t dev_malloc(void **p, size_t bytes) {
    *p = nullptr;
    if (fail_now()) return hipErrorOutOfMemory;
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
    }
    return e;
}

hipError_t host_malloc(void **p, size_t bytes, unsigned flags) {
    *p = nullptr;
    if (fail_now()) return hipErrorOutOfMemory;
    const hipError_t e = hipHostMalloc(p, bytes, flags);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
    }
    return e;
}

hipError_t stream_create(hipStream_t *s) {
    *s = nullptr;
    if (fail_now()) return hipErrorOutOfMemory;
    const hipError_t e = hipStreamCreateWithFlags(s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *s = nullptr;
    }
    return e;
}

hipError_t event_create(hipEvent_t *ev) {
    *ev = nullptr;
    if (fail_now()) return hipErrorOutOfMemory;
    const hipError_t e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ev = nullptr;
    }
    return e;
}

}  // namespace onefft
