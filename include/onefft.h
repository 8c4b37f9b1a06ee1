/*
 * OneFFT for AMD Instinct MI355X (gfx950): power-of-two, complex-to-complex, single-precision,
 * out-of-place FFT on the GPU behind the OneFFT C API.
 *
 * Drop-in: the first part of this header declares the same functions, types, macros and C++
 * wrappers as OneFFT's include/onefft.h (reference include/onefft.h:53-170), with the same
 * meaning, so code written against OneFFT compiles and runs unchanged:
 *
 *     forward:  out[k] = sum over j of in[j] * exp(-2 pi i j k / n)
 *     inverse:  out[k] = sum over j of in[j] * exp(+2 pi i j k / n)
 *
 * Data are n complex values stored as 2n interleaved floats (re, im, ...), the layout of C's
 * `float complex` and C++'s `std::complex<float>`. Nothing is normalised: forward then inverse
 * gives n times the input. n is a power of two from 1 to 2^30.
 *
 * On this implementation the transform runs on the GPU. The drop-in functions accept host
 * pointers (copied to and from the device around the transform, which then dominates the cost)
 * and device pointers (hipMalloc'd memory, transformed in place on the device); they return when
 * `out` holds the result. A call with a device-memory argument first waits, like a synchronous
 * hipMemcpy, for all work queued before it on the null (default) stream; work on other streams
 * must be synchronised by the caller first. The
 * second part of the header adds asynchronous device-pointer, batched and ragged entry points on
 * a caller-supplied HIP stream.
 *
 * Host-pointer and misaligned device-pointer calls stage data through a process-wide pool of
 * ONEFFT_STAGING_SLOTS (default 4) slots per GPU, which onefft_plan_create and
 * onefft_cache_prepare size: after them, calls of their size allocate nothing and take no lock on
 * any thread, as long as at most that many calls run at once (reference include/onefft.h:66-71).
 */
#ifndef ONEFFT_H
#define ONEFFT_H

#include <stddef.h>
#include <stdint.h>

#define ONEFFT_VERSION_MAJOR 0
#define ONEFFT_VERSION_MINOR 1
#define ONEFFT_VERSION_PATCH 0

#ifdef __cplusplus
extern "C" {
#endif

/* ==== drop-in API (reference include/onefft.h:61-116) ==================================== */

/* Transform n complex values from `in` to `out` (reference include/onefft.h:61-77).
 * The buffers must not overlap; no alignment is required. The first call for a size creates its
 * plan (device tables) and keeps it until onefft_cache_clear(); later calls find it with one
 * atomic load. Returns 0, or -1 leaving `out` untouched if n is not a power of two from 1 to
 * 2^30 or the plan could not be created. */
int onefft_forward(const float *in, float *out, size_t n);
int onefft_inverse(const float *in, float *out, size_t n);

/* Creates the plan for size n now (reference include/onefft.h:79-82) and sizes the staging pool
 * for host-pointer calls of n points, so that later onefft_forward/inverse calls of that size
 * allocate nothing and take no lock on any thread (at most ONEFFT_STAGING_SLOTS at once). Returns
 * 0, or -1 if n is invalid or on allocation failure. */
int onefft_cache_prepare(size_t n);

/* Frees every plan created by the functions above and the staging pool (reference
 * include/onefft.h:84-86). Must not run while another thread is inside any function of this
 * library. */
void onefft_cache_clear(void);

/* A plan holds the device tables for one size and serves both directions
 * (reference include/onefft.h:88-109). */
typedef struct onefft_plan onefft_plan;

/* NULL if n is not a power of two from 1 to 2^30 or if an allocation fails. The plan belongs to
 * the HIP device that is current when it is created. It also creates that device's staging pool
 * and, for n <= 2^22, sizes it for host-pointer calls of this size (larger plans leave the
 * buffers to the first such call, or to onefft_cache_prepare). */
onefft_plan *onefft_plan_create(size_t n);

/* Frees a plan; NULL is allowed. */
void onefft_plan_destroy(onefft_plan *plan);

/* Transform from `in` to `out` (host or device pointers, no alignment required, no overlap).
 * The plan is never modified: any number of threads may use one plan at once, in either
 * direction. Neither allocates nor locks once the staging pool serves the plan's size (see
 * onefft_plan_create) and at most ONEFFT_STAGING_SLOTS calls run at once. A failure leaves `out`
 * untouched and a message for onefft_last_error(). */
void onefft_plan_forward(const onefft_plan *plan, const float *in, float *out);
void onefft_plan_inverse(const onefft_plan *plan, const float *in, float *out);

/* The number of complex values a plan transforms. */
size_t onefft_plan_size(const onefft_plan *plan);

/* The kernels in use: "hip-gfx950". */
const char *onefft_implementation(void);

/* ==== GPU extensions (no counterpart in the reference) ===================================== */

/* hipStream_t without including the HIP headers: the same type. NULL is the default stream. */
struct ihipStream_t;

/* The last error of a call on this thread that failed on the GPU side, or "" if none. */
const char *onefft_last_error(void);

/* Device-pointer transform, asynchronous on `stream` (reference counterpart:
 * onefft_plan_forward/inverse, include/onefft.h:105-106, which run synchronously on the host).
 * d_in and d_out are device memory, 8-byte aligned, not overlapping (-1 if they do); d_in is
 * not modified.
 * Returns 0, or -1 (invalid arguments, launch failure; see onefft_last_error()). */
int onefft_plan_forward_device(const onefft_plan *plan, const float *d_in, float *d_out, struct ihipStream_t *stream);
int onefft_plan_inverse_device(const onefft_plan *plan, const float *d_in, float *d_out, struct ihipStream_t *stream);

/* `batch` transforms of plan's size in one launch sequence (no reference counterpart; the
 * reference transforms one array per call). Transform b reads d_in + 2*b*in_dist floats and
 * writes d_out + 2*b*out_dist floats; distances are in complex values, at least the size, so the
 * output regions do not overlap each other. The input and output ranges must not overlap: -1 if
 * they do. Asynchronous on `stream`. Returns 0 or -1. */
int onefft_plan_forward_batch(const onefft_plan *plan, const float *d_in, float *d_out, size_t batch,
                              size_t in_dist, size_t out_dist, struct ihipStream_t *stream);
int onefft_plan_inverse_batch(const onefft_plan *plan, const float *d_in, float *d_out, size_t batch,
                              size_t in_dist, size_t out_dist, struct ihipStream_t *stream);

/* ---- ragged batches: many transforms of mixed sizes in one execution ------------------------
 * (no reference counterpart: with the reference each array is its own onefft_forward/inverse
 * call, include/onefft.h:76-77). */

/* One array of a ragged batch: 2^log2n complex values (log2n <= 20) at `offset` complex values
 * from the start of both the input and the output buffer, transformed forward (inverse = 0) or
 * inverse (1). */
typedef struct {
    uint64_t offset;
    uint32_t log2n;
    uint32_t inverse;
} onefft_ragged_item;

typedef struct onefft_ragged_plan onefft_ragged_plan;

/* Validates the items (sizes, directions, no two overlapping), builds the task queue and the
 * tables on the current device. Items of the tuned batched size and up (the tuning table's ragged
 * line; 2^20 as shipped) are grouped per size and direction for the multi-pass kernels, which run
 * them at their offsets; the rest go to the persistent kernel's task queue. The plan may hold its
 * own stream and two events (to run the persistent kernel beside the groups). NULL on invalid
 * items or allocation failure. */
onefft_ragged_plan *onefft_ragged_plan_create(const onefft_ragged_item *items, size_t count);
void onefft_ragged_plan_destroy(onefft_ragged_plan *plan);

/* Complex values the input and output buffers must hold (largest offset + size), and the number
 * of tasks the persistent kernel runs in one execution (the grouped items are not counted). */
size_t onefft_ragged_extent(const onefft_ragged_plan *plan);
size_t onefft_ragged_task_count(const onefft_ragged_plan *plan);

/* Transforms every item, asynchronously on `stream`: the grouped items by multi-pass launches at
 * their offsets, the rest by one persistent-kernel launch, in the tuned order: on `stream` alone,
 * or with the plan's own second stream taking the persistent kernel or part of the groups (forked
 * from and joined back to `stream` by events, so work queued on `stream` afterwards, a graph
 * capture of the call included, sees every item done). d_in and
 * d_out are device memory, 8-byte aligned, each holding onefft_ragged_extent() complex values,
 * not overlapping (-1 if they do); d_in is not modified. A ragged plan runs one execution at a
 * time: executions of one plan must be ordered (same stream or synchronised). Returns 0 or -1. */
int onefft_ragged_execute(const onefft_ragged_plan *plan, const float *d_in, float *d_out, struct ihipStream_t *stream);

/* The same result with one device plan call per item, for comparison; creates the plan cache's
 * plan of each size on first use. */
int onefft_ragged_execute_per_item(const onefft_ragged_plan *plan, const float *d_in, float *d_out,
                                   struct ihipStream_t *stream);

/* Waits for `stream` and reports whether the work queued on it (such as onefft_ragged_execute)
 * completed: 0, or -1 with onefft_last_error() set. */
int onefft_ragged_wait(const onefft_ragged_plan *plan, struct ihipStream_t *stream);

/* ---- ragged batches with no allocation ------------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.
 *
 * Transforms the `count` items whose descriptors are at d_items (DEVICE memory, 8-byte aligned)
 * in one launch sequence on `stream`, with no plan and no allocation: a planner kernel builds the
 * task queue in a fixed device buffer of the library (about 5.2 MiB per GPU, held until the
 * process ends), then the persistent kernel of onefft_ragged_execute runs every item, with results
 * bit-identical to onefft_ragged_execute of a plan whose items all go to that kernel
 * (ONEFFT_RAGGED_BATCHED_MIN=21) and within FP32 rounding of the grouped execution. d_in and d_out (device memory, 8-byte aligned, not overlapping) hold
 * `extent` complex values; every item must lie within them and items must not overlap.
 * Limits: count <= 131072 items (else -1 here) and at most 262144 tasks (a 2^20-point item makes
 * 128 of them, a single-pass group of items 1). The device writes *d_status (4-byte aligned device
 * memory) when the work completes: 0 done, 1 the wait for the first passes timed out, 2 an item
 * invalid (log2n > 20, direction not 0/1, or beyond `extent`; nothing transformed), 3 too many
 * tasks (nothing transformed).
 * The first call on a GPU uploads the library's twiddle tables (synchronously; not under graph
 * capture); later calls only launch two kernels. Calls share the buffer per GPU: calls that overlap
 * on one GPU race, so order them (one stream, or synchronise). Returns 0 or -1. */
int onefft_ragged_execute_items(const onefft_ragged_item *d_items, size_t count, size_t extent, const float *d_in,
                                float *d_out, uint32_t *d_status, struct ihipStream_t *stream);

/* ---- HIP graphs: a plan execution recorded once, replayed with one launch ------------------- */

typedef struct onefft_graph onefft_graph;

/* Records the launches of one batched device transform (arguments as onefft_plan_*_batch) or of
 * one ragged execution, on the current device. The buffers are bound into the graph. NULL on
 * invalid arguments or capture failure. */
onefft_graph *onefft_graph_create_batch(const onefft_plan *plan, const float *d_in, float *d_out, size_t batch,
                                        size_t in_dist, size_t out_dist, int inverse);
onefft_graph *onefft_graph_create_ragged(const onefft_ragged_plan *plan, const float *d_in, float *d_out);

/* Replays the recorded launches on `stream`, asynchronously. Returns 0 or -1. A ragged graph, like
 * its plan, runs one replay at a time. */
int onefft_graph_launch(const onefft_graph *graph, struct ihipStream_t *stream);
void onefft_graph_destroy(onefft_graph *graph);

#ifdef __cplusplus
} /* extern "C" */

#include <complex>

/* C++ interface (reference include/onefft.h:121-170): the same names and behaviour. */
namespace onefft {

/* False if n is not a power of two from 1 to 2^30 or the plan could not be created. */
inline bool forward(const std::complex<float> *in, std::complex<float> *out, std::size_t n) {
    return onefft_forward(reinterpret_cast<const float *>(in), reinterpret_cast<float *>(out), n) == 0;
}
inline bool inverse(const std::complex<float> *in, std::complex<float> *out, std::size_t n) {
    return onefft_inverse(reinterpret_cast<const float *>(in), reinterpret_cast<float *>(out), n) == 0;
}

/* Owns a plan; move-only. */
class Plan {
public:
    explicit Plan(std::size_t n) : plan_(onefft_plan_create(n)) {}
    ~Plan() { onefft_plan_destroy(plan_); }
    Plan(Plan &&other) noexcept : plan_(other.plan_) { other.plan_ = nullptr; }
    Plan &operator=(Plan &&other) noexcept {
        if (this != &other) {
            onefft_plan_destroy(plan_);
            plan_ = other.plan_;
            other.plan_ = nullptr;
        }
        return *this;
    }
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;

    /* False if the size was invalid or the plan could not be created; such a plan must not be used. */
    explicit operator bool() const { return plan_ != nullptr; }
    std::size_t size() const { return onefft_plan_size(plan_); }

    void forward(const std::complex<float> *in, std::complex<float> *out) const {
        onefft_plan_forward(plan_, reinterpret_cast<const float *>(in), reinterpret_cast<float *>(out));
    }
    void inverse(const std::complex<float> *in, std::complex<float> *out) const {
        onefft_plan_inverse(plan_, reinterpret_cast<const float *>(in), reinterpret_cast<float *>(out));
    }
    void forward(const float *in, float *out) const { onefft_plan_forward(plan_, in, out); }
    void inverse(const float *in, float *out) const { onefft_plan_inverse(plan_, in, out); }

    /* The underlying C plan, for the GPU extensions above. */
    const onefft_plan *get() const { return plan_; }

private:
    onefft_plan *plan_;
};

} /* namespace onefft */
#endif /* __cplusplus */

#endif /* ONEFFT_H */
