/*
 * Harness around the reference's own CPU transform. TEST AND BENCHMARK INFRASTRUCTURE ONLY.
 *
 * oracle/Makefile compiles this file with the reference's single-file distribution
 * (dist/onefft.c of the reference repository, found through -I) included below, and renames the
 * reference's public symbols to ref_* with -D options, so that the reference can sit in one
 * process next to the library under test. Nothing of the reference is copied into this
 * repository: the recipe reads it where it lies.
 *
 * Added entry points:
 *   ref_root / ref_root_double   the reference's static twiddle functions
 *                                (src/onefft.c:115-141), for bitwise twiddle parity;
 *   ref_batch                    many transforms through one shared plan on T threads;
 *   ref_time_batch               the same, timed the way bench/bench.c:151-166 times (warm-up,
 *                                minimum over samples), for the CPU baseline.
 */
#define _GNU_SOURCE
#include "onefft.c"

#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <time.h>

void ref_root(uint64_t k, unsigned log2size, float *re, float *im) { onefft_root(k, log2size, re, im); }

void ref_root_double(uint64_t k, unsigned log2size, double *re, double *im) {
    onefft_root_double(k, log2size, re, im);
}

typedef struct {
    const ref_onefft_plan *plan;
    const float *in;
    float *out;
    size_t n, first, count, dist;
    int inverse, cpu;
    pthread_barrier_t *start;
    double seconds;
    size_t reps;
} ref_job;

static double ref_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void ref_pin(int cpu) {
    if (cpu < 0) return;
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(cpu, &set);
    pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
}

static void *ref_worker(void *arg) {
    ref_job *job = (ref_job *)arg;
    ref_pin(job->cpu);
    pthread_barrier_wait(job->start);
    double t0 = ref_now();
    for (size_t r = 0; r < job->reps; r++) {
        for (size_t b = job->first; b < job->first + job->count; b++) {
            const float *x = job->in + 2 * b * job->dist;
            float *y = job->out + 2 * b * job->dist;
            if (job->inverse) ref_onefft_plan_inverse(job->plan, x, y);
            else ref_onefft_plan_forward(job->plan, x, y);
        }
    }
    job->seconds = ref_now() - t0;
    return NULL;
}

/* The CPUs this process may run on, in order, up to max. */
static int ref_cpus(int *cpus, int max) {
    cpu_set_t set;
    int count = 0;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) {
        for (int c = 0; c < CPU_SETSIZE && count < max; c++)
            if (CPU_ISSET(c, &set)) cpus[count++] = c;
    }
    return count;
}

/* Runs `batch` transforms of n points (transform b at in + 2*b*dist floats) on `threads` threads
 * sharing one plan, each thread a contiguous share of the batch, `reps` times. Returns the wall
 * time of the slowest thread, or -1 on error. */
static double ref_run(const float *in, float *out, size_t n, size_t batch, size_t dist, int inverse,
                      int threads, size_t reps) {
    ref_onefft_plan *plan = ref_onefft_plan_create(n);
    if (!plan || threads < 1) return -1;
    int cpus[1024];
    int ncpu = ref_cpus(cpus, 1024);
    pthread_t tid[256];
    ref_job jobs[256];
    if (threads > 256) threads = 256;
    pthread_barrier_t start;
    pthread_barrier_init(&start, NULL, (unsigned)threads);
    size_t per = batch / (size_t)threads, extra = batch % (size_t)threads, first = 0;
    for (int t = 0; t < threads; t++) {
        size_t count = per + ((size_t)t < extra ? 1 : 0);
        jobs[t] = (ref_job){plan, in, out, n, first, count, dist, inverse, ncpu >= threads ? cpus[t] : -1,
                            &start, 0.0, reps};
        first += count;
        pthread_create(&tid[t], NULL, ref_worker, &jobs[t]);
    }
    double worst = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(tid[t], NULL);
        if (jobs[t].seconds > worst) worst = jobs[t].seconds;
    }
    pthread_barrier_destroy(&start);
    ref_onefft_plan_destroy(plan);
    return worst;
}

int ref_batch(const float *in, float *out, size_t n, size_t batch, size_t dist, int inverse, int threads) {
    return ref_run(in, out, n, batch, dist, inverse, threads, 1) < 0 ? -1 : 0;
}

/* Seconds for one pass over the batch: 0.5 s of warm-up work, then the repetition count is
 * calibrated so one sample takes about sample_seconds, and the minimum over `samples` samples
 * is returned (bench/bench.c:115-118, 151-166). */
double ref_time_batch(const float *in, float *out, size_t n, size_t batch, size_t dist, int inverse,
                      int threads, double sample_seconds, int samples) {
    double t0 = ref_now();
    while (ref_now() - t0 < 0.5) {
        if (ref_run(in, out, n, batch, dist, inverse, threads, 1) < 0) return -1;
    }
    size_t reps = 1;
    for (;;) {
        double dt = ref_run(in, out, n, batch, dist, inverse, threads, reps);
        if (dt < 0) return -1;
        if (dt >= 0.5 * sample_seconds || reps >= ((size_t)1 << 30)) break;
        reps *= 2;
    }
    double best = 1e300;
    for (int s = 0; s < samples; s++) {
        double dt = ref_run(in, out, n, batch, dist, inverse, threads, reps) / (double)reps;
        if (dt < best) best = dt;
    }
    return best;
}

int ref_online_cpus(void) {
    int cpus[1024];
    return ref_cpus(cpus, 1024);
}

/* ---- ragged batches: every item through the reference's cache API ---------------------------- */

typedef struct {
    uint64_t offset; /* complex values */
    uint32_t log2n;
    uint32_t inverse;
} ref_item;

typedef struct {
    const float *in;
    float *out;
    const ref_item *items;
    size_t count;
    size_t *next;
    int cpu;
    pthread_barrier_t *start;
    int failures;
} ref_ragged_job;

static void *ref_ragged_worker(void *arg) {
    ref_ragged_job *job = (ref_ragged_job *)arg;
    ref_pin(job->cpu);
    pthread_barrier_wait(job->start);
    for (;;) {
        size_t i = __atomic_fetch_add(job->next, 1, __ATOMIC_RELAXED);
        if (i >= job->count) break;
        const ref_item *it = &job->items[i];
        const float *x = job->in + 2 * it->offset;
        float *y = job->out + 2 * it->offset;
        size_t n = (size_t)1 << it->log2n;
        int status = it->inverse ? ref_onefft_inverse(x, y, n) : ref_onefft_forward(x, y, n);
        if (status) job->failures++;
    }
    return NULL;
}

/* Transforms every item on `threads` threads pulling items from an atomic counter, after
 * ref_onefft_cache_prepare for every size used. Returns the seconds of the fastest of `reps`
 * passes over the items, or -1 on error. */
double ref_ragged(const float *in, float *out, const ref_item *items, size_t count, int threads, int reps) {
    for (size_t i = 0; i < count; i++)
        if (ref_onefft_cache_prepare((size_t)1 << items[i].log2n)) return -1;
    int cpus[1024];
    int ncpu = ref_cpus(cpus, 1024);
    if (threads > 256) threads = 256;
    if (threads < 1) threads = 1;
    double best = 1e300;
    for (int r = 0; r < (reps < 1 ? 1 : reps); r++) {
        pthread_t tid[256];
        ref_ragged_job jobs[256];
        size_t next = 0;
        pthread_barrier_t start;
        pthread_barrier_init(&start, NULL, (unsigned)threads + 1);
        for (int t = 0; t < threads; t++) {
            jobs[t] = (ref_ragged_job){in, out, items, count, &next, ncpu >= threads ? cpus[t] : -1, &start, 0};
            pthread_create(&tid[t], NULL, ref_ragged_worker, &jobs[t]);
        }
        pthread_barrier_wait(&start);
        double t0 = ref_now();
        int failures = 0;
        for (int t = 0; t < threads; t++) {
            pthread_join(tid[t], NULL);
            failures += jobs[t].failures;
        }
        double dt = ref_now() - t0;
        pthread_barrier_destroy(&start);
        if (failures) return -1;
        if (dt < best) best = dt;
    }
    return best;
}
