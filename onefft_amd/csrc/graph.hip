// Plan executions captured as HIP graphs: the launch sequence of one (plan, buffers, batch,
// direction) or of one ragged execution is recorded once and replayed with a single
// hipGraphLaunch. No counterpart in the reference (a CPU library has no launch cost to save).
//
// Captured work uses kernels only (the ragged queue reset is a kernel, not a memset), so a graph
// replays correctly any number of times.
#include <hip/hip_runtime.h>

#include <new>

#include "../../include/onefft.h"
#include "runtime.hpp"

struct onefft_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Records `record`'s launches on a capture stream of `device` (the plan's), which stays current
// while the graph is built.
template <typename Record>
onefft_graph *capture(int device, Record record) {
    onefft::DeviceGuard guard(device);
    if (!guard.ok) {
        onefft::set_error("cannot select the plan's device");
        return nullptr;
    }
    onefft_graph *g = new (std::nothrow) onefft_graph;
    if (!g) return nullptr;
    g->device = device;
    hipStream_t s = nullptr;
    hipError_t e = onefft::stream_create(&s);
    if (e != hipSuccess) {
        onefft::set_error("graph stream", e);
        delete g;
        return nullptr;
    }
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? record(s) : -1;
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess || rc != 0 || e2 != hipSuccess || !graph) {
        if (e != hipSuccess) onefft::set_error("graph capture", e);
        else if (e2 != hipSuccess) onefft::set_error("graph capture end", e2);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        delete g;
        return nullptr;
    }
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        onefft::set_error("graph instantiate", e);
        (void)hipGraphDestroy(graph);
        delete g;
        return nullptr;
    }
    g->graph = graph;
    return g;
}

}  // namespace

extern "C" onefft_graph *onefft_graph_create_batch(const onefft_plan *plan, const float *d_in, float *d_out,
                                                   size_t batch, size_t in_dist, size_t out_dist, int inverse) {
    if (!onefft::device_args_ok(plan, d_in, d_out, batch, in_dist, out_dist)) return nullptr;
    return capture(onefft::plan_device(plan), [&](hipStream_t s) {
        return onefft::execute_device(plan, reinterpret_cast<const onefft::cf *>(d_in), reinterpret_cast<onefft::cf *>(d_out),
                                      batch, in_dist, out_dist, inverse != 0, s);
    });
}

extern "C" onefft_graph *onefft_graph_create_ragged(const onefft_ragged_plan *plan, const float *d_in, float *d_out) {
    if (!plan) {
        onefft::set_error("null plan");
        return nullptr;
    }
    return capture(onefft::ragged_plan_device(plan),
                   [&](hipStream_t s) { return onefft_ragged_execute(plan, d_in, d_out, s); });
}

extern "C" int onefft_graph_launch(const onefft_graph *graph, struct ihipStream_t *stream) {
    if (!graph) {
        onefft::set_error("null graph");
        return -1;
    }
    onefft::DeviceGuard guard(graph->device);
    if (!guard.ok) {
        onefft::set_error("cannot select the graph's device");
        return -1;
    }
    hipError_t e = hipGraphLaunch(graph->exec, stream);
    if (e != hipSuccess) {
        onefft::set_error("graph launch", e);
        return -1;
    }
    return 0;
}

extern "C" void onefft_graph_destroy(onefft_graph *graph) {
    if (!graph) return;
    onefft::DeviceGuard guard(graph->device);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
}
