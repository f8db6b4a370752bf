// HIP-graph capture of the batched HomMult+relin pipeline.
//
// kbfv_mult_relin launches ~40 kernels per chunk from the host and neither
// allocates nor synchronises, so the whole call can be recorded by stream
// capture on a private stream and replayed as one graph launch.  The reference
// has no counterpart (every kernel is launched from the host on each call,
// include/KeySwitch_23.cuh:244-310).
#include "kbfv_internal.h"

struct kbfv_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int kbfv_mult_relin_graph_create(const kbfv_context *c, const uint64_t *rlk, const uint64_t *ct1,
                                            const uint64_t *ct2, uint64_t *out, int batch, void *ws,
                                            size_t ws_bytes, kbfv_graph **graph) {
    KB_REQUIRE(graph, "null argument");
    *graph = nullptr;
    KB_REQUIRE(c && rlk && ct1 && ct2 && out && ws && batch > 0, "null argument");
    hipStream_t s = nullptr;
    KB_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s);
        kb_set_error("stream capture: %s", hipGetErrorString(e));
        return KBFV_EHIP;
    }
    const int rc = kbfv_mult_relin(c, rlk, ct1, ct2, out, batch, ws, ws_bytes, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    if (rc != KBFV_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != KBFV_OK) return rc;
        kb_set_error("stream capture: %s", hipGetErrorString(e));
        return KBFV_EHIP;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        kb_set_error("graph instantiate: %s", hipGetErrorString(e));
        return KBFV_EHIP;
    }
    *graph = new kbfv_graph{g, x};
    return KBFV_OK;
}

extern "C" int kbfv_graph_launch(kbfv_graph *graph, void *stream) {
    KB_REQUIRE(graph && graph->exec, "null graph");
    KB_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return KBFV_OK;
}

extern "C" void kbfv_graph_destroy(kbfv_graph *graph) {
    if (!graph) return;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
}
