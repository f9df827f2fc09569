// lh_frame_graph (include/lh/lh.h): the single-view frame of lh_frame
// (lhk::record_frame, the reference's render() order, src/rendering.cpp:79-502)
// captured once into a hipGraph and replayed with one launch.  The recorded
// passes clear their targets and counters with kernels, not memset nodes, so
// a replay starts from the same state as a direct frame.
#include "capi_common.h"

struct lh_frame_graph {
    lh_scene scene{};
    lh_view view{};
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Best effort: nothing is left to retry when teardown fails.
void release(lh_frame_graph* g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->capture_stream) (void)hipStreamDestroy(g->capture_stream);
    delete g;
}

}  // namespace

extern "C" {

lh_status lh_frame_graph_create(const lh_scene* s, const lh_view* v, lh_frame_graph** out) {
    if (!out) return LH_ERR_INVALID;
    *out = nullptr;
    if (!lhc::valid_frame(s, v)) return LH_ERR_INVALID;
    lh_frame_graph* g = new lh_frame_graph();
    g->scene = *s;
    g->view = *v;
    if (hipStreamCreateWithFlags(&g->capture_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamBeginCapture(g->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        release(g);
        return LH_ERR_HIP;
    }
    const hipError_t e = lhk::record_frame(g->scene, g->view, g->capture_stream);
    const hipError_t e2 = hipStreamEndCapture(g->capture_stream, &g->graph);
    if (e != hipSuccess || e2 != hipSuccess || !g->graph ||
        hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        release(g);
        return LH_ERR_HIP;
    }
    *out = g;
    return LH_OK;
}

lh_status lh_frame_graph_run(lh_frame_graph* g, lh_stream stream) {
    if (!g || !g->exec) return LH_ERR_INVALID;
    return lhc::hip_status(hipGraphLaunch(g->exec, lhc::stream_of(stream)));
}

void lh_frame_graph_destroy(lh_frame_graph* g) {
    if (g) release(g);
}

}  // extern "C"
