// Batched multi-view frames (C ABI in include/lh/lh.h; no counterpart in the
// reference, which renders one view per frame, src/rendering.cpp:28-548).
//
// A batch is a fixed set of views (cameras, each with its four cascades) over
// one scene.  lh_batch_create copies the view table to the device and plans
// the flattened resolve tiles once; every frame then runs each pass ONCE for
// all views:
//   reset + clears (blockIdx.z = view) -> instance cull -> meshlet cull ->
//   plan (device scan of per-view draw counts) -> one raster launch over the
//   concatenated draws of all views -> depth bounds -> cascade fit + shadow
//   instance cull -> cascade meshlet cull -> plan -> one raster launch over
//   all views x cascades -> one resolve launch over all views' tiles.
// The per-view arithmetic is the single-view kernels' own (same templates),
// so every view's keys, shadow depths and radiance are bit-identical to the
// per-view loop (lh_frame per view), which the batch can also run for
// comparison.  Either sequence can be captured once into a hipGraph and
// replayed; capture happens on a batch-owned stream (default priority).
#include <vector>

#include "capi_common.h"

struct lh_batch {
    lh_scene scene{};
    std::vector<lh_view> views;
    lh_view* d_views = nullptr;
    uint32_t* d_seg = nullptr;
    uint32_t* d_tiles = nullptr;
    uint32_t total_tiles = 0;
    uint64_t max_pixels = 0, max_shadow = 0;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    lhk::Profiler prof;
};

namespace {

lhd::ViewTable table(const lh_batch& b) {
    return lhd::ViewTable{b.d_views, uint32_t(b.views.size()), b.d_seg, b.d_tiles};
}

hipError_t record_batched(const lh_batch& b, hipStream_t st) {
    using namespace lhk;
    const lhd::ViewTable vt = table(b);
    const lh_scene& s = b.scene;
    hipError_t e = view_reset(vt, st);
    if (e == hipSuccess) e = clear_targets(vt, std::max(b.max_pixels, b.max_shadow), st);
    if (e == hipSuccess) e = cull_instances(s, vt, false, st);
    if (e == hipSuccess) e = cull_meshlets(s, vt, 0, 1, g_tunings[kBatchTuningClass].cull_blocks, st);
    if (e == hipSuccess) e = plan_segments(vt, 0, 1, st);
    if (e == hipSuccess) e = raster(s, vt, 0, 1, g_tunings[kBatchTuningClass].raster, st);
    if (e == hipSuccess) e = depth_bounds(vt, b.max_pixels, st);
    if (e == hipSuccess) e = fit_cull_instances_shadow(s, vt, st);
    if (e == hipSuccess) e = cull_meshlets(s, vt, 1, LH_NUM_CASCADES, g_tunings[kBatchTuningClass].cull_blocks, st);
    if (e == hipSuccess) e = plan_segments(vt, 1, LH_NUM_CASCADES, st);
    if (e == hipSuccess) e = raster(s, vt, 1, LH_NUM_CASCADES, g_tunings[kBatchTuningClass].raster, st);
    if (e == hipSuccess) e = resolve(s, vt, b.total_tiles, lh_resolve_params{}, st);
    return e;
}

hipError_t record_loop(const lh_batch& b, hipStream_t st) {
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < b.views.size() && e == hipSuccess; i++) e = lhk::record_frame(b.scene, b.views[i], st);
    return e;
}

hipError_t record(const lh_batch& b, uint32_t mode, hipStream_t st) {
    return mode == LH_BATCH_BATCHED ? record_batched(b, st) : record_loop(b, st);
}

// Teardown is best effort: errors here cannot be reported through the void
// lh_batch_destroy and leave nothing to retry.
void release(lh_batch* b) {
    for (int m = 0; m < 2; m++) {
        if (b->exec[m]) (void)hipGraphExecDestroy(b->exec[m]);
        if (b->graph[m]) (void)hipGraphDestroy(b->graph[m]);
    }
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_seg) (void)hipFree(b->d_seg);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->prof.created)
        for (int i = 0; i < lhk::Profiler::kMax; i++) (void)hipEventDestroy(b->prof.ev[i]);
    delete b;
}

}  // namespace

extern "C" {

lh_status lh_batch_create(const lh_scene* scene, const lh_view* views, uint32_t n_views, lh_batch** out) {
    if (!out || !lhc::valid_scene(scene) || !views || n_views == 0 || n_views > LH_MAX_BATCH_VIEWS)
        return LH_ERR_INVALID;
    *out = nullptr;
    for (uint32_t i = 0; i < n_views; i++)
        if (!lhc::valid_frame(scene, &views[i])) return LH_ERR_INVALID;
    lh_batch* b = new lh_batch();
    b->scene = *scene;
    b->views.assign(views, views + n_views);
    std::vector<uint32_t> tiles(n_views + 1, 0);
    for (uint32_t i = 0; i < n_views; i++) {
        const lh_view& v = views[i];
        tiles[i + 1] = tiles[i] + ((v.width + 7) / 8) * ((v.height + 7) / 8);
        b->max_pixels = std::max<uint64_t>(b->max_pixels, uint64_t(v.width) * v.height);
        b->max_shadow = std::max<uint64_t>(b->max_shadow, uint64_t(v.shadow_size) * v.shadow_size);
    }
    b->total_tiles = tiles[n_views];
    const size_t seg_n = size_t(LH_NUM_CASCADES) * n_views + 1;
    bool ok = hipMalloc(&b->d_views, sizeof(lh_view) * n_views) == hipSuccess &&
              hipMalloc(&b->d_seg, sizeof(uint32_t) * seg_n) == hipSuccess &&
              hipMalloc(&b->d_tiles, sizeof(uint32_t) * (n_views + 1)) == hipSuccess &&
              hipMemcpy(b->d_views, views, sizeof(lh_view) * n_views, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b->d_tiles, tiles.data(), sizeof(uint32_t) * (n_views + 1), hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        release(b);
        return LH_ERR_HIP;
    }
    *out = b;
    return LH_OK;
}

lh_status lh_batch_run(lh_batch* b, uint32_t mode, uint32_t use_graph, lh_stream stream) {
    if (!b || mode > LH_BATCH_BATCHED) return LH_ERR_INVALID;
    hipStream_t st = lhc::stream_of(stream);
    if (!use_graph) return lhc::hip_status(record(*b, mode, st));
    if (!b->exec[mode]) {
        hipStream_t cs = b->capture_stream;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return LH_ERR_HIP;
        hipError_t e = record(*b, mode, cs);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        if (e != hipSuccess || e2 != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return LH_ERR_HIP;
        }
        b->graph[mode] = g;
        if (hipGraphInstantiate(&b->exec[mode], g, nullptr, nullptr, 0) != hipSuccess) return LH_ERR_HIP;
    }
    return lhc::hip_status(hipGraphLaunch(b->exec[mode], st));
}

lh_status lh_batch_profiled(lh_batch* b, uint32_t mode, lh_stream stream, lh_kernel_time* out, uint32_t capacity,
                            uint32_t* count) {
    if (!b || mode > LH_BATCH_BATCHED || !out || !count) return LH_ERR_INVALID;
    return lhc::profile(b->prof, stream, out, capacity, count, [&](hipStream_t st) { return record(*b, mode, st); });
}

void lh_batch_destroy(lh_batch* b) {
    if (b) release(b);
}

}  // extern "C"
