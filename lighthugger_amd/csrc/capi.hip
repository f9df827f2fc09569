// C ABI of liblh.so (include/lh/lh.h): single-view passes, the frame
// recorder and its profiled variant.
//
// lh_frame issues the passes in the order of the reference's render()
// (src/rendering.cpp:79-502): reset -> instance cull -> meshlet cull ->
// visibility raster -> depth reduction -> cascade matrices + shadow instance
// cull -> meshlet cull of the four cascades -> shadow raster of the four
// cascades -> resolve (+ display transform).  Vulkan barriers become stream
// order; the reference's indirect dispatch sizes (reset_buffers_b/c) become
// persistent grids that read the device-side counts.
#include <cstring>

#include "capi_common.h"

namespace lhk {
Profiler* g_profiler = nullptr;
Tuning g_tunings[LH_TUNING_CLASSES];
}  // namespace lhk

namespace {

lhk::Profiler g_frame_profiler;

lhd::OneView one(const lh_view* v) { return lhd::OneView{*v}; }

}  // namespace

namespace lhk {

hipError_t raster_main(const lh_scene& s, const lhd::OneView& vs, const Tuning& t, hipStream_t st) {
    hipError_t e = raster(s, vs, 0, 1, t.raster, st);
    if (e == hipSuccess && (vs.v.flags & LH_VIEW_HIZ)) {
        e = hiz_cull(s, vs, t.cull_blocks, st);
        if (e == hipSuccess) e = raster(s, vs, 0, 1, t.raster, st, true);
        if (e == hipSuccess) e = update_history(s, vs, st);
    }
    return e;
}

hipError_t record_frame(const lh_scene& s, const lh_view& v, hipStream_t st) {
    const lhd::OneView vs{v};
    const uint64_t px = uint64_t(v.width) * v.height;
    // The attachment clears (src/rendering.cpp:206-211, :384) ride on the
    // meshlet culls: the view's cull clears the visibility buffer, the
    // cascades' cull the shadow maps, each before the raster that writes them.
    hipError_t e = view_reset(vs, st);
    if (e == hipSuccess) e = cull_instances(s, vs, false, st);
    if (e == hipSuccess) e = cull_meshlets(s, vs, 0, 1, tuning_for(v).cull_blocks, st, kClearVisbuffer);
    if (e == hipSuccess) e = raster_main(s, vs, tuning_for(v), st);
    if (e == hipSuccess) e = depth_bounds(vs, px, st);
    if (e == hipSuccess) e = fit_cull_instances_shadow(s, vs, st);
    if (e == hipSuccess) e = cull_meshlets(s, vs, 1, LH_NUM_CASCADES, tuning_for(v).cull_blocks, st, kClearShadowMaps);
    if (e == hipSuccess) e = raster(s, vs, 1, LH_NUM_CASCADES, tuning_for(v).raster, st);
    if (e == hipSuccess) e = resolve(s, vs, 0, lh_resolve_params{}, st);
    return e;
}

}  // namespace lhk

extern "C" {

uint64_t lh_scratch_bytes(const lh_view* v) { return v ? lhs::scratch_layout(*v).end : 0; }

lh_status lh_view_reset(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v)) return LH_ERR_INVALID;
    hipStream_t st = lhc::stream_of(stream);
    hipError_t e = lhk::view_reset(one(v), st);
    if (e == hipSuccess && v->visbuffer && (v->shadow_maps || v->shadow_size == 0)) {
        lh_view w = *v;
        if (!w.shadow_maps) w.shadow_size = 0;
        e = lhk::clear_targets(lhd::OneView{w},
                               std::max<uint64_t>(uint64_t(w.width) * w.height, uint64_t(w.shadow_size) * w.shadow_size),
                               st);
    }
    return lhc::hip_status(e);
}

lh_status lh_cull_instances(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v)) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::cull_instances(*s, one(v), false, lhc::stream_of(stream)));
}

lh_status lh_cull_meshlets(const lh_scene* s, const lh_view* v, uint32_t pass, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v) || pass >= LH_NUM_PASSES) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::cull_meshlets(*s, one(v), pass, 1, lhk::tuning_for(*v).cull_blocks, lhc::stream_of(stream)));
}

lh_status lh_raster_view(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v) || !v->visbuffer || !lhc::valid_raster(v)) return LH_ERR_INVALID;
    if ((v->flags & LH_VIEW_HIZ) && !v->radiance) return LH_ERR_INVALID;  // the pyramid lives there
    return lhc::hip_status(lhk::raster_main(*s, one(v), lhk::tuning_for(*v), lhc::stream_of(stream)));
}

lh_status lh_depth_bounds(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v) || !v->visbuffer) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::depth_bounds(one(v), uint64_t(v->width) * v->height, lhc::stream_of(stream)));
}

lh_status lh_fit_cascades(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v)) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::fit_cascades(one(v), lhc::stream_of(stream)));
}

lh_status lh_cull_instances_shadow(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v)) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::cull_instances(*s, one(v), true, lhc::stream_of(stream)));
}

lh_status lh_raster_shadow(const lh_scene* s, const lh_view* v, uint32_t cascade, lh_stream stream) {
    if (!lhc::valid_scene(s) || !lhc::valid_view(v) || cascade >= LH_NUM_CASCADES || !v->shadow_maps ||
        v->shadow_size == 0 || !lhc::valid_raster(v))
        return LH_ERR_INVALID;
    return lhc::hip_status(lhk::raster(*s, one(v), 1 + cascade, 1, lhk::tuning_for(*v).raster, lhc::stream_of(stream)));
}

static bool valid_resolve(const lh_scene* s, const lh_view* v) {
    return lhc::valid_scene(s) && lhc::valid_view(v) && v->visbuffer && v->radiance && v->shadow_maps &&
           v->shadow_size != 0;
}

lh_status lh_resolve(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_resolve(s, v)) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::resolve(*s, one(v), 0, lh_resolve_params{}, lhc::stream_of(stream)));
}

lh_status lh_resolve_ex(const lh_scene* s, const lh_view* v, const lh_resolve_params* params, lh_stream stream) {
    if (!valid_resolve(s, v)) return LH_ERR_INVALID;
    const lh_resolve_params p = params ? *params : lh_resolve_params{};
    return lhc::hip_status(lhk::resolve(*s, one(v), 0, p, lhc::stream_of(stream)));
}

lh_status lh_resolve_backward(const lh_scene* s, const lh_view* v, const lh_resolve_params* params,
                              const float* grad_radiance, const lh_resolve_grads* grads, lh_stream stream) {
    if (!valid_resolve(s, v) || !grad_radiance || !grads) return LH_ERR_INVALID;
    if (grads->geometry && (grads->geometry_bytes == 0 || grads->geometry_bytes % 4)) return LH_ERR_INVALID;
    if (!v->scratch || v->scratch_bytes < lhk::kBackwardScratchBytes) return LH_ERR_CAPACITY;
    const lh_resolve_params p = params ? *params : lh_resolve_params{};
    return lhc::hip_status(lhk::resolve_backward(*s, one(v), p, grad_radiance, *grads, lhc::stream_of(stream)));
}

lh_status lh_tonemap(const lh_view* v, lh_stream stream) {
    if (!v || !v->radiance || !v->ldr || v->width == 0 || v->height == 0) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::tonemap(one(v), uint64_t(v->width) * v->height, lhc::stream_of(stream)));
}

lh_status lh_frame(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!lhc::valid_frame(s, v)) return LH_ERR_INVALID;
    return lhc::hip_status(lhk::record_frame(*s, *v, lhc::stream_of(stream)));
}

lh_status lh_frame_profiled(const lh_scene* s, const lh_view* v, lh_stream stream, lh_kernel_time* out,
                            uint32_t capacity, uint32_t* count) {
    if (!lhc::valid_frame(s, v) || !out || !count) return LH_ERR_INVALID;
    return lhc::profile(g_frame_profiler, stream, out, capacity, count,
                        [&](hipStream_t st) { return lhk::record_frame(*s, *v, st); });
}

lh_status lh_view_status(const lh_view* v, uint32_t* host_counters, lh_stream stream) {
    if (!v || !v->counters || !host_counters) return LH_ERR_INVALID;
    hipStream_t st = lhc::stream_of(stream);
    if (hipMemcpyAsync(host_counters, v->counters, LH_NUM_COUNTERS * sizeof(uint32_t), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return LH_ERR_HIP;
    return host_counters[LH_CNT_OVERFLOW] ? LH_ERR_CAPACITY : LH_OK;
}

lh_status lh_set_tuning_class(uint32_t cls, const lh_tuning* t) {
    if (cls >= LH_TUNING_CLASSES || !t || t->small_extent_px == 0 || t->small_extent_px > 32 ||
        t->meshlet_blocks == 0 || t->big_blocks == 0 || t->clip_blocks == 0 || t->cull_blocks == 0 ||
        t->medium_extent_px < t->small_extent_px || t->medium_extent_px > 64 ||
        (t->tile_item_entries && (t->tile_item_entries < 8 || t->tile_item_entries > 512)) || t->tile_raster > 1 ||
        t->shadow_small_extent_px > 32 || t->shadow_medium_extent_px > 64 ||
        (t->shadow_medium_extent_px &&
         t->shadow_medium_extent_px < (t->shadow_small_extent_px ? t->shadow_small_extent_px : t->small_extent_px)))
        return LH_ERR_INVALID;
    lhk::Tuning& g = lhk::g_tunings[cls];
    g.raster.small_extent_px = t->small_extent_px;
    g.raster.medium_extent_px = t->medium_extent_px;
    g.raster.tile_item_entries = t->tile_item_entries ? t->tile_item_entries : 64u;
    g.raster.tile_blocks = t->tile_blocks ? t->tile_blocks : 1024u;
    g.raster.tile_raster = t->tile_raster;
    g.raster.shadow_small_extent_px = t->shadow_small_extent_px;
    g.raster.shadow_medium_extent_px = t->shadow_medium_extent_px;
    g.raster.meshlet_blocks = t->meshlet_blocks;
    g.raster.big_blocks = t->big_blocks;
    g.raster.clip_blocks = t->clip_blocks;
    g.cull_blocks = t->cull_blocks;
    return LH_OK;
}

lh_status lh_get_tuning_class(uint32_t cls, lh_tuning* t) {
    if (cls >= LH_TUNING_CLASSES || !t) return LH_ERR_INVALID;
    const lhk::Tuning& g = lhk::g_tunings[cls];
    t->small_extent_px = g.raster.small_extent_px;
    t->medium_extent_px = g.raster.medium_extent_px;
    t->tile_item_entries = g.raster.tile_item_entries;
    t->tile_blocks = g.raster.tile_blocks;
    t->tile_raster = g.raster.tile_raster;
    t->shadow_small_extent_px = g.raster.shadow_small_extent_px;
    t->shadow_medium_extent_px = g.raster.shadow_medium_extent_px;
    t->meshlet_blocks = g.raster.meshlet_blocks;
    t->big_blocks = g.raster.big_blocks;
    t->clip_blocks = g.raster.clip_blocks;
    t->cull_blocks = g.cull_blocks;
    return LH_OK;
}

lh_status lh_set_tuning(const lh_tuning* t) {
    for (uint32_t c = 0; c < LH_TUNING_CLASSES; c++) {
        lh_status s = lh_set_tuning_class(c, t);
        if (s) return s;
    }
    return LH_OK;
}

// Class 1 (views of 0.6 to 2.5 MP, 1080p among them); lh_get_tuning_class
// reads any one class.
lh_status lh_get_tuning(lh_tuning* t) { return lh_get_tuning_class(1, t); }

uint32_t lh_tuning_class_of(uint32_t width, uint32_t height) { return lhk::tuning_class(width, height); }

const char* lh_version(void) { return "lighthugger_amd 0.2 (gfx950)"; }

}  // extern "C"
