// C ABI of liblh.so (include/lh/lh.h) and the frame recorder.
//
// lh_frame issues the passes in the order of the reference's render()
// (src/rendering.cpp:79-502): reset -> instance cull -> meshlet cull ->
// visibility raster -> depth reduction -> cascade matrices -> shadow instance
// cull -> per cascade (meshlet cull, shadow raster) -> resolve (+ display
// transform).  Vulkan barriers become stream order; the reference's indirect
// dispatch sizes (reset_buffers_b/c) become persistent grids that read the
// device-side counts.
#include <cstring>

#include "kernels.h"

namespace {

lhk::Tuning g_tuning;

lh_status hip_status(hipError_t e) { return e == hipSuccess ? LH_OK : LH_ERR_HIP; }

bool valid_scene(const lh_scene* s) {
    return s && (s->num_instances == 0 || (s->instances && s->instance_meshlet_base)) && s->mesh_infos &&
           s->factors && s->meshlet_table && s->total_meshlets <= LH_MAX_SCENE_MESHLETS &&
           (s->num_textures == 0 || s->textures);
}

bool valid_view(const lh_view* v) {
    if (!v || !v->uniforms || !v->misc || !v->prefix_sums || !v->draws || !v->counters) return false;
    if (v->width == 0 || v->height == 0 || v->width > LH_MAX_VIEW_SIDE || v->height > LH_MAX_VIEW_SIDE) return false;
    if (v->shadow_size > LH_MAX_VIEW_SIDE) return false;
    if (v->instance_capacity == 0 || v->draw_capacity == 0 || v->draw_capacity > (1u << 25)) return false;
    return true;
}

bool valid_raster(const lh_view* v) {
    return v->scratch && v->scratch_bytes >= lh_scratch_bytes(v->big_capacity, v->clip_capacity) &&
           v->big_capacity > 0 && v->clip_capacity > 0;
}

hipStream_t stream_of(lh_stream s) { return static_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

uint64_t lh_scratch_bytes(uint32_t big_capacity, uint32_t clip_capacity) {
    uint64_t clip = (uint64_t(clip_capacity) * LH_CLIP_ENTRY_BYTES + 255) & ~uint64_t(255);
    return clip + uint64_t(big_capacity) * LH_BIG_ENTRY_BYTES;
}

lh_status lh_view_reset(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v)) return LH_ERR_INVALID;
    hipStream_t st = stream_of(stream);
    hipError_t e = lhk::view_reset(*v, st);
    if (e == hipSuccess && v->visbuffer)
        e = lhk::fill_u64(v->visbuffer, uint64_t(v->width) * v->height, LH_VISBUFFER_CLEAR, st);
    if (e == hipSuccess && v->shadow_maps && v->shadow_size)
        e = lhk::fill_u32(v->shadow_maps, uint64_t(LH_NUM_CASCADES) * v->shadow_size * v->shadow_size,
                          LH_SHADOW_CLEAR, st);
    return hip_status(e);
}

lh_status lh_cull_instances(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v)) return LH_ERR_INVALID;
    return hip_status(lhk::cull_instances(*s, *v, false, stream_of(stream)));
}

lh_status lh_cull_meshlets(const lh_scene* s, const lh_view* v, uint32_t pass, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v) || pass >= LH_NUM_PASSES) return LH_ERR_INVALID;
    return hip_status(lhk::cull_meshlets(*s, *v, pass, 1, g_tuning.cull_blocks, stream_of(stream)));
}

lh_status lh_raster_view(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v) || !v->visbuffer || !valid_raster(v)) return LH_ERR_INVALID;
    return hip_status(lhk::raster(*s, *v, 0, g_tuning.raster, stream_of(stream)));
}

lh_status lh_depth_bounds(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v) || !v->visbuffer) return LH_ERR_INVALID;
    return hip_status(lhk::depth_bounds(*v, stream_of(stream)));
}

lh_status lh_fit_cascades(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v)) return LH_ERR_INVALID;
    return hip_status(lhk::fit_cascades(*v, stream_of(stream)));
}

lh_status lh_cull_instances_shadow(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v)) return LH_ERR_INVALID;
    return hip_status(lhk::cull_instances(*s, *v, true, stream_of(stream)));
}

lh_status lh_raster_shadow(const lh_scene* s, const lh_view* v, uint32_t cascade, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v) || cascade >= LH_NUM_CASCADES || !v->shadow_maps || v->shadow_size == 0 ||
        !valid_raster(v))
        return LH_ERR_INVALID;
    return hip_status(lhk::raster(*s, *v, 1 + cascade, g_tuning.raster, stream_of(stream)));
}

lh_status lh_resolve(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v) || !v->visbuffer || !v->radiance || !v->shadow_maps ||
        v->shadow_size == 0)
        return LH_ERR_INVALID;
    return hip_status(lhk::resolve(*s, *v, stream_of(stream)));
}

lh_status lh_tonemap(const lh_view* v, lh_stream stream) {
    if (!v || !v->radiance || !v->ldr || v->width == 0 || v->height == 0) return LH_ERR_INVALID;
    return hip_status(lhk::tonemap(*v, stream_of(stream)));
}

lh_status lh_frame(const lh_scene* s, const lh_view* v, lh_stream stream) {
    if (!valid_scene(s) || !valid_view(v) || !v->visbuffer || !v->radiance || !v->shadow_maps ||
        v->shadow_size == 0 || !valid_raster(v))
        return LH_ERR_INVALID;
    hipStream_t st = stream_of(stream);
    lh_status r;
    if ((r = lh_view_reset(s, v, stream))) return r;
    hipError_t e = lhk::cull_instances(*s, *v, false, st);
    if (e == hipSuccess) e = lhk::cull_meshlets(*s, *v, 0, 1, g_tuning.cull_blocks, st);
    if (e == hipSuccess) e = lhk::raster(*s, *v, 0, g_tuning.raster, st);
    if (e == hipSuccess) e = lhk::depth_bounds(*v, st);
    if (e == hipSuccess) e = lhk::fit_cascades(*v, st);
    if (e == hipSuccess) e = lhk::cull_instances(*s, *v, true, st);
    if (e == hipSuccess) e = lhk::cull_meshlets(*s, *v, 1, LH_NUM_CASCADES, g_tuning.cull_blocks, st);
    for (uint32_t c = 0; c < LH_NUM_CASCADES && e == hipSuccess; c++) e = lhk::raster(*s, *v, 1 + c, g_tuning.raster, st);
    if (e == hipSuccess) e = lhk::resolve(*s, *v, st);
    return hip_status(e);
}

lh_status lh_view_status(const lh_view* v, uint32_t* host_counters, lh_stream stream) {
    if (!v || !v->counters || !host_counters) return LH_ERR_INVALID;
    hipStream_t st = stream_of(stream);
    if (hipMemcpyAsync(host_counters, v->counters, LH_NUM_COUNTERS * sizeof(uint32_t), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return LH_ERR_HIP;
    return host_counters[LH_CNT_OVERFLOW] ? LH_ERR_CAPACITY : LH_OK;
}

lh_status lh_set_tuning(const lh_tuning* t) {
    if (!t || t->small_extent_px == 0 || t->small_extent_px > 32 || t->meshlet_blocks == 0 || t->big_blocks == 0 ||
        t->clip_blocks == 0 || t->cull_blocks == 0)
        return LH_ERR_INVALID;
    g_tuning.raster.small_extent_px = t->small_extent_px;
    g_tuning.raster.meshlet_blocks = t->meshlet_blocks;
    g_tuning.raster.big_blocks = t->big_blocks;
    g_tuning.raster.clip_blocks = t->clip_blocks;
    g_tuning.cull_blocks = t->cull_blocks;
    return LH_OK;
}

lh_status lh_get_tuning(lh_tuning* t) {
    if (!t) return LH_ERR_INVALID;
    t->small_extent_px = g_tuning.raster.small_extent_px;
    t->meshlet_blocks = g_tuning.raster.meshlet_blocks;
    t->big_blocks = g_tuning.raster.big_blocks;
    t->clip_blocks = g_tuning.raster.clip_blocks;
    t->cull_blocks = g_tuning.cull_blocks;
    return LH_OK;
}

const char* lh_version(void) { return "lighthugger_amd 0.1 (gfx950)"; }

}  // extern "C"
