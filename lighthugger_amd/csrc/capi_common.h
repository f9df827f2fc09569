// Argument validation and profiling helpers shared by the C-ABI translation
// units (capi.hip: single view; batch.hip: batched views).
#pragma once

#include <cstring>

#include "kernels.h"
#include "scratch.h"

namespace lhk {
// Launch tunables per resolution class (lh_tuning_class_of) and for batches.
extern Tuning g_tunings[LH_TUNING_CLASSES];
inline uint32_t tuning_class(uint32_t w, uint32_t h) {
    const uint64_t px = uint64_t(w) * h;
    return px < 600000 ? 0u : px < 2500000 ? 1u : px < 9000000 ? 2u : 3u;
}
inline const Tuning& tuning_for(const lh_view& v) { return g_tunings[tuning_class(v.width, v.height)]; }
constexpr uint32_t kBatchTuningClass = 4;
// The single-view frame (render() order) on one stream.
hipError_t record_frame(const lh_scene& s, const lh_view& v, hipStream_t st);
}  // namespace lhk

namespace lhc {

inline lh_status hip_status(hipError_t e) { return e == hipSuccess ? LH_OK : LH_ERR_HIP; }
inline hipStream_t stream_of(lh_stream s) { return static_cast<hipStream_t>(s); }

inline bool valid_scene(const lh_scene* s) {
    return s && (s->num_instances == 0 || (s->instances && s->instance_meshlet_base)) && s->mesh_infos &&
           s->factors && s->meshlet_table && s->total_meshlets <= LH_MAX_SCENE_MESHLETS &&
           (s->num_textures == 0 || s->textures);
}

inline bool valid_view(const lh_view* v) {
    if (!v || !v->uniforms || !v->misc || !v->prefix_sums || !v->draws || !v->counters) return false;
    if (v->width == 0 || v->height == 0 || v->width > LH_MAX_VIEW_SIDE || v->height > LH_MAX_VIEW_SIDE) return false;
    if (v->shadow_size > LH_MAX_VIEW_SIDE) return false;
    if (v->instance_capacity == 0 || v->draw_capacity == 0 || v->draw_capacity > (1u << 25)) return false;
    return true;
}

inline bool valid_raster(const lh_view* v) {
    return v->scratch && v->scratch_bytes >= lhs::scratch_layout(*v).end && v->big_capacity > 0 &&
           v->clip_capacity > 0 && v->bin_capacity <= v->draw_capacity;
}

inline bool valid_frame(const lh_scene* s, const lh_view* v) {
    return valid_scene(s) && valid_view(v) && v->visbuffer && v->radiance && v->shadow_maps && v->shadow_size != 0 &&
           valid_raster(v);
}

// Runs `record(stream)` with the profiler installed, then converts the event
// intervals into per-kernel times.
template <class F>
lh_status profile(lhk::Profiler& p, lh_stream stream, lh_kernel_time* out, uint32_t capacity, uint32_t* count,
                  F&& record) {
    if (!p.created) {
        for (int i = 0; i < lhk::Profiler::kMax; i++)
            if (hipEventCreate(&p.ev[i]) != hipSuccess) return LH_ERR_HIP;
        p.created = true;
    }
    hipStream_t st = stream_of(stream);
    p.n = 0;
    lhk::g_profiler = &p;
    hipError_t e = record(st);
    lhk::mark("end", st);
    lhk::g_profiler = nullptr;
    if (e != hipSuccess) return LH_ERR_HIP;
    if (hipEventSynchronize(p.ev[p.n - 1]) != hipSuccess) return LH_ERR_HIP;
    uint32_t n = 0;
    for (int i = 0; i + 1 < p.n && n < capacity; i++, n++) {
        std::memset(out[n].name, 0, sizeof(out[n].name));
        std::strncpy(out[n].name, p.name[i], sizeof(out[n].name) - 1);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.ev[i], p.ev[i + 1]) != hipSuccess) return LH_ERR_HIP;
        out[n].ms = ms;
    }
    *count = n;
    return LH_OK;
}

}  // namespace lhc
