// Internal launcher interface between the C ABI (capi.hip, batch.hip) and the
// kernels.  Launchers are templated on the view source (lhd::OneView for one
// view, lhd::ViewTable for a batch) and instantiated for both.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "lh/lh.h"
#include "lh_device.h"

namespace lhk {

// Raster tunables (chosen by the autotuner; output is independent of them).
struct RasterTuning {
    uint32_t small_extent_px = 4;   // in-lane raster when the vertex bbox is within this many pixels
    uint32_t medium_extent_px = 64; // wave-cooperative raster up to this extent
    uint32_t meshlet_blocks = 8192; // persistent grid of the per-meshlet kernel (64-thread blocks)
    uint32_t big_blocks = 2048;     // persistent grid of the large-triangle kernel (256-thread blocks)
    uint32_t clip_blocks = 256;     // persistent grid of the clip kernel (64-thread blocks)
    uint32_t tile_item_entries = 64; // tiled main-view raster: binned meshlets per work item (>= 8)
    uint32_t tile_blocks = 1024;    // tiled main-view raster: persistent grid (512-thread blocks)
    uint32_t tile_raster = 1;       // main view through the tiled raster (when bin_capacity > 0)
    uint32_t shadow_small_extent_px = 0;  // the cascades' small_extent_px (0: the same)
    uint32_t shadow_medium_extent_px = 0; // the cascades' medium_extent_px (0: the same)
};

struct Tuning {
    RasterTuning raster;
    uint32_t cull_blocks = 1024;    // persistent grid of the meshlet cull (256-thread blocks)
};

// Optional per-kernel timing: when a Profiler is installed, every launcher
// records an event on the launch stream right before each kernel, so the
// interval between consecutive events is that kernel's device time.
struct Profiler {
    static constexpr int kMax = 2048;
    hipEvent_t ev[kMax];
    const char* name[kMax];
    int n = 0;
    bool created = false;
};
extern Profiler* g_profiler;
inline void mark(const char* name, hipStream_t st) {
    Profiler* p = g_profiler;
    if (p && p->n < Profiler::kMax) {
        (void)hipEventRecord(p->ev[p->n], st);  // a failed record shows up as a failed elapsed-time query
        p->name[p->n++] = name;
    }
}

template <class VS>
hipError_t view_reset(const VS& vs, hipStream_t st);
// Clears visibility buffers and shadow maps; max_pixels = largest W*H of the views.
template <class VS>
hipError_t clear_targets(const VS& vs, uint64_t max_pixels, hipStream_t st);
template <class VS>
hipError_t cull_instances(const lh_scene& s, const VS& vs, bool shadow, hipStream_t st);
// generate_shadow_matrices + cull_instances_shadows in one launch.
template <class VS>
hipError_t fit_cull_instances_shadow(const lh_scene& s, const VS& vs, hipStream_t st);
// Render targets a meshlet-cull launch clears after its culling (the frame's
// clears, moved into the latency-bound cull launches that precede the
// rasters writing those targets).
constexpr uint32_t kClearVisbuffer = 1, kClearShadowMaps = 2;
template <class VS>
hipError_t cull_meshlets(const lh_scene& s, const VS& vs, uint32_t pass0, uint32_t npasses, uint32_t blocks,
                         hipStream_t st, uint32_t clear = 0);
// Rasterizes passes [pass0, pass0 + npasses) in one set of launches (the main
// view alone, or any run of cascades).
template <class VS>
hipError_t raster(const lh_scene& s, const VS& vs, uint32_t pass0, uint32_t npasses, const RasterTuning& t,
                  hipStream_t st, bool phase2 = false);
// LH_VIEW_HIZ: depth pyramid of the phase-1 raster and the occlusion test of
// the staged candidates (appended to draws[0] for the phase-2 raster).
hipError_t hiz_cull(const lh_scene& s, const lhd::OneView& vs, uint32_t blocks, hipStream_t st);
// LH_VIEW_HIZ: rewrites the view's visibility history from its visibility buffer.
hipError_t update_history(const lh_scene& s, const lhd::OneView& vs, hipStream_t st);
// The main-view raster of one view: phase 1, and with LH_VIEW_HIZ the
// pyramid, the occlusion test and phase 2.
hipError_t raster_main(const lh_scene& s, const lhd::OneView& vs, const Tuning& t, hipStream_t st);
template <class VS>
hipError_t depth_bounds(const VS& vs, uint64_t max_pixels, hipStream_t st);
template <class VS>
hipError_t fit_cascades(const VS& vs, hipStream_t st);
// total_tiles: 8x8 tiles over all views of a batch (ignored for one view);
// params: the differentiable-parameter overrides (all NULL = scene values).
template <class VS>
hipError_t resolve(const lh_scene& s, const VS& vs, uint32_t total_tiles, const lh_resolve_params& params,
                   hipStream_t st);
// Adjoint of the single-view resolve (lh_resolve_backward); uses the first
// kBackwardScratchBytes of the view's scratch for per-block partial sums.
constexpr uint32_t kBackwardBlocks = 2048;
constexpr uint64_t kBackwardScratchBytes = uint64_t(kBackwardBlocks) * 8 * sizeof(double);
hipError_t resolve_backward(const lh_scene& s, const lhd::OneView& vs, const lh_resolve_params& params,
                            const float* grad_radiance, const lh_resolve_grads& grads, hipStream_t st);
template <class VS>
hipError_t tonemap(const VS& vs, uint64_t max_pixels, hipStream_t st);

// Batch plan: exclusive scan of the per-(view, pass) surviving meshlet counts.
hipError_t plan_segments(const lhd::ViewTable& vt, uint32_t pass0, uint32_t npasses, hipStream_t st);

}  // namespace lhk
