// Internal launcher interface between the C ABI (capi.hip) and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "lh/lh.h"

namespace lhk {

// Raster tunables (chosen by the autotuner; output is independent of them).
struct RasterTuning {
    uint32_t small_extent_px = 8;   // in-lane raster when the vertex bbox is within this many pixels
    uint32_t meshlet_blocks = 4096; // persistent grid of the per-meshlet kernel (64-thread blocks)
    uint32_t big_blocks = 2048;     // persistent grid of the large-triangle kernel (256-thread blocks)
    uint32_t clip_blocks = 256;     // persistent grid of the clip kernel (64-thread blocks)
};

struct Tuning {
    RasterTuning raster;
    uint32_t cull_blocks = 1024;    // persistent grid of the meshlet cull (256-thread blocks)
};

hipError_t view_reset(const lh_view& v, hipStream_t st);
hipError_t fill_u64(uint64_t* p, uint64_t n, uint64_t value, hipStream_t st);
hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t value, hipStream_t st);
hipError_t cull_instances(const lh_scene& s, const lh_view& v, bool shadow, hipStream_t st);
hipError_t cull_meshlets(const lh_scene& s, const lh_view& v, uint32_t pass0, uint32_t npasses, uint32_t blocks,
                         hipStream_t st);
hipError_t raster(const lh_scene& s, const lh_view& v, uint32_t pass, const RasterTuning& t, hipStream_t st);
hipError_t depth_bounds(const lh_view& v, hipStream_t st);
hipError_t fit_cascades(const lh_view& v, hipStream_t st);
hipError_t resolve(const lh_scene& s, const lh_view& v, hipStream_t st);
hipError_t tonemap(const lh_view& v, hipStream_t st);

}  // namespace lhk
