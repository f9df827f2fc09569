// Depth bounds and cascade fitting (gfx950).
//
//   k_depth_bounds   read_depth                src/shaders/compute/read_depth.glsl:3-83
//   k_fit_cascades   generate_shadow_matrices  src/shaders/compute/generate_shadow_matrices.glsl:15-110
//
// The reference gathers 4x4 texels of the D32 depth buffer per thread and
// reduces with subgroupMin / subgroupMax; here the depth bits come straight
// from the visibility-buffer keys (depth = ~(key >> 32)), each thread streams
// 16-byte loads over a slice of the buffer, the wave reduces with cross-lane
// shuffles and one atomic pair per block (at most 256 blocks per view, so the
// two same-address atomics stay cheap) lands in MiscStorage.  min ignores
// zero (sky) exactly like min_if_not_zero (read_depth.glsl:3-18).
#include "cascade.h"
#include "kernels.h"

using namespace lhd;

template <class VS>
__global__ void __launch_bounds__(256) k_depth_bounds(VS vs) {
    const lh_view& v = vs.get(blockIdx.z);
    const uint64_t n = uint64_t(v.width) * v.height;
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x * 2;
    uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 2;
    for (; i + 1 < n; i += stride) {
        ulonglong2 k = *reinterpret_cast<const ulonglong2*>(v.visbuffer + i);
        uint32_t a = ~uint32_t(k.x >> 32), b = ~uint32_t(k.y >> 32);
        if (a) mn = min(mn, a);
        if (b) mn = min(mn, b);
        mx = max(mx, max(a, b));
    }
    if (i < n) {
        uint32_t a = ~uint32_t(v.visbuffer[i] >> 32);
        if (a) mn = min(mn, a);
        mx = max(mx, a);
    }
    mn = wave_min_u32(mn);
    mx = wave_max_u32(mx);
    __shared__ uint32_t smin[4], smax[4];
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        smin[w] = mn;
        smax[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        mn = min(min(smin[0], smin[1]), min(smin[2], smin[3]));
        mx = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
        if (mn != 0xFFFFFFFFu) atomicMin(&v.misc->min_depth, mn);
        atomicMax(&v.misc->max_depth, mx);
    }
}

// One lane per cascade, one block per view.
template <class VS>
__global__ void __launch_bounds__(64) k_fit_cascades(VS vs) {
    const lh_view& v = vs.get(blockIdx.z);
    const uint32_t c = threadIdx.x;
    const uint32_t min_bits = v.misc->min_depth, max_bits = v.misc->max_depth;
    if (c == 0 && min_bits == 0xFFFFFFFFu) atomicOr(&v.counters[LH_CNT_FLAGS], LH_FLAG_EMPTY_VIEW);
    if (c >= LH_NUM_CASCADES) return;
    cascade::store(*v.misc, c, cascade::fit(*v.uniforms, min_bits, max_bits, c));
}

namespace lhk {

template <class VS>
hipError_t depth_bounds(const VS& vs, uint64_t max_pixels, hipStream_t st) {
    // ~16 keys per thread, at most one block per CU and view.
    uint32_t blocks = uint32_t(std::min<uint64_t>((max_pixels + 4095) / 4096, 256));
    blocks = std::max(blocks, 1u);
    mark("depth_bounds", st);
    hipLaunchKernelGGL(k_depth_bounds<VS>, dim3(blocks, 1, vs.count()), dim3(256), 0, st, vs);
    return hipGetLastError();
}

template <class VS>
hipError_t fit_cascades(const VS& vs, hipStream_t st) {
    mark("fit_cascades", st);
    hipLaunchKernelGGL(k_fit_cascades<VS>, dim3(1, 1, vs.count()), dim3(64), 0, st, vs);
    return hipGetLastError();
}

template hipError_t depth_bounds<OneView>(const OneView&, uint64_t, hipStream_t);
template hipError_t depth_bounds<ViewTable>(const ViewTable&, uint64_t, hipStream_t);
template hipError_t fit_cascades<OneView>(const OneView&, hipStream_t);
template hipError_t fit_cascades<ViewTable>(const ViewTable&, hipStream_t);

}  // namespace lhk
