// Visibility-buffer shading resolve and display transform (gfx950).
//
//   k_resolve   render_geometry main()   src/shaders/render_geometry.comp:70-266
//               CalcFullBary / interpolate  src/shaders/common/vbuffer.glsl:11-119
//               BRDF + Filament terms    src/shaders/common/pbr.glsl:1-96
//   k_tonemap   display_transform        src/shaders/display_transform.comp:5-81
//
// One thread per pixel, 8x8 blocks (the reference's workgroup shape).  The
// per-pixel dependent chain key -> meshlet table -> instance -> mesh info ->
// meshlet -> micro indices -> remap -> vertices is the reference's; the key
// carries a stable meshlet id, resolved through the scene's meshlet table.
// Differences from the reference, all documented in the build's design notes:
// factors for roughness / metallic and a vec3 sun intensity (identity defaults
// reproduce the reference), a constant sky in place of the absent skybox,
// isotropic trilinear filtering in place of 16x anisotropy, and the absent
// tony-mc-mapface LUT replaced by its own x/(x+1) pre-encoding.
#include "lh_device.h"
#include "kernels.h"
#include "shading.h"

using namespace lhd;

__global__ void __launch_bounds__(64) k_resolve(lh_scene s, lh_view v) {
    const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7), y = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (x >= v.width || y >= v.height) return;
    const f3 c = shading::resolve_pixel(s, v, x, y);
    const size_t o = size_t(y) * v.width + x;
    reinterpret_cast<float4*>(v.radiance)[o] = make_float4(c.x, c.y, c.z, 1.0f);
    if (v.ldr) v.ldr[o] = shading::tonemap_pixel(c);
}

__global__ void __launch_bounds__(256) k_tonemap(lh_view v) {
    const uint64_t n = uint64_t(v.width) * v.height;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        float4 c = reinterpret_cast<const float4*>(v.radiance)[i];
        v.ldr[i] = shading::tonemap_pixel(f3{c.x, c.y, c.z});
    }
}

namespace lhk {

hipError_t resolve(const lh_scene& s, const lh_view& v, hipStream_t st) {
    dim3 grid((v.width + 7) / 8, (v.height + 7) / 8);
    hipLaunchKernelGGL(k_resolve, grid, dim3(64), 0, st, s, v);
    return hipGetLastError();
}

hipError_t tonemap(const lh_view& v, hipStream_t st) {
    uint64_t n = uint64_t(v.width) * v.height;
    uint32_t blocks = uint32_t(std::min<uint64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_tonemap, dim3(std::max(blocks, 1u)), dim3(256), 0, st, v);
    return hipGetLastError();
}

}  // namespace lhk
