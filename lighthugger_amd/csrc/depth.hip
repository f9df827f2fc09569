// Depth bounds and cascade fitting (gfx950).
//
//   k_depth_bounds   read_depth                src/shaders/compute/read_depth.glsl:3-83
//   k_fit_cascades   generate_shadow_matrices  src/shaders/compute/generate_shadow_matrices.glsl:15-110
//                    lookAt / OrthographicProjection  src/shaders/common/matrices.glsl:9-40
//
// The reference gathers 4x4 texels of the D32 depth buffer per thread and
// reduces with subgroupMin / subgroupMax; here the depth bits come straight
// from the visibility-buffer keys (depth = ~(key >> 32)), each thread streams
// 16-byte loads, the wave reduces with cross-lane shuffles and one atomic per
// block lands in MiscStorage.  min ignores zero (sky) exactly like
// min_if_not_zero (read_depth.glsl:3-18).
#include "lh_device.h"
#include "kernels.h"

using namespace lhd;

__global__ void __launch_bounds__(256) k_depth_bounds(lh_view v) {
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

namespace {

// Integer power for pow(0.5 | 0.75, cascade_split_pow); powf for a
// non-integral exponent.
LH_DEV float split_pow(float base, float p) {
    if (p >= 0.0f && p <= 64.0f && floorf(p) == p) {
        float r = 1.0f;
        for (int i = 0; i < int(p); i++) r = r * base;
        return r;
    }
    return powf(base, p);
}

}  // namespace

// One lane per cascade.
__global__ void __launch_bounds__(64) k_fit_cascades(lh_view v) {
    const uint32_t c = threadIdx.x;
    const Uniforms& u = *v.uniforms;
    MiscStorage& m = *v.misc;
    const uint32_t min_bits = m.min_depth, max_bits = m.max_depth;
    if (c == 0 && min_bits == 0xFFFFFFFFu) atomicOr(&v.counters[LH_CNT_FLAGS], LH_FLAG_EMPTY_VIEW);
    if (c >= 4) return;
    const float min_depth = __uint_as_float(min_bits), max_depth = __uint_as_float(max_bits);
    const float nearp = near_plane(u);
    const float min_distance = nearp / max_depth, max_distance = nearp / min_depth;
    float splits[4] = {max_depth, nearp / mixf(min_distance, max_distance, split_pow(0.5f, u.cascade_split_pow)),
                       nearp / mixf(min_distance, max_distance, split_pow(0.75f, u.cascade_split_pow)), min_depth};
    const float last = c > 0 ? splits[c - 1] : max_depth, split = splits[c];
    f3 corners[8] = {{-1, 1, last}, {1, 1, last}, {1, -1, last}, {-1, -1, last},
                     {-1, 1, split}, {1, 1, split}, {1, -1, split}, {-1, -1, split}};
    for (int i = 0; i < 8; i++) {
        f4 q = mat4_mul_vec4(u.inv_perspective_view, f4{corners[i].x, corners[i].y, corners[i].z, 1.0f});
        corners[i] = f3{q.x / q.w, q.y / q.w, q.z / q.w};
    }
    f3 centre{0, 0, 0};
    for (int i = 0; i < 8; i++) centre = add(centre, corners[i]);
    centre = f3{centre.x / 8.0f, centre.y / 8.0f, centre.z / 8.0f};
    float radius = 0.0f;
    for (int i = 0; i < 8; i++) radius = fmaxf(radius, length3(sub(corners[i], centre)));
    radius = ceilf(radius * 16.0f) / 16.0f;
    const float dist = u.shadow_cam_distance;
    const f3 eye = add(mul(mk3(u.sun_dir), dist), centre);
    const f3 f = normalize3(sub(centre, eye));
    const f3 r = normalize3(cross3(f, f3{0, 1, 0}));
    const f3 up = cross3(r, f);
    LhMat4 view;
    view.c[0] = LhVec4{r.x, up.x, -f.x, 0};
    view.c[1] = LhVec4{r.y, up.y, -f.y, 0};
    view.c[2] = LhVec4{r.z, up.z, -f.z, 0};
    view.c[3] = LhVec4{-dot3(r, eye), -dot3(up, eye), dot3(f, eye), 1};
    const float left = -radius, right = radius, bottom = -radius, top = radius, zn = 0.0f, zf = dist * 2.0f;
    const float rml = right - left, rpl = right + left, tmb = top - bottom, tpb = top + bottom, fmn = zf - zn;
    LhMat4 proj;
    proj.c[0] = LhVec4{2.0f / rml, 0, 0, 0};
    proj.c[1] = LhVec4{0, -2.0f / tmb, 0, 0};
    proj.c[2] = LhVec4{0, 0, -1.0f / fmn, 0};
    proj.c[3] = LhVec4{-(rpl / rml), -(tpb / tmb), -(zn / fmn), 1};
    const LhMat4 mat = mat4_mul(proj, view);
    LhMat4 bias;
    bias.c[0] = LhVec4{0.5f, 0, 0, 0};
    bias.c[1] = LhVec4{0, 0.5f, 0, 0};
    bias.c[2] = LhVec4{0, 0, 1, 0};
    bias.c[3] = LhVec4{0.5f, 0.5f, 0, 1};
    m.shadow_matrices[c] = mat;
    m.uv_space_shadow_matrices[c] = mat4_mul(bias, mat);
    m.shadow_view_matrices[c] = view;
    m.shadow_sphere_radii[c] = radius;
}

namespace lhk {

hipError_t depth_bounds(const lh_view& v, hipStream_t st) {
    uint64_t n = uint64_t(v.width) * v.height;
    uint32_t blocks = uint32_t(std::min<uint64_t>((n / 2 + 255) / 256, 2048));
    blocks = std::max(blocks, 1u);
    hipLaunchKernelGGL(k_depth_bounds, dim3(blocks), dim3(256), 0, st, v);
    return hipGetLastError();
}

hipError_t fit_cascades(const lh_view& v, hipStream_t st) {
    hipLaunchKernelGGL(k_fit_cascades, dim3(1), dim3(64), 0, st, v);
    return hipGetLastError();
}

}  // namespace lhk
