// Device-side helpers shared by the HIP kernels (gfx950, wave64).
//
// Arithmetic conventions - every bit-exact result depends on them and the
// build disables FMA contraction (-ffp-contract=off) and uses IEEE fp32
// division / square root:
//   dot3(a, b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v) = v / sqrt(dot3(v, v)), component by component
//   (M*v)[r]     = ((M[0][r]*v.x + M[1][r]*v.y) + M[2][r]*v.z) + M[3][r]*v.w
//   mix(x, y, a) = x*(1-a) + y*a
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lh/lh.h"

#define LH_DEV __device__ __forceinline__
#define LH_HD __host__ __device__ __forceinline__

// Occupancy targets (waves per SIMD) of the two register-heavy kernels; a
// build may override them.  0 leaves the register budget to the compiler.
#ifndef LH_RASTER_WAVES
#define LH_RASTER_WAVES 8
#endif
#ifndef LH_RESOLVE_WAVES
#define LH_RESOLVE_WAVES 4
#endif
#define LH_WAVES_ATTR(n) __attribute__((amdgpu_waves_per_eu(n, 8)))
#if LH_RASTER_WAVES > 0
#define LH_RASTER_OCCUPANCY LH_WAVES_ATTR(LH_RASTER_WAVES)
#else
#define LH_RASTER_OCCUPANCY
#endif
#if LH_RESOLVE_WAVES > 0
#define LH_RESOLVE_OCCUPANCY LH_WAVES_ATTR(LH_RESOLVE_WAVES)
#else
#define LH_RESOLVE_OCCUPANCY
#endif

namespace lhd {

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};
struct f4 {
    float x, y, z, w;
};

LH_DEV f3 mk3(const LhVec3& v) { return {v.x, v.y, v.z}; }
LH_DEV f4 mk4(const LhVec4& v) { return {v.x, v.y, v.z, v.w}; }
LH_DEV f3 xyz(const LhVec4& v) { return {v.x, v.y, v.z}; }
LH_DEV f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LH_DEV f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LH_DEV f3 mul(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LH_DEV f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LH_DEV f3 neg(f3 a) { return {-a.x, -a.y, -a.z}; }
LH_DEV float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
LH_DEV float length3(f3 v) { return sqrtf(dot3(v, v)); }
LH_DEV f3 normalize3(f3 v) {
    float l = length3(v);
    return {v.x / l, v.y / l, v.z / l};
}
LH_DEV f3 cross3(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LH_DEV float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
LH_DEV float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
// 1.0f / a, correctly rounded, for a in [1, 2^30]: the hardware reciprocal
// (within 1 ulp) and one Newton step on fused multiply-adds (Markstein's
// reciprocal iteration), 3 instructions instead of the IEEE division's ~11.
// Equal to the IEEE quotient for every float of that range (an exhaustive
// check on the device: lhp_rcp_check in probe.hip); the raster takes it for
// 1 / area of triangles within 64 px, whose areas are integers below 2^28.
LH_DEV float rcp_exact(float a) {
    const float y = __builtin_amdgcn_rcpf(a);
    return __builtin_fmaf(__builtin_fmaf(-a, y, 1.0f), y, y);
}

LH_DEV f4 mat4_mul_vec4(const LhMat4& m, f4 v) {
    return {((m.c[0].x * v.x + m.c[1].x * v.y) + m.c[2].x * v.z) + m.c[3].x * v.w,
            ((m.c[0].y * v.x + m.c[1].y * v.y) + m.c[2].y * v.z) + m.c[3].y * v.w,
            ((m.c[0].z * v.x + m.c[1].z * v.y) + m.c[2].z * v.z) + m.c[3].z * v.w,
            ((m.c[0].w * v.x + m.c[1].w * v.y) + m.c[2].w * v.z) + m.c[3].w * v.w};
}
LH_DEV f3 mat4_point(const LhMat4& m, f3 p) {
    f4 r = mat4_mul_vec4(m, f4{p.x, p.y, p.z, 1.0f});
    return {r.x, r.y, r.z};
}
LH_DEV f3 mat3_mul_vec3(const LhMat3& m, f3 v) {
    return {(m.c[0].x * v.x + m.c[1].x * v.y) + m.c[2].x * v.z, (m.c[0].y * v.x + m.c[1].y * v.y) + m.c[2].y * v.z,
            (m.c[0].z * v.x + m.c[1].z * v.y) + m.c[2].z * v.z};
}
LH_DEV LhMat4 mat4_mul(const LhMat4& a, const LhMat4& b) {
    LhMat4 r;
    for (int c = 0; c < 4; c++) {
        f4 v = mat4_mul_vec4(a, mk4(b.c[c]));
        r.c[c] = LhVec4{v.x, v.y, v.z, v.w};
    }
    return r;
}

LH_DEV float near_plane(const Uniforms& u) { return u.perspective.c[3].z; }

// World-space bounding sphere of a mesh-space sphere under an instance
// transform: centre transformed, radius times the largest column length
// (src/shaders/common/culling.glsl:4-17).
LH_DEV void world_sphere(const Instance& in, const LhVec4& s, f3& c, float& r) {
    c = mat4_point(in.transform, f3{s.x, s.y, s.z});
    float sx = length3(f3{in.transform.c[0].x, in.transform.c[0].y, in.transform.c[0].z});
    float sy = length3(f3{in.transform.c[1].x, in.transform.c[1].y, in.transform.c[1].z});
    float sz = length3(f3{in.transform.c[2].x, in.transform.c[2].y, in.transform.c[2].z});
    r = s.w * fmaxf(fmaxf(sx, sy), sz);
}

// ---- wave64 helpers ----
LH_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
// Number of set bits of `mask` below this lane.
LH_DEV uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}
LH_DEV uint64_t ballot(bool p) { return __ballot(p); }
LH_DEV uint32_t first_lane(uint64_t mask) { return uint32_t(__builtin_ctzll(mask)); }
template <class T>
LH_DEV T shfl(T v, int src) {
    return __shfl(v, src, 64);
}
LH_DEV uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Inclusive scan of a u32 across the wave (lanes in order); inactive lanes
// must contribute 0.
LH_DEV uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = lane_id();
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= uint32_t(d)) v += o;
    }
    return v;
}
// Inclusive scans across the wave with DPP lane moves (no LDS round trip):
// shifts of 1, 2, 4, 8 lanes within each row of 16, then row 15 / row 31
// broadcasts into the rows above.  Lanes a move has no source for keep the
// identity.
template <class Op>
LH_DEV int32_t wave_scan_dpp(int32_t v, int32_t identity, Op op) {
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}
LH_DEV uint32_t wave_scan_add(uint32_t v) {
    return uint32_t(wave_scan_dpp(int32_t(v), 0, [](int32_t a, int32_t b) { return a + b; }));
}
LH_DEV int32_t wave_scan_max(int32_t v) {
    return wave_scan_dpp(v, INT32_MIN, [](int32_t a, int32_t b) { return max(a, b); });
}
// Sum of a float over the wave, returned to every lane: the DPP inclusive
// scan leaves it in lane 63.  The wave must be converged.
LH_DEV float wave_sum_f32(float v) {
    const int32_t s = wave_scan_dpp(__float_as_int(v), 0, [](int32_t a, int32_t b) {
        return __float_as_int(__int_as_float(a) + __int_as_float(b));
    });
    return __int_as_float(__builtin_amdgcn_readlane(s, 63));
}
LH_DEV uint32_t wave_sum_u32(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
LH_DEV uint32_t wave_min_u32(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}
LH_DEV uint32_t wave_max_u32(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

// Wave-aggregated atomicAdd: returns the slot of each active lane with
// `pred` set; one atomic per wave.
LH_DEV uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t m = ballot(pred);
    if (m == 0) return 0;
    uint32_t leader = first_lane(m);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, uint32_t(__popcll(m)));
    base = shfl(base, int(leader));
    return base + lanes_below(m);
}

LH_DEV const MeshInfo& mesh_info_at(uint64_t addr) { return *reinterpret_cast<const MeshInfo*>(addr); }
template <class T>
LH_DEV const T* ptr(uint64_t addr) {
    return reinterpret_cast<const T*>(addr);
}

LH_DEV uint64_t* sector(const lh_view& v, uint32_t pass) {
    return v.prefix_sums + size_t(pass) * (1 + v.instance_capacity);
}

// View sources: every kernel takes one of these and reads its view with
// get(index).  OneView carries a single view by value (index ignored);
// ViewTable reads entry `index` of a device-resident lh_view array, which is
// how one launch covers every view of a batch with the same arithmetic.
struct OneView {
    lh_view v;
    LH_HD uint32_t count() const { return 1; }
    LH_DEV const lh_view& get(uint32_t) const { return v; }
};
struct ViewTable {
    const lh_view* views;
    uint32_t n;
    // Per-batch plan arrays (device): flattened raster segments and resolve tiles.
    uint32_t* seg_offsets;        // [LH_NUM_PASSES * n + 1]
    const uint32_t* tile_offsets; // [n + 1] resolve tiles (8x8) per view, exclusive scan
    LH_HD uint32_t count() const { return n; }
    LH_DEV const lh_view& get(uint32_t i) const { return views[i]; }
};

}  // namespace lhd
