// Cascade fitting of one shadow cascade (device code).
//
// generate_shadow_matrices (src/shaders/compute/generate_shadow_matrices.glsl:15-110)
// with lookAt / OrthographicProjection (src/shaders/common/matrices.glsl:9-40):
// split the visible depth range {max_depth, f(0.5^p), f(0.75^p), min_depth}
// (cascade 0 is the slice at the nearest depth), unproject the slice's eight
// NDC corners, bound them by a sphere whose radius is rounded up to 1/16,
// look at its centre from shadow_cam_distance along sun_dir and fit an
// orthographic box of +-radius, depth [0, 2 * distance].
#pragma once

#include "lh_device.h"

namespace cascade {

using namespace lhd;

struct Fit {
    LhMat4 matrix, uv_matrix, view;
    float radius;
};

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

LH_DEV Fit fit(const Uniforms& u, uint32_t min_bits, uint32_t max_bits, uint32_t c) {
    const float min_depth = __uint_as_float(min_bits), max_depth = __uint_as_float(max_bits);
    const float nearp = near_plane(u);
    const float min_distance = nearp / max_depth, max_distance = nearp / min_depth;
    const float s1 = nearp / mixf(min_distance, max_distance, split_pow(0.5f, u.cascade_split_pow));
    const float s2 = nearp / mixf(min_distance, max_distance, split_pow(0.75f, u.cascade_split_pow));
    const float last = c == 0 ? max_depth : c == 1 ? max_depth : c == 2 ? s1 : s2;
    const float split = c == 0 ? max_depth : c == 1 ? s1 : c == 2 ? s2 : min_depth;
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
    Fit o;
    o.view.c[0] = LhVec4{r.x, up.x, -f.x, 0};
    o.view.c[1] = LhVec4{r.y, up.y, -f.y, 0};
    o.view.c[2] = LhVec4{r.z, up.z, -f.z, 0};
    o.view.c[3] = LhVec4{-dot3(r, eye), -dot3(up, eye), dot3(f, eye), 1};
    const float left = -radius, right = radius, bottom = -radius, top = radius, zn = 0.0f, zf = dist * 2.0f;
    const float rml = right - left, rpl = right + left, tmb = top - bottom, tpb = top + bottom, fmn = zf - zn;
    LhMat4 proj;
    proj.c[0] = LhVec4{2.0f / rml, 0, 0, 0};
    proj.c[1] = LhVec4{0, -2.0f / tmb, 0, 0};
    proj.c[2] = LhVec4{0, 0, -1.0f / fmn, 0};
    proj.c[3] = LhVec4{-(rpl / rml), -(tpb / tmb), -(zn / fmn), 1};
    o.matrix = mat4_mul(proj, o.view);
    LhMat4 bias;
    bias.c[0] = LhVec4{0.5f, 0, 0, 0};
    bias.c[1] = LhVec4{0, 0.5f, 0, 0};
    bias.c[2] = LhVec4{0, 0, 1, 0};
    bias.c[3] = LhVec4{0.5f, 0.5f, 0, 1};
    o.uv_matrix = mat4_mul(bias, o.matrix);
    o.radius = radius;
    return o;
}

LH_DEV void store(MiscStorage& m, uint32_t c, const Fit& f) {
    m.shadow_matrices[c] = f.matrix;
    m.uv_space_shadow_matrices[c] = f.uv_matrix;
    m.shadow_view_matrices[c] = f.view;
    m.shadow_sphere_radii[c] = f.radius;
}

}  // namespace cascade
