// Two-phase occlusion culling of the main view (LH_VIEW_HIZ), device side.
//
// Phase 1 rasterizes the surviving meshlets whose screen footprint is large
// (or unbounded: touching the near plane); a one-level depth pyramid of 8x8
// pixel tiles then holds, per tile, the farthest (smallest, reverse-Z) depth
// phase 1 left there, 0 where any pixel is still empty.  Phase 2 tests every
// small meshlet against the tiles its conservative screen rectangle touches
// and rasterizes only those that could still be in front.
//
// Conservative by construction, so the final keys are those of the frame
// without it: a culled meshlet's fragments all have depth <= zmax < the
// tile minimum <= the depth phase 1 left at that pixel, so their keys
// (~bits(depth) << 32 | id) are larger and lose every atomicMin.
#pragma once

#include "lh_device.h"

namespace hiz {

using namespace lhd;

constexpr uint32_t kTile = 8;           // pixels per pyramid texel side
constexpr int32_t kMaxRectPx = 32;      // larger footprints are phase-1 occluders
constexpr float kDepthMargin = 1e-4f;   // relative slack on the meshlet's nearest depth

struct ScreenBound {
    bool valid;            // false: part of the bound is at or behind w = 0 (no rectangle)
    int32_t x0, y0, x1, y1;  // inclusive pixel rectangle, clamped to the image (x0 > x1: off screen)
    float zmax;            // largest NDC depth of the bound (nearest under reverse-Z)
};

// Bound of a world-space sphere: the 8 corners of its enclosing cube through
// `m` (clip space), viewport x_f = (W/2) x/w + W/2 as in the raster.  Every
// point of the sphere projects inside the corners' bounding rectangle and has
// NDC depth <= the corners' maximum (w is linear and positive over the cube;
// z/w = near/w decreases with w).  The rectangle is widened by one pixel on
// each side for subpixel snapping.
LH_DEV ScreenBound sphere_bound(const LhMat4& m, f3 c, float r, uint32_t W, uint32_t H) {
    ScreenBound b{false, 0, 0, -1, -1, 0.0f};
    float xmin = 3.0e38f, xmax = -3.0e38f, ymin = 3.0e38f, ymax = -3.0e38f, zmax = -3.0e38f;
    const float hw = 0.5f * float(W), hh = 0.5f * float(H);
    for (int k = 0; k < 8; k++) {
        const f3 p{c.x + ((k & 1) ? r : -r), c.y + ((k & 2) ? r : -r), c.z + ((k & 4) ? r : -r)};
        const f4 q = mat4_mul_vec4(m, f4{p.x, p.y, p.z, 1.0f});
        if (!(q.w > 1e-6f) || !isfinite(q.x) || !isfinite(q.y) || !isfinite(q.z)) return b;
        const float xf = (q.x / q.w) * hw + hw, yf = (q.y / q.w) * hh + hh, z = q.z / q.w;
        xmin = fminf(xmin, xf);
        xmax = fmaxf(xmax, xf);
        ymin = fminf(ymin, yf);
        ymax = fmaxf(ymax, yf);
        zmax = fmaxf(zmax, z);
    }
    // keep the float -> int conversion in range before clamping
    xmin = fmaxf(xmin, -1.0e6f);
    ymin = fmaxf(ymin, -1.0e6f);
    xmax = fminf(xmax, 1.0e6f);
    ymax = fminf(ymax, 1.0e6f);
    b.valid = true;
    b.x0 = max(int32_t(floorf(xmin)) - 1, 0);
    b.y0 = max(int32_t(floorf(ymin)) - 1, 0);
    b.x1 = min(int32_t(ceilf(xmax)) + 1, int32_t(W) - 1);
    b.y1 = min(int32_t(ceilf(ymax)) + 1, int32_t(H) - 1);
    b.zmax = zmax;
    return b;
}

}  // namespace hiz
