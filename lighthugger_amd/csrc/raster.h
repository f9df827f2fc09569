// Device helpers shared by the raster kernels (raster.hip: per-meshlet raster
// with device-scope depth atomics; tile.hip: the LDS-binned screen-tile
// raster): the software restatement of the Vulkan rasterization rules, the
// render targets of a pass, the raster queues inside lh_view.scratch, and the
// flattened draw-index space of a launch.
#pragma once

#include <type_traits>

#include "kernels.h"
#include "lh_device.h"
#include "scratch.h"
#include "shading.h"

namespace lhr {

using namespace lhd;

constexpr float kGuard = 4.0f;
constexpr uint32_t kBadVertex = 64u;  // non-finite, or w <= 0 inside every plane

struct BigTri {
    int32_t X[3];
    int32_t Y[3];
    float z[3];
    uint32_t id;
    uint16_t xlo, ylo, xhi, yhi;  // xlo bits 13..15: pass index within the set; ylo bit 15: alpha clip
};
static_assert(sizeof(BigTri) == LH_BIG_ENTRY_BYTES, "big queue entry");

struct SV {
    int32_t X, Y;
    float z;
};

// One render target of a launch.
struct Pass {
    uint32_t pass, W, H;
    uint64_t* vis;
    uint32_t* depth;
    const MeshletReference* draws;
};

// A launch rasterizes passes [pass0, pass0 + n) of every view of its source;
// the passes of one view share that view's queues, counted at
// LH_CNT_CLIP / LH_CNT_BIG + pass0.
struct PassSet {
    uint32_t pass0, n;
    int32_t small_extent;   // subpixels: in-lane raster up to this vertex-bbox extent
    int32_t medium_extent;  // subpixels: wave-cooperative raster up to this extent (<= 64 px)
    uint32_t phase2;        // main view, second occlusion phase: draws from LH_CNT_HIZ_PHASE1 on
    uint32_t direct;        // single view: rasterize only the draws the tiled raster left on its direct
                            // list (flattened index | kBinDirect when their clip triangles are queued)
};

template <int KIND>
LH_DEV const LhMat4& view_matrix(const lh_view& v, uint32_t pass) {
    return KIND == 0 ? v.uniforms->combined_perspective_view : v.misc->shadow_matrices[pass - 1];
}

// The alpha test of one covered sample (alpha-clip geometry only).
template <int KIND>
LH_DEV bool alpha_pass(const lh_scene& s, const lh_view& v, const Pass& p, uint32_t id, int32_t px, int32_t py) {
    return shading::alpha_visible(s, view_matrix<KIND>(v, p.pass), p.W, p.H, id, px, py);
}

LH_DEV Pass make_pass(const lh_view& v, uint32_t pass) {
    Pass p;
    p.pass = pass;
    if (pass == 0) {
        p.W = v.width;
        p.H = v.height;
        p.vis = v.visbuffer;
        p.depth = nullptr;
    } else {
        p.W = p.H = v.shadow_size;
        p.vis = nullptr;
        p.depth = v.shadow_maps + size_t(pass - 1) * v.shadow_size * v.shadow_size;
    }
    p.draws = v.draws + size_t(pass) * v.draw_capacity;
    return p;
}

// Queue layout inside lh_view.scratch: clip entries {draw << 7 | triangle,
// pass index within the set | alpha clip << 31}, then (256-byte aligned) the
// big-triangle entries.
constexpr uint32_t kClipAlpha = 0x80000000u;
constexpr uint16_t kBigAlpha = 0x8000u;
struct Queues {
    uint2* clip;
    BigTri* big;
};
LH_DEV Queues queues(const lh_view& v) {
    const lhs::ScratchLayout L = lhs::scratch_layout(v);
    return {reinterpret_cast<uint2*>(static_cast<char*>(v.scratch) + L.clip),
            reinterpret_cast<BigTri*>(static_cast<char*>(v.scratch) + L.big)};
}

LH_DEV float plane_dist(f4 v, int k) {
    switch (k) {
        case 0: return v.w - v.z;
        case 1: return v.z;
        case 2: return kGuard * v.w - v.x;
        case 3: return kGuard * v.w + v.x;
        case 4: return kGuard * v.w - v.y;
        default: return kGuard * v.w + v.y;
    }
}
LH_DEV uint32_t outcode(f4 v) {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 6; k++)
        if (!(plane_dist(v, k) >= 0.0f)) c |= 1u << k;
    return c;
}
LH_DEV bool finite4(f4 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w); }

LH_DEV bool project(f4 c, uint32_t W, uint32_t H, SV& o) {
    if (!(c.w > 0.0f)) return false;
    // x / 1 is x exactly: the orthographic cascades (w == 1 at every vertex)
    // skip the three correctly rounded divisions.
    float xn = c.x, yn = c.y, zn = c.z;
    if (c.w != 1.0f) {
        xn = c.x / c.w;
        yn = c.y / c.w;
        zn = c.z / c.w;
    }
    float hw = 0.5f * float(W), hh = 0.5f * float(H);
    float xf = xn * hw + hw, yf = yn * hh + hh;
    o.X = int32_t(rintf(xf * 256.0f));
    o.Y = int32_t(rintf(yf * 256.0f));
    o.z = zn;
    return true;
}

LH_DEV int64_t edge64(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t px, int32_t py) {
    return int64_t(bx - ax) * int64_t(py - ay) - int64_t(by - ay) * int64_t(px - ax);
}
LH_DEV bool top_left(int32_t ax, int32_t ay, int32_t bx, int32_t by) {
    int32_t dx = bx - ax, dy = by - ay;
    return dy < 0 || (dy == 0 && dx > 0);
}

// Culls by facing (alpha-clip geometry is double-sided: only zero area is
// culled, src/pipelines.cpp:22-25), orients to positive area and computes the
// pixel bbox.  Every output is written on every path (the bbox first: it does
// not depend on the vertex order).
template <int KIND>
LH_DEV bool setup(SV& v0, SV& v1, SV& v2, uint32_t W, uint32_t H, int64_t& area, int32_t& xlo, int32_t& xhi,
                  int32_t& ylo, int32_t& yhi, int32_t& extx, int32_t& exty, bool two_sided = false,
                  bool* swapped = nullptr) {
    const int32_t minX = min(v0.X, min(v1.X, v2.X)), maxX = max(v0.X, max(v1.X, v2.X));
    const int32_t minY = min(v0.Y, min(v1.Y, v2.Y)), maxY = max(v0.Y, max(v1.Y, v2.Y));
    xlo = max((minX + 127) >> 8, 0);
    xhi = min((maxX - 128) >> 8, int32_t(W) - 1);
    ylo = max((minY + 127) >> 8, 0);
    yhi = min((maxY - 128) >> 8, int32_t(H) - 1);
    extx = maxX - minX;
    exty = maxY - minY;
    area = edge64(v0.X, v0.Y, v1.X, v1.Y, v2.X, v2.Y);
    if (swapped) *swapped = area < 0;
    if (two_sided ? area == 0 : (KIND == 0 ? area >= 0 : area <= 0)) return false;
    if (area < 0) {
        SV t = v1;
        v1 = v2;
        v2 = t;
        area = -area;
    }
    return xlo <= xhi && ylo <= yhi;
}

// Raster statistics (LH_VIEW_STATISTICS): covered samples and issued depth
// atomics per lane, wave-reduced into the view's counters.  Tally<false>
// compiles to nothing.
template <bool ON>
struct Tally {
    LH_DEV void hit(bool) {}
    LH_DEV void flush(const lh_view&, uint32_t) {}
};
template <>
struct Tally<true> {
    uint32_t frags = 0, atomics = 0;
    LH_DEV void hit(bool issued) {
        frags++;
        atomics += issued ? 1u : 0u;
    }
    // Called by the whole wave.
    LH_DEV void flush(const lh_view& v, uint32_t pass) {
        const uint32_t f = wave_sum_u32(frags), a = wave_sum_u32(atomics);
        if (lane_id() == 0 && f != 0) {
            atomicAdd(&v.counters[LH_CNT_FRAGMENTS + pass], f);
            atomicAdd(&v.counters[LH_CNT_ATOMICS + pass], a);
        }
        frags = atomics = 0;
    }
};

// Depth test of one covered sample; returns whether an atomic was issued.
// The atomic is issued without reading the target first: a read-before-
// atomic early-out saves atomics but makes every sample wait for a load,
// and measured slower (the raster is latency-bound, not atomic-bound).
template <int KIND>
LH_DEV bool fragment(const Pass& p, int32_t px, int32_t py, float z, uint32_t id) {
    z = fminf(fmaxf(z, 0.0f), 1.0f);
    size_t o = size_t(py) * p.W + uint32_t(px);
    if (KIND == 0) {
        if (!(z > 0.0f)) return false;  // depth GREATER against the 0.0 clear
        unsigned long long key = ((unsigned long long)(~__float_as_uint(z)) << 32) | id;
        atomicMin(reinterpret_cast<unsigned long long*>(p.vis + o), key);
    } else {
        atomicMin(p.depth + o, __float_as_uint(z));
    }
    return true;
}

// In-lane raster of a triangle whose vertex bbox is within small_extent
// (<= 32 px): every edge value fits in int32 and the edge functions are
// stepped exactly in integers.
template <int KIND, bool ST, bool ALPHA>
LH_DEV void raster_small(const lh_scene& s, const lh_view& v, const Pass& p, SV v0, SV v1, SV v2, int32_t area,
                         int32_t xlo, int32_t xhi, int32_t ylo, int32_t yhi, uint32_t id, Tally<ST>& tally) {
    const bool tl0 = top_left(v1.X, v1.Y, v2.X, v2.Y), tl1 = top_left(v2.X, v2.Y, v0.X, v0.Y),
               tl2 = top_left(v0.X, v0.Y, v1.X, v1.Y);
    const float inv_area = rcp_exact(float(area));
    const float d10 = v1.z - v0.z, d20 = v2.z - v0.z;
    // E_i(Px + 256) = E_i(Px) - 256 (Yb - Ya); E_i(Py + 256) = E_i(Py) + 256 (Xb - Xa).
    const int32_t sx0 = -(v2.Y - v1.Y) * 256, sx1 = -(v0.Y - v2.Y) * 256, sx2 = -(v1.Y - v0.Y) * 256;
    const int32_t Px0 = xlo * 256 + 128;
    for (int32_t py = ylo; py <= yhi; py++) {
        const int32_t Py = py * 256 + 128;
        int32_t e0 = (v2.X - v1.X) * (Py - v1.Y) - (v2.Y - v1.Y) * (Px0 - v1.X);
        int32_t e1 = (v0.X - v2.X) * (Py - v2.Y) - (v0.Y - v2.Y) * (Px0 - v2.X);
        int32_t e2 = (v1.X - v0.X) * (Py - v0.Y) - (v1.Y - v0.Y) * (Px0 - v0.X);
        for (int32_t px = xlo; px <= xhi; px++) {
            bool in = (e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) && (e2 > 0 || (e2 == 0 && tl2));
            if (in && (!ALPHA || alpha_pass<KIND>(s, v, p, id, px, py))) {
                float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
                tally.hit(fragment<KIND>(p, px, py, (v0.z + l1 * d10) + l2 * d20, id));
            }
            e0 += sx0;
            e1 += sx1;
            e2 += sx2;
        }
    }
}

// Wave-cooperative raster of one medium triangle (vertex bbox within 64 px,
// so every edge value fits in int32): the 64 lanes cover its pixel bbox in
// 8x8 tiles.  All arguments are wave-uniform.
template <int KIND, bool ST, bool ALPHA>
LH_DEV void raster_wave(const lh_scene& s, const lh_view& v, const Pass& p, SV v0, SV v1, SV v2, int32_t xlo,
                        int32_t xhi, int32_t ylo, int32_t yhi, uint32_t id, uint32_t lane, Tally<ST>& tally) {
    const int32_t area = (v1.X - v0.X) * (v2.Y - v0.Y) - (v1.Y - v0.Y) * (v2.X - v0.X);
    const bool tl0 = top_left(v1.X, v1.Y, v2.X, v2.Y), tl1 = top_left(v2.X, v2.Y, v0.X, v0.Y),
               tl2 = top_left(v0.X, v0.Y, v1.X, v1.Y);
    const float inv_area = rcp_exact(float(area));
    const float d10 = v1.z - v0.z, d20 = v2.z - v0.z;
    const int32_t lx = int32_t(lane & 7), ly = int32_t(lane >> 3);
    for (int32_t ty = ylo; ty <= yhi; ty += 8) {
        for (int32_t tx = xlo; tx <= xhi; tx += 8) {
            const int32_t px = tx + lx, py = ty + ly;
            if (px > xhi || py > yhi) continue;
            const int32_t Px = px * 256 + 128, Py = py * 256 + 128;
            const int32_t e0 = (v2.X - v1.X) * (Py - v1.Y) - (v2.Y - v1.Y) * (Px - v1.X);
            const int32_t e1 = (v0.X - v2.X) * (Py - v2.Y) - (v0.Y - v2.Y) * (Px - v2.X);
            const int32_t e2 = (v1.X - v0.X) * (Py - v0.Y) - (v1.Y - v0.Y) * (Px - v0.X);
            if ((e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) && (e2 > 0 || (e2 == 0 && tl2)) &&
                (!ALPHA || alpha_pass<KIND>(s, v, p, id, px, py))) {
                const float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
                tally.hit(fragment<KIND>(p, px, py, (v0.z + l1 * d10) + l2 * d20, id));
            }
        }
    }
}

LH_DEV int32_t readlane(int32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, int(l)); }
LH_DEV float readlane(float v, uint32_t l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), int(l))); }

LH_DEV void store_big(BigTri* q, const SV& v0, const SV& v1, const SV& v2, uint32_t id, int32_t xlo, int32_t xhi,
                      int32_t ylo, int32_t yhi, uint32_t pi, bool alpha) {
    BigTri b;
    b.X[0] = v0.X, b.X[1] = v1.X, b.X[2] = v2.X;
    b.Y[0] = v0.Y, b.Y[1] = v1.Y, b.Y[2] = v2.Y;
    b.z[0] = v0.z, b.z[1] = v1.z, b.z[2] = v2.z;
    b.id = id;
    b.xlo = uint16_t(uint32_t(xlo) | (pi << 13)), b.ylo = uint16_t(uint32_t(ylo) | (alpha ? kBigAlpha : 0u));
    b.xhi = uint16_t(xhi), b.yhi = uint16_t(yhi);
    uint4* dst = reinterpret_cast<uint4*>(q);
    const uint4* src = reinterpret_cast<const uint4*>(&b);
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
}

// Vertex fetch (load_vertex_data, src/shaders/rasterization.glsl:12-34, and
// calculate_world_pos, src/shaders/common/util.glsl:11-16).
LH_DEV uint32_t load_index(const MeshInfo& mi, uint32_t i) {
    if (mi.flags & MESH_INFO_FLAGS_32_BIT_INDICES) return ptr<uint32_t>(mi.indices)[i];
    return ptr<uint16_t>(mi.indices)[i];
}
LH_DEV f3 load_position(const MeshInfo& mi, uint32_t v) {
    if (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) {
        const float* q = ptr<float>(mi.positions) + 3 * size_t(v);
        return {q[0], q[1], q[2]};
    }
    const uint16_t* q = ptr<uint16_t>(mi.positions) + 3 * size_t(v);
    return {float(q[0]), float(q[1]), float(q[2])};
}

// Segment lookup of a flattened draw index g.  One view: the segments are the
// passes of the set, offsets computed from the view's counters.  A batch: the
// segments are (view, pass) pairs, offsets planned on the device beforehand.
template <class VS>
struct Segments;

template <>
struct Segments<OneView> {
    // off[i]: first flattened index of pass i; only constant indices, so the
    // array lives in registers.
    uint32_t off[LH_NUM_CASCADES + 1];
    uint32_t first;  // draws before this index were rasterized by an earlier phase
    LH_DEV Segments(const OneView& vs, const PassSet& ps) {
        const lh_view& v = vs.v;
        first = ps.phase2 ? min(v.counters[LH_CNT_HIZ_PHASE1], v.draw_capacity) : 0u;
        off[0] = 0;
#pragma unroll
        for (uint32_t i = 0; i < LH_NUM_CASCADES; i++)
            off[i + 1] = off[i] + (i < ps.n ? min(v.counters[LH_CNT_DRAWS + ps.pass0 + i], v.draw_capacity) : 0u);
        if (ps.phase2) {  // one pass: [first, count) of the main view's draws
#pragma unroll
            for (uint32_t i = 1; i <= LH_NUM_CASCADES; i++) off[i] = off[i] > first ? off[i] - first : 0u;
        }
    }
    LH_DEV uint32_t total() const { return off[LH_NUM_CASCADES]; }
    static constexpr uint32_t kChunk = 1;  // lookups are register-only: plain grid stride
    struct Cursor {};
    LH_DEV Cursor cursor(uint32_t) const { return {}; }
    LH_DEV void at(Cursor&, uint32_t g, uint32_t& view, uint32_t& pi, uint32_t& d) const { find(g, view, pi, d); }
    LH_DEV void find(uint32_t g, uint32_t& view, uint32_t& pi, uint32_t& d) const {
        // the last pass whose range starts at or before g (empty passes are skipped)
        uint32_t base = 0;
        pi = 0;
#pragma unroll
        for (uint32_t i = 1; i < LH_NUM_CASCADES; i++)
            if (g >= off[i]) {
                pi = i;
                base = off[i];
            }
        view = 0;
        d = g - base + first;  // first is 0 unless this is the main view's second phase
    }
};

template <>
struct Segments<ViewTable> {
    const uint32_t* off;
    uint32_t nseg, np;
    LH_DEV Segments(const ViewTable& vs, const PassSet& ps) : off(vs.seg_offsets), nseg(vs.n * ps.n), np(ps.n) {}
    LH_DEV uint32_t total() const { return off[nseg]; }
    // A block takes chunks of kChunk consecutive draws: one binary search over
    // the segment table per chunk, then the cursor walks forward.
    static constexpr uint32_t kChunk = 8;
    struct Cursor {
        uint32_t k, begin, end;  // segment k covers draws [begin, end)
    };
    LH_DEV Cursor cursor(uint32_t g) const {
        uint32_t lo = 0, hi = nseg;  // last k with off[k] <= g
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (off[mid] <= g)
                lo = mid;
            else
                hi = mid;
        }
        return {lo, off[lo], off[lo + 1]};
    }
    // g < total() and not below the cursor's segment
    LH_DEV void at(Cursor& c, uint32_t g, uint32_t& view, uint32_t& pi, uint32_t& d) const {
        while (g >= c.end) {
            c.k++;
            c.begin = c.end;
            c.end = off[c.k + 1];
        }
        view = c.k / np;
        pi = c.k % np;
        d = g - c.begin;
    }
};

}  // namespace lhr

namespace lhk {
// LDS-binned tiled raster of the main view (tile.hip): binning, scan, fill and
// the tile kernel; the caller then runs the per-meshlet raster over the
// direct list (PassSet::direct) and the clip / large-triangle launches.
template <bool ST>
hipError_t raster_tiles(const lh_scene& s, const lhd::OneView& vs, const lhr::PassSet& ps, const RasterTuning& t,
                        hipStream_t st);
}  // namespace lhk
