// Software rasterizer of the GPU-driven frame (gfx950, wave64).
//
// Replaces the reference's fixed-function raster of its meshlet draws:
//   visbuffer_opaque_vertex/pixel   src/shaders/rasterization.glsl:38-58
//   shadowmap_opaque_vertex         src/shaders/rasterization.glsl:101-111
//   pipeline state                  src/pipelines.cpp:10-77 (cull back / front,
//                                   depth GREATER / LESS, CCW front faces)
//   draws                           src/rendering.cpp:235-242, :405-412
// with the Vulkan rasterization rules restated for software:
//   - clip to 0 <= z <= w and to the guard band |x|,|y| <= 4w
//     (Sutherland-Hodgman, intersections computed from the inside vertex);
//   - x_f = (W/2) x/w + W/2 (same for y), snapped to 8 subpixel bits (rint);
//   - pixel centres (x + 0.5, y + 0.5); exact integer edge functions;
//     top-left fill rule in y-down framebuffer coordinates;
//   - depth = z0 + l1 (z1 - z0) + l2 (z2 - z0), l_i = E_i / area, clamped to [0, 1].
// Main view: 64-bit atomicMin of ~bits(z) << 32 | meshlet << 7 | tri (nearest
// wins under reverse-Z); shadows: 32-bit atomicMin of bits(z).
//
// One launch covers a pass set - the main view, or all four cascades - whose
// draw lists are concatenated into one index space, so the persistent grid is
// balanced across cascades of very different load.  Work split:
//   k_raster_meshlets  one wave per surviving meshlet: its <= 64 vertices are
//                      transformed once (one per lane) into LDS, then each lane
//                      sets up <= 2 of its <= 124 triangles.  Triangles whose
//                      vertex bbox fits in `small_extent` pixels are rasterized
//                      in the lane with int32 edge functions; larger ones go to
//                      the big queue, clip-crossing ones to the clip queue
//                      (both appended with one atomic per wave, tagged with
//                      the pass).
//   k_raster_clip      clips queued triangles and queues the fan pieces.
//   k_raster_big       one 256-thread block per queued triangle, 16x16 tiles.
#include "kernels.h"
#include "lh_device.h"
#include "shading.h"

using namespace lhd;

namespace {

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
    const uint64_t clip_bytes = (uint64_t(v.clip_capacity) * LH_CLIP_ENTRY_BYTES + 255) & ~uint64_t(255);
    return {static_cast<uint2*>(v.scratch),
            reinterpret_cast<BigTri*>(static_cast<char*>(v.scratch) + clip_bytes)};
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
    float xn = c.x / c.w, yn = c.y / c.w, zn = c.z / c.w;
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
// pixel bbox.
template <int KIND>
LH_DEV bool setup(SV& v0, SV& v1, SV& v2, uint32_t W, uint32_t H, int64_t& area, int32_t& xlo, int32_t& xhi,
                  int32_t& ylo, int32_t& yhi, int32_t& extx, int32_t& exty, bool two_sided = false) {
    area = edge64(v0.X, v0.Y, v1.X, v1.Y, v2.X, v2.Y);
    if (two_sided ? area == 0 : (KIND == 0 ? area >= 0 : area <= 0)) return false;
    if (area < 0) {
        SV t = v1;
        v1 = v2;
        v2 = t;
        area = -area;
    }
    int32_t minX = min(v0.X, min(v1.X, v2.X)), maxX = max(v0.X, max(v1.X, v2.X));
    int32_t minY = min(v0.Y, min(v1.Y, v2.Y)), maxY = max(v0.Y, max(v1.Y, v2.Y));
    xlo = max((minX + 127) >> 8, 0);
    xhi = min((maxX - 128) >> 8, int32_t(W) - 1);
    ylo = max((minY + 127) >> 8, 0);
    yhi = min((maxY - 128) >> 8, int32_t(H) - 1);
    extx = maxX - minX;
    exty = maxY - minY;
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
    const float inv_area = 1.0f / float(area);
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
    const float inv_area = 1.0f / float(area);
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

// ALPHA selects the draws of alpha-clip mesh infos (double-sided, alpha test
// per sample); the opaque instantiation skips them.  The alpha instantiation
// returns at once when the culls counted no alpha-clip draw in its passes.
template <int KIND, class VS, bool ST, bool ALPHA>
__global__ void __launch_bounds__(64) LH_RASTER_OCCUPANCY k_raster_meshlets(lh_scene s, VS vs, PassSet ps) {
    __shared__ int32_t sX[64], sY[64];
    __shared__ float sZ[64];
    __shared__ uint32_t sO[64];
    const uint32_t lane = threadIdx.x;
    if constexpr (ALPHA) {
        uint32_t n = 0;
        for (uint32_t k = lane; k < vs.count() * ps.n; k += 64)
            n += vs.get(k / ps.n).counters[LH_CNT_ALPHA_DRAWS + ps.pass0 + k % ps.n];
        if (wave_sum_u32(n) == 0) return;
    }
    const Segments<VS> seg(vs, ps);
    const uint32_t total = seg.total();
    Tally<ST> tally;
    auto body = [&](uint32_t view, uint32_t pi, uint32_t d) {
        const lh_view& v = vs.get(view);
        const Pass p = make_pass(v, ps.pass0 + pi);
        const Queues qs = queues(v);
        uint32_t* clip_counter = &v.counters[LH_CNT_CLIP + ps.pass0];
        uint32_t* big_counter = &v.counters[LH_CNT_BIG + ps.pass0];
        const LhMat4 M = view_matrix<KIND>(v, p.pass);
        const MeshletReference ref = p.draws[d];
        const Instance& in = s.instances[ref.instance_index];
        const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
        if (((mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) != 0) != ALPHA) return;  // wave-uniform
        const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
        const uint32_t mid = s.instance_meshlet_base[ref.instance_index] + ref.meshlet_index;
        const uint32_t nvert = ml.index_count, ntri = ml.triangle_count;
        // both rounds' corner indices, issued before the vertex fetch so the
        // two latencies overlap
        const uint8_t* micro = ptr<uint8_t>(mi.micro_indices) + ml.triangle_offset;
        uint32_t corners[2] = {0u, 0u};
        for (uint32_t r = 0; r < 2; r++) {
            const uint32_t t = lane + 64 * r;
            if (t < ntri) corners[r] = uint32_t(micro[3 * t]) | (uint32_t(micro[3 * t + 1]) << 8) |
                                       (uint32_t(micro[3 * t + 2]) << 16);
        }
        if (lane < nvert) {
            f3 w = mat4_point(in.transform, load_position(mi, load_index(mi, ml.index_offset + lane)));
            f4 c = mat4_mul_vec4(M, f4{w.x, w.y, w.z, 1.0f});
            uint32_t oc = finite4(c) ? outcode(c) : kBadVertex;
            if (oc == 0) {
                SV sv;
                if (project(c, p.W, p.H, sv)) {
                    sX[lane] = sv.X;
                    sY[lane] = sv.Y;
                    sZ[lane] = sv.z;
                } else {
                    oc = kBadVertex;
                }
            }
            sO[lane] = oc;
        }
        __syncthreads();
        for (uint32_t round = 0; round < 2; round++) {
            const uint32_t t = lane + 64 * round;
            if (__ballot(t < ntri) == 0) break;
            // 0 none, 1 clip queue, 2 big queue, 3 in-lane, 4 wave-cooperative
            uint32_t cls = 0;
            SV v0{}, v1{}, v2{};
            int64_t area = 0;
            int32_t xlo = 0, xhi = 0, ylo = 0, yhi = 0, ex = 0, ey = 0;
            if (t < ntri) {
                const uint32_t packed = corners[round];
                uint32_t i0 = packed & 255u, i1 = (packed >> 8) & 255u, i2 = packed >> 16;
                uint32_t o0 = sO[i0], o1 = sO[i1], o2 = sO[i2];
                if (((o0 | o1 | o2) & kBadVertex) == 0 && (o0 & o1 & o2) == 0) {
                    if (o0 | o1 | o2) {
                        cls = 1;
                    } else {
                        v0 = SV{sX[i0], sY[i0], sZ[i0]};
                        v1 = SV{sX[i1], sY[i1], sZ[i1]};
                        v2 = SV{sX[i2], sY[i2], sZ[i2]};
                        if (setup<KIND>(v0, v1, v2, p.W, p.H, area, xlo, xhi, ylo, yhi, ex, ey, ALPHA)) {
                            const int32_t e = max(ex, ey);
                            cls = e <= ps.small_extent ? 3u : e <= ps.medium_extent ? 4u : 2u;
                        }
                    }
                }
            }
            const uint32_t id = (mid << LH_KEY_TRI_BITS) | t;
            uint32_t cslot = wave_append(clip_counter, cls == 1);
            if (cls == 1) {
                if (cslot < v.clip_capacity)
                    qs.clip[cslot] = make_uint2((d << LH_KEY_TRI_BITS) | t, pi | (ALPHA ? kClipAlpha : 0u));
                else
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_CLIP);
            }
            uint32_t bslot = wave_append(big_counter, cls == 2);
            if (cls == 2) {
                if (bslot < v.big_capacity)
                    store_big(qs.big + bslot, v0, v1, v2, id, xlo, xhi, ylo, yhi, pi, ALPHA);
                else
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_BIG);
            }
            if (cls == 3)
                raster_small<KIND, ST, ALPHA>(s, v, p, v0, v1, v2, int32_t(area), xlo, xhi, ylo, yhi, id, tally);
            // Medium triangles: broadcast each lane's setup and raster it with the whole wave.
            for (uint64_t med = ballot(cls == 4); med; med &= med - 1) {
                const uint32_t l = first_lane(med);
                const SV a{readlane(v0.X, l), readlane(v0.Y, l), readlane(v0.z, l)};
                const SV b{readlane(v1.X, l), readlane(v1.Y, l), readlane(v1.z, l)};
                const SV c{readlane(v2.X, l), readlane(v2.Y, l), readlane(v2.z, l)};
                raster_wave<KIND, ST, ALPHA>(s, v, p, a, b, c, readlane(xlo, l), readlane(xhi, l), readlane(ylo, l),
                                             readlane(yhi, l), (mid << LH_KEY_TRI_BITS) | (l + 64 * round), lane,
                                             tally);
            }
        }
        tally.flush(v, p.pass);
        __syncthreads();
    };
    constexpr uint32_t C = Segments<VS>::kChunk;
    if constexpr (C == 1) {
        for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
            uint32_t view, pi, d;
            seg.find(g, view, pi, d);
            body(view, pi, d);
        }
    } else {
        for (uint32_t c0 = blockIdx.x * C; c0 < total; c0 += gridDim.x * C) {
            auto cur = seg.cursor(c0);
            const uint32_t c1 = min(c0 + C, total);
            for (uint32_t g = c0; g < c1; g++) {
                uint32_t view, pi, d;
                seg.at(cur, g, view, pi, d);
                body(view, pi, d);
            }
        }
    }
}

// Clip polygon in homogeneous coordinates (Sutherland-Hodgman).
LH_DEV f4 clip_point(f4 a, f4 b, float da, float db) {
    float t = da / (da - db);
    return {a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t};
}

// View blockIdx.z; grid-stride over that view's clip queue.
template <int KIND, class VS>
__global__ void __launch_bounds__(64) k_raster_clip(lh_scene s, VS vs, PassSet ps) {
    const lh_view& v = vs.get(blockIdx.z);
    const Queues qs = queues(v);
    const uint32_t count = min(v.counters[LH_CNT_CLIP + ps.pass0], v.clip_capacity);
    uint32_t* big_counter = &v.counters[LH_CNT_BIG + ps.pass0];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        const uint2 e = qs.clip[q];
        const uint32_t d = e.x >> LH_KEY_TRI_BITS, t = e.x & ((1u << LH_KEY_TRI_BITS) - 1), pi = e.y & ~kClipAlpha;
        const bool alpha = (e.y & kClipAlpha) != 0;
        const Pass p = make_pass(v, ps.pass0 + pi);
        const LhMat4 M = view_matrix<KIND>(v, p.pass);
        const MeshletReference ref = p.draws[d];
        const Instance& in = s.instances[ref.instance_index];
        const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
        const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
        const uint32_t mid = s.instance_meshlet_base[ref.instance_index] + ref.meshlet_index;
        const uint8_t* micro = ptr<uint8_t>(mi.micro_indices) + ml.triangle_offset;
        f4 poly[12], tmp[12];
        for (int j = 0; j < 3; j++) {
            f3 w = mat4_point(in.transform, load_position(mi, load_index(mi, ml.index_offset + micro[3 * t + j])));
            poly[j] = mat4_mul_vec4(M, f4{w.x, w.y, w.z, 1.0f});
        }
        int n = 3;
        for (int k = 0; k < 6 && n >= 3; k++) {
            int m = 0;
            for (int i = 0; i < n; i++) {
                f4 a = poly[i], b = poly[i + 1 == n ? 0 : i + 1];
                float da = plane_dist(a, k), db = plane_dist(b, k);
                bool ia = da >= 0.0f, ib = db >= 0.0f;
                if (ia) tmp[m++] = a;
                if (ia != ib) tmp[m++] = ia ? clip_point(a, b, da, db) : clip_point(b, a, db, da);
            }
            for (int i = 0; i < m; i++) poly[i] = tmp[i];
            n = m;
        }
        if (n < 3) continue;
        SV sv[12];
        bool ok = true;
        for (int i = 0; i < n; i++) ok = ok && project(poly[i], p.W, p.H, sv[i]);
        if (!ok) continue;
        const uint32_t id = (mid << LH_KEY_TRI_BITS) | t;
        for (int i = 1; i + 1 < n; i++) {
            SV a = sv[0], b = sv[i], c = sv[i + 1];
            int64_t area;
            int32_t xlo, xhi, ylo, yhi, ex, ey;
            if (!setup<KIND>(a, b, c, p.W, p.H, area, xlo, xhi, ylo, yhi, ex, ey, alpha)) continue;
            uint32_t slot = atomicAdd(big_counter, 1u);
            if (slot < v.big_capacity)
                store_big(qs.big + slot, a, b, c, id, xlo, xhi, ylo, yhi, pi, alpha);
            else
                atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_BIG);
        }
    }
}

// View blockIdx.z; one 256-thread block per queued triangle; the bbox is
// walked in 16x16 tiles and each pixel evaluates the int64 edge functions.
// ALPHA selects the entries of alpha-clip geometry (alpha test per sample),
// so the opaque instantiation keeps its small register footprint; the alpha
// one returns at once when the view's culls counted no alpha-clip draw.
template <int KIND, class VS, bool ST, bool ALPHA>
__global__ void __launch_bounds__(256) k_raster_big(lh_scene s, VS vs, PassSet ps) {
    const lh_view& v = vs.get(blockIdx.z);
    if constexpr (ALPHA) {
        uint32_t n = 0;
        for (uint32_t k = 0; k < ps.n; k++) n += v.counters[LH_CNT_ALPHA_DRAWS + ps.pass0 + k];
        if (n == 0) return;
    }
    const Queues qs = queues(v);
    const uint32_t count = min(v.counters[LH_CNT_BIG + ps.pass0], v.big_capacity);
    const int32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    Tally<ST> tally;
    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const BigTri b = qs.big[e];
        if (((b.ylo & kBigAlpha) != 0) != ALPHA) continue;  // block-uniform
        const Pass p = make_pass(v, ps.pass0 + (uint32_t(b.xlo) >> 13));
        const int32_t X0 = b.X[0], X1 = b.X[1], X2 = b.X[2], Y0 = b.Y[0], Y1 = b.Y[1], Y2 = b.Y[2];
        const int64_t area = edge64(X0, Y0, X1, Y1, X2, Y2);
        const float inv_area = 1.0f / float(area);
        const float z0 = b.z[0], d10 = b.z[1] - b.z[0], d20 = b.z[2] - b.z[0];
        const bool tl0 = top_left(X1, Y1, X2, Y2), tl1 = top_left(X2, Y2, X0, Y0), tl2 = top_left(X0, Y0, X1, Y1);
        const int32_t xlo = b.xlo & 0x1FFF, xhi = b.xhi, ylo = b.ylo & ~kBigAlpha, yhi = b.yhi;
        for (int32_t by = ylo; by <= yhi; by += 16) {
            for (int32_t bx = xlo; bx <= xhi; bx += 16) {
                const int32_t px = bx + tx, py = by + ty;
                if (px > xhi || py > yhi) continue;
                const int32_t Px = px * 256 + 128, Py = py * 256 + 128;
                const int64_t e0 = edge64(X1, Y1, X2, Y2, Px, Py);
                const int64_t e1 = edge64(X2, Y2, X0, Y0, Px, Py);
                const int64_t e2 = edge64(X0, Y0, X1, Y1, Px, Py);
                if (!(e0 > 0 || (e0 == 0 && tl0))) continue;
                if (!(e1 > 0 || (e1 == 0 && tl1))) continue;
                if (!(e2 > 0 || (e2 == 0 && tl2))) continue;
                if (ALPHA && !alpha_pass<KIND>(s, v, p, b.id, px, py)) continue;
                const float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
                tally.hit(fragment<KIND>(p, px, py, (z0 + l1 * d10) + l2 * d20, b.id));
            }
        }
        tally.flush(v, p.pass);
    }
}

// Batch plan: off[k + 1] = off[k] + surviving meshlets of segment k
// (segment = view k / np, pass pass0 + k % np), scanned in LDS by one block.
__global__ void __launch_bounds__(256) k_plan_segments(ViewTable vs, uint32_t pass0, uint32_t np) {
    __shared__ uint32_t cnt[LH_MAX_BATCH_SEGMENTS];
    const uint32_t nseg = vs.n * np;
    for (uint32_t k = threadIdx.x; k < nseg; k += blockDim.x) {
        const lh_view& v = vs.views[k / np];
        cnt[k] = min(v.counters[LH_CNT_DRAWS + pass0 + k % np], v.draw_capacity);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < nseg; k++) {
            uint32_t c = cnt[k];
            cnt[k] = acc;
            acc += c;
        }
        vs.seg_offsets[nseg] = acc;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nseg; k += blockDim.x) vs.seg_offsets[k] = cnt[k];
}

}  // namespace

namespace lhk {

template <int KIND, bool ST, class VS>
void launch_kernels(const lh_scene& s, const VS& vs, const PassSet& ps, uint32_t meshlet_blocks, uint32_t clip_blocks,
                    uint32_t big_blocks, hipStream_t st) {
    const uint32_t nv = vs.count();
    mark(KIND == 0 ? "raster_meshlets_view" : "raster_meshlets_shadow", st);
    hipLaunchKernelGGL((k_raster_meshlets<KIND, VS, ST, false>), dim3(meshlet_blocks), dim3(64), 0, st, s, vs, ps);
    const bool alpha = !(s.flags & LH_SCENE_OPAQUE_ONLY);
    if (alpha) {
        mark(KIND == 0 ? "raster_alpha_view" : "raster_alpha_shadow", st);
        hipLaunchKernelGGL((k_raster_meshlets<KIND, VS, ST, true>), dim3(meshlet_blocks), dim3(64), 0, st, s, vs, ps);
    }
    mark(KIND == 0 ? "raster_clip_view" : "raster_clip_shadow", st);
    hipLaunchKernelGGL((k_raster_clip<KIND, VS>), dim3(clip_blocks, 1, nv), dim3(64), 0, st, s, vs, ps);
    mark(KIND == 0 ? "raster_big_view" : "raster_big_shadow", st);
    hipLaunchKernelGGL((k_raster_big<KIND, VS, ST, false>), dim3(big_blocks, 1, nv), dim3(256), 0, st, s, vs, ps);
    if (alpha) {
        mark(KIND == 0 ? "raster_big_alpha_view" : "raster_big_alpha_shadow", st);
        hipLaunchKernelGGL((k_raster_big<KIND, VS, ST, true>), dim3(big_blocks, 1, nv), dim3(256), 0, st, s, vs, ps);
    }
}

// Statistics are a single-view feature (LH_VIEW_STATISTICS).
template <int KIND>
void launch_set(const lh_scene& s, const OneView& vs, const PassSet& ps, uint32_t mb, uint32_t cb, uint32_t bb,
                hipStream_t st) {
    if (vs.v.flags & LH_VIEW_STATISTICS)
        launch_kernels<KIND, true>(s, vs, ps, mb, cb, bb, st);
    else
        launch_kernels<KIND, false>(s, vs, ps, mb, cb, bb, st);
}
template <int KIND>
void launch_set(const lh_scene& s, const ViewTable& vs, const PassSet& ps, uint32_t mb, uint32_t cb, uint32_t bb,
                hipStream_t st) {
    launch_kernels<KIND, false>(s, vs, ps, mb, cb, bb, st);
}

template <class VS>
hipError_t raster(const lh_scene& s, const VS& vs, uint32_t pass0, uint32_t npasses, const RasterTuning& t,
                  hipStream_t st, bool phase2) {
    if (npasses == 0 || npasses > LH_NUM_CASCADES || pass0 + npasses > LH_NUM_PASSES || (pass0 == 0 && npasses != 1))
        return hipErrorInvalidValue;
    const uint32_t small_px = std::min<uint32_t>(std::max<uint32_t>(t.small_extent_px, 1), 32);
    const uint32_t medium_px = std::min<uint32_t>(std::max(t.medium_extent_px, small_px), 64);
    PassSet ps{pass0, npasses, int32_t(small_px * 256), int32_t(medium_px * 256), phase2 ? 1u : 0u};
    const uint32_t nv = vs.count();
    // per-view persistent grids of the queue kernels shrink with the batch size
    const uint32_t clip_blocks = nv > 1 ? std::max(t.clip_blocks / 16, 4u) : t.clip_blocks;
    const uint32_t big_blocks = nv > 1 ? std::max(t.big_blocks / 16, 32u) : t.big_blocks;
    if (pass0 == 0)
        launch_set<0>(s, vs, ps, t.meshlet_blocks, clip_blocks, big_blocks, st);
    else
        launch_set<1>(s, vs, ps, t.meshlet_blocks, clip_blocks, big_blocks, st);
    return hipGetLastError();
}

template hipError_t raster<OneView>(const lh_scene&, const OneView&, uint32_t, uint32_t, const RasterTuning&,
                                    hipStream_t, bool);
template hipError_t raster<ViewTable>(const lh_scene&, const ViewTable&, uint32_t, uint32_t, const RasterTuning&,
                                      hipStream_t, bool);

hipError_t plan_segments(const ViewTable& vt, uint32_t pass0, uint32_t npasses, hipStream_t st) {
    if (vt.n * npasses > LH_MAX_BATCH_SEGMENTS) return hipErrorInvalidValue;
    mark("plan_segments", st);
    hipLaunchKernelGGL(k_plan_segments, dim3(1), dim3(256), 0, st, vt, pass0, npasses);
    return hipGetLastError();
}

}  // namespace lhk
