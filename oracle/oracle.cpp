// Scalar C++ oracle of the GPU-driven frame.  TEST INFRASTRUCTURE ONLY (see
// oracle.h): the checker for the HIP kernels, never part of the product path.
//
// Arithmetic conventions (shared with the kernels by definition, not by code):
//   dot3(a, b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   length(v)       = sqrt(dot3(v, v)); normalize(v) = (v.x/l, v.y/l, v.z/l)
//   (M*v)[r]        = ((M[0][r]*v.x + M[1][r]*v.y) + M[2][r]*v.z) + M[3][r]*v.w
//   (A*B)[c]        = A * B[c]
//   mix(x, y, a)    = x*(1-a) + y*a
// compiled with -ffp-contract=off; no transcendental function feeds a
// bit-exact result (powers of the cascade split are integer powers).
#include "oracle.h"
#include "texel.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ------------------------------------------------------------------ math

struct F2 {
    float x, y;
};
struct F3 {
    float x, y, z;
};
struct F4 {
    float x, y, z, w;
};

inline F3 f3(const LhVec3& v) { return {v.x, v.y, v.z}; }
inline F4 f4(const LhVec4& v) { return {v.x, v.y, v.z, v.w}; }
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 mul(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline F3 neg(F3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float length3(F3 v) { return std::sqrt(dot3(v, v)); }
inline F3 normalize3(F3 v) {
    float l = length3(v);
    return {v.x / l, v.y / l, v.z / l};
}
inline F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }

inline float mrow(const LhVec4& c, int r) { return r == 0 ? c.x : r == 1 ? c.y : r == 2 ? c.z : c.w; }

inline F4 mat4_mul_vec4(const LhMat4& m, F4 v) {
    float o[4];
    for (int r = 0; r < 4; r++)
        o[r] = ((mrow(m.c[0], r) * v.x + mrow(m.c[1], r) * v.y) + mrow(m.c[2], r) * v.z) + mrow(m.c[3], r) * v.w;
    return {o[0], o[1], o[2], o[3]};
}
inline F3 mat4_point(const LhMat4& m, F3 p) {
    F4 r = mat4_mul_vec4(m, F4{p.x, p.y, p.z, 1.0f});
    return {r.x, r.y, r.z};
}
inline F3 mat3_mul_vec3(const LhMat3& m, F3 v) {
    const LhVec3* c = m.c;
    return {(c[0].x * v.x + c[1].x * v.y) + c[2].x * v.z, (c[0].y * v.x + c[1].y * v.y) + c[2].y * v.z,
            (c[0].z * v.x + c[1].z * v.y) + c[2].z * v.z};
}
inline LhMat4 mat4_mul(const LhMat4& a, const LhMat4& b) {
    LhMat4 r;
    for (int c = 0; c < 4; c++) {
        F4 v = mat4_mul_vec4(a, f4(b.c[c]));
        r.c[c] = LhVec4{v.x, v.y, v.z, v.w};
    }
    return r;
}
inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline float near_plane(const Uniforms* u) { return u->perspective.c[3].z; }

// Integer power for the cascade split (pow(0.5, p) at
// src/shaders/compute/generate_shadow_matrices.glsl:121-130); falls back to
// powf for a non-integral exponent.
float split_pow(float base, float p) {
    if (p >= 0.0f && p <= 64.0f && std::floor(p) == p) {
        float r = 1.0f;
        for (int i = 0; i < int(p); i++) r = r * base;
        return r;
    }
    return std::pow(base, p);
}

template <class T>
const T* at(uint64_t addr) {
    return reinterpret_cast<const T*>(static_cast<uintptr_t>(addr));
}

// ------------------------------------------------------- vertex fetching

// load_index (src/shaders/common/util.glsl:3-9).
uint32_t load_index(const MeshInfo& mi, uint32_t i) {
    if (mi.flags & MESH_INFO_FLAGS_32_BIT_INDICES) return at<uint32_t>(mi.indices)[i];
    return at<uint16_t>(mi.indices)[i];
}
// QuantizedPositionBuffer (src/shaders/common/buffer_references.glsl:33-36):
// u16x3 converted to float; FP32 positions with MESH_INFO_FLAGS_FP32_VERTICES.
F3 load_position(const MeshInfo& mi, uint32_t v) {
    if (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) {
        const float* p = at<float>(mi.positions) + 3 * size_t(v);
        return {p[0], p[1], p[2]};
    }
    const uint16_t* p = at<uint16_t>(mi.positions) + 3 * size_t(v);
    return {float(p[0]), float(p[1]), float(p[2])};
}
// QuanitizedNormals (buffer_references.glsl:43-45): i8x3 -> float3.
F3 load_normal(const MeshInfo& mi, uint32_t v) {
    if (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) {
        const float* p = at<float>(mi.normals) + 3 * size_t(v);
        return {p[0], p[1], p[2]};
    }
    const int8_t* p = at<int8_t>(mi.normals) + 3 * size_t(v);
    return {float(p[0]), float(p[1]), float(p[2])};
}
// QuanitizedUvs (buffer_references.glsl:39-41): u16x2 -> float2.
F2 load_uv(const MeshInfo& mi, uint32_t v) {
    if (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) {
        const float* p = at<float>(mi.uvs) + 2 * size_t(v);
        return {p[0], p[1]};
    }
    const uint16_t* p = at<uint16_t>(mi.uvs) + 2 * size_t(v);
    return {float(p[0]), float(p[1])};
}

const MeshInfo& mesh_info_of(const Instance& in) { return *at<MeshInfo>(in.mesh_info_address); }
const Meshlet& meshlet_of(const MeshInfo& mi, uint32_t k) { return at<Meshlet>(mi.meshlets)[k]; }
// Corner j of triangle k: micro index -> remap -> vertex index
// (load_vertex_data, src/shaders/rasterization.glsl:12-34).
uint32_t corner_vertex(const MeshInfo& mi, const Meshlet& ml, uint32_t k, uint32_t j) {
    uint8_t m = at<uint8_t>(mi.micro_indices)[size_t(ml.triangle_offset) + 3 * k + j];
    return load_index(mi, ml.index_offset + m);
}

// ------------------------------------------------------------- culling

// World-space sphere of a (mesh-space) sphere under the instance transform
// (src/shaders/common/culling.glsl:4-17).
void world_sphere(const Instance& in, const float* s, F3& centre, float& radius) {
    centre = mat4_point(in.transform, F3{s[0], s[1], s[2]});
    const LhMat4& t = in.transform;
    float sx = length3(F3{t.c[0].x, t.c[0].y, t.c[0].z});
    float sy = length3(F3{t.c[1].x, t.c[1].y, t.c[1].z});
    float sz = length3(F3{t.c[2].x, t.c[2].y, t.c[2].z});
    radius = s[3] * std::max(std::max(sx, sy), sz);
}

// cull_bounding_sphere (src/shaders/common/culling.glsl:1-46).
bool cull_sphere(const Instance& in, const float* s, const Uniforms& u) {
    F3 c;
    float r;
    world_sphere(in, s, c, r);
    F3 v = mat4_point(u.initial_view, c);
    v.z = -v.z;
    bool visible = v.z + r > near_plane(&u);
    const LhMat4& p = u.perspective;
    // rows 3, 0, 1 of the perspective matrix (transpose(P)[k])
    F3 row3{p.c[0].w, p.c[1].w, p.c[2].w}, row0{p.c[0].x, p.c[1].x, p.c[2].x}, row1{p.c[0].y, p.c[1].y, p.c[2].y};
    F3 fx = normalize3(add(row3, row0));
    F3 fy = normalize3(add(row3, row1));
    visible = visible && (v.z * fx.z + std::fabs(v.x) * fx.x < r);
    visible = visible && (v.z * fy.z - std::fabs(v.y) * fy.y < r);
    return !visible;
}

// fits_partially_inside / fits_entirely_inside (culling.glsl:48-62).
bool fits_partially(F2 p, float r, float size) { return (std::fabs(p.x) - r < size) && (std::fabs(p.y) - r < size); }
bool fits_entirely(F2 p, float r, float size) { return (std::fabs(p.x) + r < size) && (std::fabs(p.y) + r < size); }
F2 shadow_view_xy(const MiscStorage& m, uint32_t c, F3 w) {
    F3 v = mat4_point(m.shadow_view_matrices[c], w);
    return {v.x, v.y};
}

// cull_bounding_sphere_shadows (culling.glsl:82-135).
bool cull_sphere_shadow(const Instance& in, const float* s, const MiscStorage& m, uint32_t c) {
    F3 w;
    float r;
    world_sphere(in, s, w, r);
    if (!fits_partially(shadow_view_xy(m, c, w), r, m.shadow_sphere_radii[c])) return true;
    for (uint32_t k = 0; k < c; k++)
        if (fits_entirely(shadow_view_xy(m, k, w), r, m.shadow_sphere_radii[k])) return true;
    return false;
}

// cull_cone_perspective / cull_cone_orthographic (culling.glsl:137-151).
bool cull_cone(const Instance& in, const Meshlet& ml, const Uniforms& u, bool ortho) {
    F3 axis = normalize3(mat3_mul_vec3(in.normal_transform, f3(ml.cone_axis)));
    if (ortho) return dot3(f3(u.sun_dir), axis) >= ml.cone_cutoff;
    F3 apex = mat4_point(in.transform, f3(ml.cone_apex));
    return dot3(normalize3(sub(apex, f3(u.camera_pos))), normalize3(axis)) >= ml.cone_cutoff;
}

// ---------------------------------------------------------- prefix sum

uint64_t* sector(const lh_view& v, uint32_t pass) { return v.prefix_sums + size_t(pass) * (1 + v.instance_capacity); }

// prefix_sum_inclusive_append (src/shaders/common/prefix_sum.glsl:17-30).
void prefix_append(const lh_view& v, uint32_t pass, uint32_t index, uint32_t value) {
    uint64_t* s = sector(v, pass);
    uint64_t old = s[0];
    s[0] = old + (uint64_t(value) | (uint64_t(1) << 32));
    uint32_t slot = uint32_t(old >> 32);
    if (slot >= v.instance_capacity) {
        v.counters[LH_CNT_OVERFLOW] |= LH_OVERFLOW_INSTANCES;
        return;
    }
    PrefixSumValue* vals = reinterpret_cast<PrefixSumValue*>(s + 1);
    vals[slot].index = index;
    vals[slot].sum = uint32_t(old) + value;
}

// prefix_sum_binary_search (prefix_sum.glsl:33-46), std::upper_bound on sum.
bool prefix_search(const lh_view& v, uint32_t pass, uint32_t target, PrefixSumValue& out) {
    const uint64_t* s = sector(v, pass);
    uint32_t count = std::min(uint32_t(s[0] >> 32), v.instance_capacity);
    const PrefixSumValue* vals = reinterpret_cast<const PrefixSumValue*>(s + 1);
    uint32_t first = 0;
    while (count > 0) {
        uint32_t step = count / 2, cur = first + step;
        if (target >= vals[cur].sum) {
            first = cur + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    uint32_t n = std::min(uint32_t(s[0] >> 32), v.instance_capacity);
    if (first >= n) return false;
    out = vals[first];
    return true;
}

// --------------------------------------------------------------- raster

constexpr float kGuard = 4.0f;  // guard band |x|, |y| <= 4 w

struct CV {
    float x, y, z, w;
};
struct SV {
    int32_t X, Y;
    float z;
};

float plane_dist(const CV& v, int k) {
    switch (k) {
        case 0: return v.w - v.z;        // z <= w
        case 1: return v.z;              // z >= 0
        case 2: return kGuard * v.w - v.x;
        case 3: return kGuard * v.w + v.x;
        case 4: return kGuard * v.w - v.y;
        default: return kGuard * v.w + v.y;
    }
}
uint32_t outcode(const CV& v) {
    uint32_t c = 0;
    for (int k = 0; k < 6; k++)
        if (!(plane_dist(v, k) >= 0.0f)) c |= 1u << k;
    return c;
}
bool finite4(const CV& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z) && std::isfinite(v.w); }

// Intersection of the segment (in -> out) with plane k, always computed from
// the inside end so a shared edge clips to the same point in both triangles.
CV clip_point(const CV& in, const CV& out, float din, float dout) {
    float t = din / (din - dout);
    return {in.x + (out.x - in.x) * t, in.y + (out.y - in.y) * t, in.z + (out.z - in.z) * t,
            in.w + (out.w - in.w) * t};
}

// Sutherland-Hodgman against the six planes; returns the vertex count (<= 9).
int clip_polygon(CV* poly, int n) {
    CV tmp[12];
    for (int k = 0; k < 6 && n >= 3; k++) {
        int m = 0;
        for (int i = 0; i < n; i++) {
            const CV& a = poly[i];
            const CV& b = poly[(i + 1) % n];
            float da = plane_dist(a, k), db = plane_dist(b, k);
            bool ia = da >= 0.0f, ib = db >= 0.0f;
            if (ia) tmp[m++] = a;
            if (ia != ib) tmp[m++] = ia ? clip_point(a, b, da, db) : clip_point(b, a, db, da);
        }
        for (int i = 0; i < m; i++) poly[i] = tmp[i];
        n = m;
    }
    return n >= 3 ? n : 0;
}

// Viewport transform (x_f = (W/2) x_ndc + W/2) and 8-bit subpixel snapping.
bool project(const CV& c, uint32_t W, uint32_t H, SV& o) {
    if (!(c.w > 0.0f)) return false;
    float xn = c.x / c.w, yn = c.y / c.w, zn = c.z / c.w;
    float hw = 0.5f * float(W), hh = 0.5f * float(H);
    float xf = xn * hw + hw, yf = yn * hh + hh;
    o.X = int32_t(std::rint(xf * 256.0f));
    o.Y = int32_t(std::rint(yf * 256.0f));
    o.z = zn;
    return true;
}

bool alpha_visible(const lh_scene& sc, const LhMat4& m, uint32_t W, uint32_t H, uint32_t id, int32_t px,
                   int32_t py);

struct RasterTarget {
    uint32_t W, H, kind;  // kind 0 main (u64 keys), 1 shadow (u32 depth)
    uint64_t* vis;
    uint32_t* depth;
    bool shared = false;  // several threads rasterize into the target: min by compare-exchange
};

// Alpha-clip geometry (MESH_INFO_FLAGS_ALPHA_CLIP): double-sided, alpha test
// per covered sample (src/pipelines.cpp:22-25, rasterization.glsl:84-97,
// :130-145).  NULL for opaque geometry.
struct AlphaClip {
    const lh_scene* sc;
    const LhMat4* m;  // the pass's view-projection matrix
};

template <class T>
inline void store_min(T& slot, T value, bool shared) {
    if (!shared) {
        if (value < slot) slot = value;
        return;
    }
    std::atomic_ref<T> a(slot);
    T cur = a.load(std::memory_order_relaxed);
    while (value < cur && !a.compare_exchange_weak(cur, value, std::memory_order_relaxed)) {
    }
}

inline int64_t edge(int32_t ax, int32_t ay, int32_t bx, int32_t by, int32_t px, int32_t py) {
    return int64_t(bx - ax) * int64_t(py - ay) - int64_t(by - ay) * int64_t(px - ax);
}
inline bool top_left(int32_t ax, int32_t ay, int32_t bx, int32_t by) {
    int32_t dx = bx - ax, dy = by - ay;
    return dy < 0 || (dy == 0 && dx > 0);
}

// Rasterizes one snapped triangle.  Main view keeps front faces (signed area
// < 0 in y-down framebuffer coordinates = counter-clockwise on screen,
// src/pipelines.cpp:10-13 with the default VK_FRONT_FACE_COUNTER_CLOCKWISE);
// shadows keep back faces (cull FRONT, src/pipelines.cpp:15-19).
uint64_t raster_snapped(SV v0, SV v1, SV v2, const RasterTarget& t, uint32_t id, const AlphaClip* ac) {
    int64_t area = edge(v0.X, v0.Y, v1.X, v1.Y, v2.X, v2.Y);
    if (ac ? area == 0 : (t.kind == 0 ? area >= 0 : area <= 0)) return 0;
    if (area < 0) {
        std::swap(v1, v2);
        area = -area;
    }
    int32_t minX = std::min({v0.X, v1.X, v2.X}), maxX = std::max({v0.X, v1.X, v2.X});
    int32_t minY = std::min({v0.Y, v1.Y, v2.Y}), maxY = std::max({v0.Y, v1.Y, v2.Y});
    int32_t xlo = std::max((minX + 127) >> 8, 0), xhi = std::min((maxX - 128) >> 8, int32_t(t.W) - 1);
    int32_t ylo = std::max((minY + 127) >> 8, 0), yhi = std::min((maxY - 128) >> 8, int32_t(t.H) - 1);
    if (xlo > xhi || ylo > yhi) return 0;
    bool tl0 = top_left(v1.X, v1.Y, v2.X, v2.Y), tl1 = top_left(v2.X, v2.Y, v0.X, v0.Y),
         tl2 = top_left(v0.X, v0.Y, v1.X, v1.Y);
    float inv_area = 1.0f / float(area);
    float d10 = v1.z - v0.z, d20 = v2.z - v0.z;
    uint64_t frags = 0;
    for (int32_t py = ylo; py <= yhi; py++) {
        int32_t Py = py * 256 + 128;
        for (int32_t px = xlo; px <= xhi; px++) {
            int32_t Px = px * 256 + 128;
            int64_t e0 = edge(v1.X, v1.Y, v2.X, v2.Y, Px, Py);
            int64_t e1 = edge(v2.X, v2.Y, v0.X, v0.Y, Px, Py);
            int64_t e2 = edge(v0.X, v0.Y, v1.X, v1.Y, Px, Py);
            if (!(e0 > 0 || (e0 == 0 && tl0))) continue;
            if (!(e1 > 0 || (e1 == 0 && tl1))) continue;
            if (!(e2 > 0 || (e2 == 0 && tl2))) continue;
            if (ac && !alpha_visible(*ac->sc, *ac->m, t.W, t.H, id, px, py)) continue;
            float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
            float z = (v0.z + l1 * d10) + l2 * d20;
            z = std::min(std::max(z, 0.0f), 1.0f);
            frags++;
            size_t p = size_t(py) * t.W + px;
            if (t.kind == 0) {
                if (!(z > 0.0f)) continue;  // depth GREATER against the 0.0 clear
                uint64_t key = (uint64_t(~fbits(z)) << 32) | id;
                store_min(t.vis[p], key, t.shared);
            } else {
                uint32_t d = fbits(z);
                store_min(t.depth[p], d, t.shared);
            }
        }
    }
    return frags;
}

uint64_t raster_clip_triangle(const CV* c, const RasterTarget& t, uint32_t id, const AlphaClip* ac = nullptr) {
    if (!finite4(c[0]) || !finite4(c[1]) || !finite4(c[2])) return 0;
    uint32_t o0 = outcode(c[0]), o1 = outcode(c[1]), o2 = outcode(c[2]);
    if (o0 & o1 & o2) return 0;
    if ((o0 | o1 | o2) == 0) {
        SV s[3];
        for (int i = 0; i < 3; i++)
            if (!project(c[i], t.W, t.H, s[i])) return 0;
        return raster_snapped(s[0], s[1], s[2], t, id, ac);
    }
    CV poly[12] = {c[0], c[1], c[2]};
    int n = clip_polygon(poly, 3);
    if (n == 0) return 0;
    SV s[12];
    for (int i = 0; i < n; i++)
        if (!project(poly[i], t.W, t.H, s[i])) return 0;
    uint64_t frags = 0;
    for (int i = 1; i + 1 < n; i++) frags += raster_snapped(s[0], s[i], s[i + 1], t, id, ac);
    return frags;
}

// Draws [d0, d1) of pass `pass`; returns the covered samples.
uint64_t raster_draws(const lh_scene& sc, const lh_view& v, uint32_t pass, uint32_t d0, uint32_t d1, bool shared) {
    const Uniforms& u = *v.uniforms;
    const MeshletReference* draws = v.draws + size_t(pass) * v.draw_capacity;
    RasterTarget t;
    const LhMat4* m;
    if (pass == 0) {
        t = RasterTarget{v.width, v.height, 0, v.visbuffer, nullptr};
        m = &u.combined_perspective_view;
    } else {
        uint32_t c = pass - 1;
        t = RasterTarget{v.shadow_size, v.shadow_size, 1, nullptr,
                         v.shadow_maps + size_t(c) * v.shadow_size * v.shadow_size};
        m = &v.misc->shadow_matrices[c];
    }
    t.shared = shared;
    uint64_t frags = 0;
    for (uint32_t d = d0; d < d1; d++) {
        const MeshletReference& ref = draws[d];
        const Instance& in = sc.instances[ref.instance_index];
        const MeshInfo& mi = mesh_info_of(in);
        const Meshlet& ml = meshlet_of(mi, ref.meshlet_index);
        uint32_t mid = sc.instance_meshlet_base[ref.instance_index] + ref.meshlet_index;
        const AlphaClip alpha{&sc, m};
        const AlphaClip* ac = (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) ? &alpha : nullptr;
        for (uint32_t k = 0; k < ml.triangle_count; k++) {
            CV c[3];
            for (uint32_t j = 0; j < 3; j++) {
                // calculate_world_pos (util.glsl:11-16) then the view matrix.
                F3 w = mat4_point(in.transform, load_position(mi, corner_vertex(mi, ml, k, j)));
                F4 q = mat4_mul_vec4(*m, F4{w.x, w.y, w.z, 1.0f});
                c[j] = CV{q.x, q.y, q.z, q.w};
            }
            frags += raster_clip_triangle(c, t, (mid << LH_KEY_TRI_BITS) | k, ac);
        }
    }
    return frags;
}

// Splits [0, n) into `threads` contiguous parts run on std::threads.
template <class F>
void parallel_for(uint32_t n, uint32_t threads, F&& f) {
    threads = std::max(1u, std::min(threads, std::max(n, 1u)));
    if (threads == 1) {
        f(0u, n, 0u);
        return;
    }
    std::vector<std::thread> pool;
    for (uint32_t i = 0; i < threads; i++) {
        const uint32_t a = uint32_t(uint64_t(n) * i / threads), b = uint32_t(uint64_t(n) * (i + 1) / threads);
        pool.emplace_back([&f, a, b, i] { f(a, b, i); });
    }
    for (auto& th : pool) th.join();
}

// Rasterizes every triangle of every meshlet in draws[pass]
// (drawIndirectCount of src/rendering.cpp:235-242 / :405-412), with
// `threads` threads over contiguous draw ranges (the result is a minimum per
// pixel, so it does not depend on the split).
void raster_pass(const lh_scene& sc, const lh_view& v, uint32_t pass, uint32_t threads = 1) {
    const uint32_t count = std::min(v.counters[LH_CNT_DRAWS + pass], v.draw_capacity);
    std::vector<uint64_t> frags(std::max(threads, 1u), 0);
    parallel_for(count, threads, [&](uint32_t a, uint32_t b, uint32_t i) {
        frags[i] = raster_draws(sc, v, pass, a, b, threads > 1);
    });
    uint64_t total = 0;
    for (uint64_t f : frags) total += f;
    // LH_VIEW_STATISTICS: covered samples of the pass (depth-test order
    // dependent atomics counts have no sequential counterpart).
    if (v.flags & LH_VIEW_STATISTICS) v.counters[LH_CNT_FRAGMENTS + pass] += uint32_t(total);
}

// ------------------------------------------------------------- resolve

struct BaryDeriv {
    F3 lambda, ddx, ddy;
};

// CalcFullBary (src/shaders/common/vbuffer.glsl:11-63).
BaryDeriv calc_full_bary(F4 p0, F4 p1, F4 p2, F2 ndc, F2 win) {
    BaryDeriv r;
    F3 invW{1.0f / p0.w, 1.0f / p1.w, 1.0f / p2.w};
    F2 n0{p0.x * invW.x, p0.y * invW.x}, n1{p1.x * invW.y, p1.y * invW.y}, n2{p2.x * invW.z, p2.y * invW.z};
    F2 c0{n2.x - n1.x, n2.y - n1.y}, c1{n0.x - n1.x, n0.y - n1.y};
    float invDet = 1.0f / (c0.x * c1.y - c1.x * c0.y);
    r.ddx = mul(mul(F3{n1.y - n2.y, n2.y - n0.y, n0.y - n1.y}, invDet), invW);
    r.ddy = mul(mul(F3{n2.x - n1.x, n0.x - n2.x, n1.x - n0.x}, invDet), invW);
    float ddxSum = (r.ddx.x + r.ddx.y) + r.ddx.z;
    float ddySum = (r.ddy.x + r.ddy.y) + r.ddy.z;
    F2 dv{ndc.x - n0.x, ndc.y - n0.y};
    float interpInvW = (invW.x + dv.x * ddxSum) + dv.y * ddySum;
    float interpW = 1.0f / interpInvW;
    r.lambda.x = interpW * ((invW.x + dv.x * r.ddx.x) + dv.y * r.ddy.x);
    r.lambda.y = interpW * ((0.0f + dv.x * r.ddx.y) + dv.y * r.ddy.y);
    r.lambda.z = interpW * ((0.0f + dv.x * r.ddx.z) + dv.y * r.ddy.z);
    float sx = 2.0f / win.x, sy = 2.0f / win.y;
    r.ddx = mul(r.ddx, sx);
    r.ddy = mul(r.ddy, sy);
    ddxSum *= sx;
    ddySum *= sy;
    r.ddy = mul(r.ddy, -1.0f);
    ddySum *= -1.0f;
    float wdx = 1.0f / (interpInvW + ddxSum), wdy = 1.0f / (interpInvW + ddySum);
    r.ddx = F3{wdx * (r.lambda.x * interpInvW + r.ddx.x) - r.lambda.x, wdx * (r.lambda.y * interpInvW + r.ddx.y) - r.lambda.y,
               wdx * (r.lambda.z * interpInvW + r.ddx.z) - r.lambda.z};
    r.ddy = F3{wdy * (r.lambda.x * interpInvW + r.ddy.x) - r.lambda.x, wdy * (r.lambda.y * interpInvW + r.ddy.y) - r.lambda.y,
               wdy * (r.lambda.z * interpInvW + r.ddy.z) - r.lambda.z};
    return r;
}

// InterpolateWithDeriv (vbuffer.glsl:65-73): (value, d/dx, d/dy).
F3 interp1(const BaryDeriv& d, float a, float b, float c) {
    F3 m{a, b, c};
    return {dot3(m, d.lambda), dot3(m, d.ddx), dot3(m, d.ddy)};
}

struct Interp3 {
    F3 v, dx, dy;
};
struct Interp2 {
    F2 v, dx, dy;
};
Interp3 interp3(const BaryDeriv& d, F3 a, F3 b, F3 c) {
    F3 x = interp1(d, a.x, b.x, c.x), y = interp1(d, a.y, b.y, c.y), z = interp1(d, a.z, b.z, c.z);
    return {{x.x, y.x, z.x}, {x.y, y.y, z.y}, {x.z, y.z, z.z}};
}
Interp2 interp2(const BaryDeriv& d, F2 a, F2 b, F2 c) {
    F3 x = interp1(d, a.x, b.x, c.x), y = interp1(d, a.y, b.y, c.y);
    return {{x.x, y.x}, {x.y, y.y}, {x.z, y.z}};
}

// Texel fetch with repeat addressing: RGBA8, BC1 or BC7 storage, sRGB decoded
// to linear when the format says so.
F4 texel(const LhTexture& t, uint32_t level, int32_t i, int32_t j) {
    int32_t w = int32_t(std::max(1u, t.width >> level)), h = int32_t(std::max(1u, t.height >> level));
    i %= w;
    if (i < 0) i += w;
    j %= h;
    if (j < 0) j += h;
    const orc_texel::RGBA8 c =
        orc_texel::fetch(t, at<uint8_t>(t.texels), level, uint32_t(w), uint32_t(i), uint32_t(j));
    if (t.format & LH_TEXTURE_SRGB)
        return {orc_texel::srgb_to_linear<float>(c.c[0]), orc_texel::srgb_to_linear<float>(c.c[1]),
                orc_texel::srgb_to_linear<float>(c.c[2]), c.c[3] / 255.0f};
    return {c.c[0] / 255.0f, c.c[1] / 255.0f, c.c[2] / 255.0f, c.c[3] / 255.0f};
}
F4 mix4(F4 a, F4 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t)}; }
F4 bilinear(const LhTexture& t, uint32_t level, F2 uv) {
    float w = float(std::max(1u, t.width >> level)), h = float(std::max(1u, t.height >> level));
    float u = uv.x * w - 0.5f, v = uv.y * h - 0.5f;
    float fu = std::floor(u), fv = std::floor(v);
    float a = u - fu, b = v - fv;
    int32_t i = int32_t(fu), j = int32_t(fv);
    F4 t00 = texel(t, level, i, j), t10 = texel(t, level, i + 1, j), t01 = texel(t, level, i, j + 1),
       t11 = texel(t, level, i + 1, j + 1);
    return mix4(mix4(t00, t10, a), mix4(t01, t11, a), b);
}

// sample_texture (src/shaders/render_geometry.comp:59-66): textureGrad with
// the repeat sampler, restated as isotropic trilinear filtering with the LOD
// from the analytic uv derivatives (the reference sampler asks for 16x
// anisotropy, src/main.cpp:550-560, which is implementation defined).
// An unused texture index reads as opaque white.
struct Lod {
    uint32_t l0, l1;
    float f;
};
Lod texture_lod(const LhTexture& t, const Interp2& uv) {
    float W = float(t.width), H = float(t.height);
    float ax = uv.dx.x * W, ay = uv.dx.y * H, bx = uv.dy.x * W, by = uv.dy.y * H;
    float rho = std::max(std::sqrt(ax * ax + ay * ay), std::sqrt(bx * bx + by * by));
    float lambda = rho > 0.0f ? std::log2(rho) : 0.0f;
    lambda = clampf(lambda, 0.0f, float(t.levels - 1));
    float d0 = std::floor(lambda);
    uint32_t l0 = uint32_t(d0);
    return {l0, std::min(l0 + 1, t.levels - 1), lambda - d0};
}
bool has_texture(const lh_scene& sc, uint32_t index) { return index != UNUSED_TEXTURE_INDEX && index < sc.num_textures; }
F4 sample_texture(const lh_scene& sc, uint32_t index, const Interp2& uv) {
    if (!has_texture(sc, index)) return {1, 1, 1, 1};
    const LhTexture& t = sc.textures[index];
    Lod lod = texture_lod(t, uv);
    F4 c0 = bilinear(t, lod.l0, uv.v);
    if (lod.f == 0.0f) return c0;
    return mix4(c0, bilinear(t, lod.l1, uv.v), lod.f);
}

// Alpha test of the alpha-clip pipelines (visbuffer_alpha_clip_pixel,
// shadowmap_alpha_clipped_pixel, src/shaders/rasterization.glsl:84-97,
// :130-145): base-colour alpha at the pixel centre, discarded below 0.5; the
// uv is interpolated perspective-correctly with the LOD from its analytic
// derivatives (CalcFullBary) in place of the hardware's 2x2-quad
// differences.  A triangle with a vertex at or behind w = 0, or a non-finite
// uv, passes.
bool alpha_visible(const lh_scene& sc, const LhMat4& m, uint32_t W, uint32_t H, uint32_t id, int32_t px,
                   int32_t py) {
    const MeshletReference ref = sc.meshlet_table[id >> LH_KEY_TRI_BITS];
    const Instance& in = sc.instances[ref.instance_index];
    const MeshInfo& mi = mesh_info_of(in);
    if (!has_texture(sc, mi.base_color_texture_index)) return true;
    const Meshlet& ml = meshlet_of(mi, ref.meshlet_index);
    const uint32_t tri = id & ((1u << LH_KEY_TRI_BITS) - 1);
    F4 clip[3];
    F2 uvs[3];
    for (uint32_t j = 0; j < 3; j++) {
        const uint32_t vi = corner_vertex(mi, ml, tri, j);
        F3 w = mat4_point(in.transform, load_position(mi, vi));
        clip[j] = mat4_mul_vec4(m, F4{w.x, w.y, w.z, 1.0f});
        if (!(clip[j].w > 0.0f)) return true;
        F2 t = load_uv(mi, vi);
        uvs[j] = {t.x * mi.texture_scale.x, t.y * mi.texture_scale.y};
    }
    F2 win{float(W), float(H)};
    F2 ndc{(float(px) + 0.5f) / win.x * 2.0f - 1.0f, (float(py) + 0.5f) / win.y * 2.0f - 1.0f};
    BaryDeriv bary = calc_full_bary(clip[0], clip[1], clip[2], ndc, win);
    Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);
    uv.v.x += mi.texture_offset.x;
    uv.v.y += mi.texture_offset.y;
    if (!(std::isfinite(uv.v.x) && std::isfinite(uv.v.y) && std::isfinite(uv.dx.x) && std::isfinite(uv.dx.y) &&
          std::isfinite(uv.dy.x) && std::isfinite(uv.dy.y)))
        return true;
    return !(sample_texture(sc, mi.base_color_texture_index, uv).w < 0.5f);
}

// One tap of sampler2DArrayShadow with linear filtering, clamp to edge and
// compare LESS (result 1 where ref < texel), render_geometry.comp:211-218.
float shadow_tap(const lh_view& v, uint32_t layer, F2 uv, float ref) {
    const uint32_t S = v.shadow_size;
    const uint32_t* map = v.shadow_maps + size_t(layer) * S * S;
    float u = uv.x * float(S) - 0.5f, w = uv.y * float(S) - 0.5f;
    float fu = std::floor(u), fw = std::floor(w);
    float a = u - fu, b = w - fw;
    int32_t i = int32_t(fu), j = int32_t(fw);
    auto cmp = [&](int32_t x, int32_t y) {
        x = std::min(std::max(x, 0), int32_t(S) - 1);
        y = std::min(std::max(y, 0), int32_t(S) - 1);
        return ref < bitsf(map[size_t(y) * S + x]) ? 1.0f : 0.0f;
    };
    return mixf(mixf(cmp(i, j), cmp(i + 1, j), a), mixf(cmp(i, j + 1), cmp(i + 1, j + 1), a), b);
}

constexpr float kPi = 3.14159265358979323846f;
inline float pow5(float x) {
    float x2 = x * x;
    return (x2 * x2) * x;
}
// Filament terms (src/shaders/common/pbr.glsl:1-42).
float F_Schlick1(float u, float f0, float f90) { return f0 + (f90 - f0) * pow5(1.0f - u); }
float Fd_Burley(float NoV, float NoL, float LoH, float roughness) {
    float f90 = 0.5f + ((2.0f * roughness) * LoH) * LoH;
    float ls = F_Schlick1(NoL, 1.0f, f90), vs = F_Schlick1(NoV, 1.0f, f90);
    return (ls * vs) * (1.0f / kPi);
}
float D_GGX(float NoH, float a) {
    float a2 = a * a;
    float f = (NoH * a2 - NoH) * NoH + 1.0f;
    return a2 / ((kPi * f) * f);
}
F3 F_Schlick3(float u, F3 f0) {
    float p = pow5(1.0f - u);
    return {f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
}
float V_SmithGGXCorrelated(float NoV, float NoL, float a) {
    float a2 = a * a;
    float GGXL = NoV * std::sqrt((-NoL * a2 + NoL) * NoL + a2);
    float GGXV = NoL * std::sqrt((-NoV * a2 + NoV) * NoV + a2);
    return 0.5f / (GGXV + GGXL);
}
F3 diffuse_color(F3 base, float metallic) { return mul(base, 1.0f - metallic); }

// BRDF (src/shaders/common/pbr.glsl:44-96).
F3 brdf(F3 v, F3 l, F3 n, float pr, float metallic, F3 base, bool thin) {
    F3 h = normalize3(add(v, l));
    F3 diffuse = diffuse_color(base, metallic);
    const float refl = ((1.5f - 1.0f) / (1.5f + 1.0f)) * ((1.5f - 1.0f) / (1.5f + 1.0f));
    F3 f0 = add(F3{refl * (1.0f - metallic), refl * (1.0f - metallic), refl * (1.0f - metallic)}, mul(base, metallic));
    float NoV = clampf(dot3(n, v), 1e-5f, 1.0f);
    float NoL = clampf(dot3(n, l), 0.0f, 1.0f);
    float NoH = clampf(dot3(n, h), 0.0f, 1.0f);
    float LoH = clampf(dot3(l, h), 0.0f, 1.0f);
    float a = pr * pr;
    float D = D_GGX(NoH, a);
    F3 F = F_Schlick3(LoH, f0);
    float V = V_SmithGGXCorrelated(NoV, NoL, a);
    F3 Fr = mul(F, D * V);
    F3 Fd = mul(diffuse, Fd_Burley(NoV, NoL, LoH, a));
    F3 irr = mul(add(Fd, Fr), NoL);
    irr = F3{std::max(irr.x, 0.0f), std::max(irr.y, 0.0f), std::max(irr.z, 0.0f)};
    if (thin) {
        F3 nn = neg(n);
        float rNoV = clampf(dot3(nn, v), 1e-5f, 1.0f), rNoL = clampf(dot3(nn, l), 0.0f, 1.0f);
        irr = add(irr, mul(diffuse, Fd_Burley(rNoV, rNoL, LoH, a) * rNoL));
    }
    return {std::isnan(irr.x) ? 0.0f : irr.x, std::isnan(irr.y) ? 0.0f : irr.y, std::isnan(irr.z) ? 0.0f : irr.z};
}

// Sky without a skybox cubemap (lh_view.skybox NULL): a constant colour
// times the 2^3 exposure of render_geometry.comp:93-97.
const F3 kSky{0.35f * 8.0f, 0.50f * 8.0f, 0.80f * 8.0f};

// Skybox lookup (render_geometry.comp:86-101): the ray through the pixel
// corner unprojected by perspective_inverse at z = 0, normalized, rotated to
// world space by view_inverse; cube face by the major axis (ties: x, then y)
// with the (sc, tc, rc) table of the Vulkan specification's cube map face
// selection; s = (sc / |rc| + 1) / 2; bilinear at LOD 0 with the face's edge
// clamped; times 2^3.
F3 sky_texel(const lh_view& v, uint32_t f, int32_t i, int32_t j) {
    const int32_t n = int32_t(v.skybox_size);
    i = std::min(std::max(i, 0), n - 1);
    j = std::min(std::max(j, 0), n - 1);
    const float* t = v.skybox + 4 * ((size_t(f) * uint32_t(n) + uint32_t(j)) * uint32_t(n) + uint32_t(i));
    return {t[0], t[1], t[2]};
}
F3 mix3(F3 a, F3 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }
F3 sky_radiance(const lh_view& v, F2 ndc) {
    if (!v.skybox || v.skybox_size == 0) return kSky;
    const Uniforms& u = *v.uniforms;
    const F4 un = mat4_mul_vec4(u.perspective_inverse, F4{ndc.x, ndc.y, 0.0f, 1.0f});
    const F3 local = normalize3(F3{un.x, un.y, un.z});
    const F4 r = mat4_mul_vec4(u.view_inverse, F4{local.x, local.y, local.z, 0.0f});
    const float ax = std::fabs(r.x), ay = std::fabs(r.y), az = std::fabs(r.z);
    uint32_t face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        face = r.x >= 0.0f ? 0u : 1u;
        sc = r.x >= 0.0f ? -r.z : r.z;
        tc = -r.y;
        ma = ax;
    } else if (ay >= az) {
        face = r.y >= 0.0f ? 2u : 3u;
        sc = r.x;
        tc = r.y >= 0.0f ? r.z : -r.z;
        ma = ay;
    } else {
        face = r.z >= 0.0f ? 4u : 5u;
        sc = r.z >= 0.0f ? r.x : -r.x;
        tc = -r.y;
        ma = az;
    }
    const float n = float(v.skybox_size);
    const float fs = (sc / ma + 1.0f) * 0.5f, ft = (tc / ma + 1.0f) * 0.5f;
    const float uu = fs * n - 0.5f, vv = ft * n - 0.5f;
    const float fu = std::floor(uu), fv = std::floor(vv);
    const float a = uu - fu, b = vv - fv;
    const int32_t i = int32_t(fu), j = int32_t(fv);
    const F3 c = mix3(mix3(sky_texel(v, face, i, j), sky_texel(v, face, i + 1, j), a),
                      mix3(sky_texel(v, face, i, j + 1), sky_texel(v, face, i + 1, j + 1), a), b);
    return mul(c, 8.0f);
}

// Cascade selection (render_geometry.comp:183-196; the loop may end at 4,
// which the array layer clamps to 3) and 3x3 PCF (:206-226).
float shadow_factor(const lh_view& v, const MeshInfo& mi, F3 world, uint32_t* cascade_out = nullptr) {
    const MiscStorage& misc = *v.misc;
    uint32_t cascade;
    F4 sc4{0, 0, 0, 0};
    for (cascade = 0; cascade < 4; cascade++) {
        sc4 = mat4_mul_vec4(misc.uv_space_shadow_matrices[cascade], F4{world.x, world.y, world.z, 1.0f});
        float ox = sc4.x - 0.5f, oy = sc4.y - 0.5f;
        if (std::fabs(ox) < 0.5f && std::fabs(oy) < 0.5f) break;
    }
    if (cascade_out) *cascade_out = cascade;
    float bias = (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) ? 0.000002f : 0.0f;
    sc4 = F4{sc4.x / sc4.w, sc4.y / sc4.w, sc4.z / sc4.w, sc4.w / sc4.w};
    uint32_t layer = std::min(cascade, 3u);
    float shadow_sum = 0.0f;
    const float S = float(v.shadow_size);
    for (int ox = -1; ox <= 1; ox++)
        for (int oy = -1; oy <= 1; oy++) {
            F2 o{float(ox) / S, float(oy) / S};
            shadow_sum += shadow_tap(v, layer, F2{sc4.x + o.x, sc4.y + o.y}, sc4.z - bias);
        }
    shadow_sum /= 9.0f;
    if (sc4.z > 1.0f) shadow_sum = 1.0f;
    return shadow_sum;
}

// Normal mapping (render_geometry.comp:238-246): the sampled texel decoded
// as map = texel * 255 / 127 - 128 / 127, rotated by the cotangent frame of
// compute_cotangent_frame (:25-38) - the screen-space derivatives of the
// world position and uv give T = dp/dy_perp * du/dx + dp/dx_perp * du/dy and
// B likewise for v, both scaled by 1 / max(|T|, |B|) - and normalized.
F3 normal_map(F3 nrm, const Interp3& world, const Interp2& uv, F4 mn) {
    F3 map{mn.x * 255.0f / 127.0f - 128.0f / 127.0f, mn.y * 255.0f / 127.0f - 128.0f / 127.0f,
           mn.z * 255.0f / 127.0f - 128.0f / 127.0f};
    F3 dpy_perp = cross3(world.dy, nrm), dpx_perp = cross3(nrm, world.dx);
    F3 t = add(mul(dpy_perp, uv.dx.x), mul(dpx_perp, uv.dy.x));
    F3 b = add(mul(dpy_perp, uv.dx.y), mul(dpx_perp, uv.dy.y));
    float invmax = 1.0f / std::sqrt(std::max(dot3(t, t), dot3(b, b)));
    t = mul(t, invmax);
    b = mul(b, invmax);
    return normalize3(F3{(t.x * map.x + b.x * map.y) + nrm.x * map.z, (t.y * map.x + b.y * map.y) + nrm.y * map.z,
                         (t.z * map.x + b.z * map.y) + nrm.z * map.z});
}

// DEBUG_COLOURS (src/shaders/common/debug.glsl:1-11).
F3 debug_colour(uint32_t i) {
    static const float k[10][3] = {{0, 2, 91},    {0, 108, 251}, {0, 221, 221}, {51, 221, 0},  {255, 252, 0},
                                   {255, 180, 0}, {255, 104, 0}, {226, 22, 0},  {191, 0, 83}, {145, 0, 65}};
    const float* c = k[std::min(i, 9u)];
    return {c[0] / 255.0f, c[1] / 255.0f, c[2] / 255.0f};
}

// GLSL smoothstep and temperature() (src/shaders/common/debug.glsl:15-43),
// the heat map of the shader-clock debug view.
float smoothstepf(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
F3 temperature(float t) {
    const float s = t * 10.0f;
    const int32_t cur = int32_t(s) <= 9 ? int32_t(s) : 9;
    const int32_t prv = cur >= 1 ? cur - 1 : 0, nxt = cur < 9 ? cur + 1 : 9;
    const float blur = 0.8f;
    const float wc = smoothstepf(float(cur) - blur, float(cur) + blur, s) *
                     (1.0f - smoothstepf(float(cur + 1) - blur, float(cur + 1) + blur, s));
    const float wp = 1.0f - smoothstepf(float(cur) - blur, float(cur) + blur, s);
    const float wn = smoothstepf(float(cur + 1) - blur, float(cur + 1) + blur, s);
    const F3 a = debug_colour(uint32_t(cur)), b = debug_colour(uint32_t(prv)), c = debug_colour(uint32_t(nxt));
    return {clampf((wc * a.x + wp * b.x) + wn * c.x, 0.0f, 1.0f), clampf((wc * a.y + wp * b.y) + wn * c.y, 0.0f, 1.0f),
            clampf((wc * a.z + wp * b.z) + wn * c.z, 0.0f, 1.0f)};
}

// render_geometry main() (src/shaders/render_geometry.comp:70-266) for one pixel.
F3 resolve_pixel(const lh_scene& sc, const lh_view& v, uint32_t x, uint32_t y) {
    const Uniforms& u = *v.uniforms;
    F2 win{float(u.window_size.x), float(u.window_size.y)};
    F2 ndc{float(x) / win.x * 2.0f - 1.0f, float(y) / win.y * 2.0f - 1.0f};  // pixel corner (:82-84)
    uint64_t key = v.visbuffer[size_t(y) * v.width + x];
    if (~uint32_t(key >> 32) == 0u) return sky_radiance(v, ndc);  // depth 0.0
    uint32_t mid = uint32_t(key >> LH_KEY_TRI_BITS) & (LH_MAX_SCENE_MESHLETS - 1);
    uint32_t tri = uint32_t(key) & ((1u << LH_KEY_TRI_BITS) - 1);
    const MeshletReference& ref = sc.meshlet_table[mid];
    const Instance& in = sc.instances[ref.instance_index];
    const MeshInfo& mi = mesh_info_of(in);
    const Meshlet& ml = meshlet_of(mi, ref.meshlet_index);
    uint32_t vi[3];
    F3 pos[3];
    F4 clip[3];
    for (int j = 0; j < 3; j++) {
        vi[j] = corner_vertex(mi, ml, tri, j);
        pos[j] = mat4_point(in.transform, load_position(mi, vi[j]));
        clip[j] = mat4_mul_vec4(u.combined_perspective_view, F4{pos[j].x, pos[j].y, pos[j].z, 1.0f});
    }
    BaryDeriv bary = calc_full_bary(clip[0], clip[1], clip[2], ndc, win);
    Interp3 world = interp3(bary, pos[0], pos[1], pos[2]);
    F2 ts{mi.texture_scale.x, mi.texture_scale.y};
    F2 uvs[3];
    for (int j = 0; j < 3; j++) {
        F2 t = load_uv(mi, vi[j]);
        uvs[j] = {t.x * ts.x, t.y * ts.y};
    }
    Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);
    uv.v.x += mi.texture_offset.x;
    uv.v.y += mi.texture_offset.y;
    F3 view_vector = normalize3(sub(f3(u.camera_pos), world.v));
    F3 nrm = interp3(bary, normalize3(load_normal(mi, vi[0])), normalize3(load_normal(mi, vi[1])),
                     normalize3(load_normal(mi, vi[2])))
                 .v;
    nrm = normalize3(mat3_mul_vec3(in.normal_transform, nrm));
    F3 gn = cross3(sub(pos[1], pos[0]), sub(pos[2], pos[0]));
    if (dot3(gn, view_vector) < 0.0f) nrm = neg(nrm);
    uint32_t cascade = 0;
    float shadow_sum = shadow_factor(v, mi, world.v, &cascade);
    // Material (:228-246).
    const LhMaterialFactors& fac = sc.factors[(in.mesh_info_address - uint64_t(uintptr_t(sc.mesh_infos))) / sizeof(MeshInfo)];
    F4 bc = sample_texture(sc, mi.base_color_texture_index, uv);
    F3 base = mul(F3{bc.x, bc.y, bc.z}, f3(mi.base_color_factor));
    F4 mr = sample_texture(sc, mi.metallic_roughness_texture_index, uv);
    float roughness = mr.y * fac.roughness_factor;
    float metallic = mr.z * fac.metallic_factor;
    if (mi.normal_texture_index != UNUSED_TEXTURE_INDEX)
        nrm = normal_map(nrm, world, uv, sample_texture(sc, mi.normal_texture_index, uv));
    // Debug views (:268-308); the shader-clock heat map has no clock here and
    // shades normally (the kernel's heat map is checked against temperature()).
    if (u.debug == UNIFORMS_DEBUG_CASCADES) return mul(base, debug_colour(cascade));
    if (u.debug == UNIFORMS_DEBUG_TRIANGLE_INDEX) return debug_colour(tri % 10u);
    if (u.debug == UNIFORMS_DEBUG_INSTANCE_INDEX) return debug_colour(uint32_t(ref.meshlet_index) % 10u);
    if (u.debug == UNIFORMS_DEBUG_NORMALS) return {nrm.x * 0.5f + 0.5f, nrm.y * 0.5f + 0.5f, nrm.z * 0.5f + 0.5f};
    // Ambient SH (:40-57, :248-249).
    const float L = 0.025f;
    float shv = (((((((L + 0.0f * nrm.x) + L * nrm.y) + 0.0f * nrm.z) + 0.0f * (nrm.x * nrm.z)) + 0.0f * (nrm.z * nrm.y)) +
                  0.0f * (nrm.y * nrm.x)) +
                 0.0f * ((3.0f * nrm.z) * nrm.z - 1.0f)) +
                0.0f * (nrm.x * nrm.x - nrm.y * nrm.y);
    F3 ambient = mul(diffuse_color(base, metallic), shv);
    float sun_scale = kPi * shadow_sum;
    F3 sun = mul(mul(brdf(view_vector, f3(u.sun_dir), nrm, roughness, metallic, base,
                          (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) != 0),
                     sun_scale),
                 f3(u.sun_intensity));
    return add(sun, ambient);
}

// display_transform (src/shaders/display_transform.comp:5-33): the
// tony-mc-mapface LUT is absent, so the LUT's own pre-encoding x/(x+1) is
// used as the tone curve, followed by the sRGB OETF and BGRA8 packing.
float srgb_oetf(float x) { return x <= 0.003130f ? x * 12.92f : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f; }
uint32_t tonemap_pixel(F3 c) {
    auto enc = [](float x) {
        x = std::max(x, 0.0f);
        float t = srgb_oetf(x / (x + 1.0f));
        t = clampf(t, 0.0f, 1.0f);
        return uint32_t(t * 255.0f + 0.5f);
    };
    return enc(c.z) | (enc(c.y) << 8) | (enc(c.x) << 16) | (255u << 24);
}

// textureLod of the shadow-map array at level 0: linear filter, clamp to edge.
float shadow_depth(const lh_view& v, uint32_t layer, F2 uv) {
    const int32_t S = int32_t(v.shadow_size);
    const uint32_t* map = v.shadow_maps + size_t(layer) * v.shadow_size * v.shadow_size;
    float fx = uv.x * float(S) - 0.5f, fy = uv.y * float(S) - 0.5f;
    float x0 = std::floor(fx), y0 = std::floor(fy);
    float a = fx - x0, b = fy - y0;
    int32_t i = int32_t(x0), j = int32_t(y0);
    auto at = [&](int32_t ii, int32_t jj) {
        ii = std::min(std::max(ii, 0), S - 1);
        jj = std::min(std::max(jj, 0), S - 1);
        return bitsf(map[size_t(jj) * v.shadow_size + uint32_t(ii)]);
    };
    return mixf(mixf(at(i, j), at(i + 1, j), a), mixf(at(i, j + 1), at(i + 1, j + 1), a), b);
}

// tony_mc_mapface (display_transform.comp:5-17) with the view's LUT:
// x / (x + 1) aligned to texel centres, trilinear, clamp to edge.
F3 lut_lookup(const float* lut, F3 c) {
    const int32_t N = int32_t(LH_TONEMAP_LUT_DIMS);
    const float dims = float(LH_TONEMAP_LUT_DIMS);
    const float x[3] = {std::max(c.x, 0.0f), std::max(c.y, 0.0f), std::max(c.z, 0.0f)};
    int32_t i0[3];
    float fr[3];
    for (int k = 0; k < 3; k++) {
        float uvw = (x[k] / (x[k] + 1.0f)) * ((dims - 1.0f) / dims) + 0.5f / dims;
        float t = uvw * dims - 0.5f, fl = std::floor(t);
        fr[k] = t - fl;
        i0[k] = int32_t(fl);
    }
    auto at = [&](int32_t r, int32_t g, int32_t b) {
        r = std::min(std::max(r, 0), N - 1);
        g = std::min(std::max(g, 0), N - 1);
        b = std::min(std::max(b, 0), N - 1);
        const float* q = lut + 4 * ((size_t(b) * N + g) * N + r);
        return F3{q[0], q[1], q[2]};
    };
    auto mix3 = [](F3 a, F3 b, float t) { return F3{mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; };
    F3 plane[2];
    for (int db = 0; db < 2; db++) {
        F3 g0 = mix3(at(i0[0], i0[1], i0[2] + db), at(i0[0] + 1, i0[1], i0[2] + db), fr[0]);
        F3 g1 = mix3(at(i0[0], i0[1] + 1, i0[2] + db), at(i0[0] + 1, i0[1] + 1, i0[2] + db), fr[0]);
        plane[db] = mix3(g0, g1, fr[1]);
    }
    return mix3(plane[0], plane[1], fr[2]);
}
uint32_t encode_display(float x) { return uint32_t(clampf(srgb_oetf(x), 0.0f, 1.0f) * 255.0f + 0.5f); }

// display_transform (:37-81) for pixel (x, y): with Uniforms.debug_shadowmaps
// the top-left quarter shows the four cascades (2 x 2), depth d as
// (d - 0.5) * 100 + 0.5 without the tone curve (:51-75).
uint32_t display_pixel(const lh_view& v, uint32_t x, uint32_t y, F3 c) {
    const Uniforms& u = *v.uniforms;
    if (u.debug_shadowmaps) {
        float ux = (float(x) + 0.5f) / float(u.window_size.x), uy = (float(y) + 0.5f) / float(u.window_size.y);
        if (ux < 0.5f && uy < 0.5f) {
            float sx = ux * 2.0f / 0.5f, sy = uy * 2.0f / 0.5f;
            uint32_t layer = (sy > 1.0f ? 2u : 0u) + (sx > 1.0f ? 1u : 0u);
            float g = (shadow_depth(v, layer, F2{sx - std::floor(sx), sy - std::floor(sy)}) - 0.5f) * 100.0f + 0.5f;
            uint32_t e = uint32_t(clampf(srgb_oetf(std::max(g, 0.0f)), 0.0f, 1.0f) * 255.0f + 0.5f);
            return e | (e << 8) | (e << 16) | (255u << 24);
        }
    }
    if (v.tonemap_lut) {
        F3 d = lut_lookup(v.tonemap_lut, c);
        return encode_display(d.z) | (encode_display(d.y) << 8) | (encode_display(d.x) << 16) | (255u << 24);
    }
    return tonemap_pixel(c);
}

bool valid(const lh_scene* sc, const lh_view* v) {
    return sc && v && v->uniforms && v->misc && v->prefix_sums && v->draws && v->counters && v->width > 0 &&
           v->height > 0 && v->width <= LH_MAX_VIEW_SIDE && v->height <= LH_MAX_VIEW_SIDE;
}

// Quantities the differentiable resolve holds fixed, computed in FP32 exactly
// as resolve_pixel computes them.
void frozen_pixel(const lh_scene& sc, const lh_view& v, uint32_t x, uint32_t y, orc_frozen& o) {
    o = orc_frozen{};
    const Uniforms& u = *v.uniforms;
    F2 win{float(u.window_size.x), float(u.window_size.y)};
    F2 ndc{float(x) / win.x * 2.0f - 1.0f, float(y) / win.y * 2.0f - 1.0f};
    uint64_t key = v.visbuffer[size_t(y) * v.width + x];
    if (~uint32_t(key >> 32) == 0u) return;  // sky
    uint32_t mid = uint32_t(key >> LH_KEY_TRI_BITS) & (LH_MAX_SCENE_MESHLETS - 1);
    uint32_t tri = uint32_t(key) & ((1u << LH_KEY_TRI_BITS) - 1);
    const MeshletReference& ref = sc.meshlet_table[mid];
    const Instance& in = sc.instances[ref.instance_index];
    const MeshInfo& mi = mesh_info_of(in);
    const Meshlet& ml = meshlet_of(mi, ref.meshlet_index);
    if (!(mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) || (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) ||
        mi.normal_texture_index != UNUSED_TEXTURE_INDEX) {
        o.kind = 2;
        return;
    }
    F3 pos[3];
    F4 clip[3];
    uint32_t vi[3];
    for (int j = 0; j < 3; j++) {
        vi[j] = corner_vertex(mi, ml, tri, j);
        pos[j] = mat4_point(in.transform, load_position(mi, vi[j]));
        clip[j] = mat4_mul_vec4(u.combined_perspective_view, F4{pos[j].x, pos[j].y, pos[j].z, 1.0f});
    }
    BaryDeriv bary = calc_full_bary(clip[0], clip[1], clip[2], ndc, win);
    Interp3 world = interp3(bary, pos[0], pos[1], pos[2]);
    F2 ts{mi.texture_scale.x, mi.texture_scale.y};
    F2 uvs[3];
    for (int j = 0; j < 3; j++) {
        F2 t = load_uv(mi, vi[j]);
        uvs[j] = {t.x * ts.x, t.y * ts.y};
    }
    Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);
    F3 view_vector = normalize3(sub(f3(u.camera_pos), world.v));
    F3 gn = cross3(sub(pos[1], pos[0]), sub(pos[2], pos[0]));
    o.kind = 1;
    o.flip = dot3(gn, view_vector) < 0.0f ? -1.0f : 1.0f;
    o.shadow = shadow_factor(v, mi, world.v);
    const uint32_t tex[2] = {mi.base_color_texture_index, mi.metallic_roughness_texture_index};
    for (int k = 0; k < 2; k++) {
        if (!has_texture(sc, tex[k])) continue;
        Lod lod = texture_lod(sc.textures[tex[k]], uv);
        o.lod_l0[k] = lod.l0;
        o.lod_l1[k] = lod.l1;
        o.lod_f[k] = lod.f;
    }
}

// Conservative screen bound of a world-space sphere for the two-phase
// occlusion test (restated from the build's device rule): the 8 corners of the
// enclosing cube through m, pixel rectangle of their projections widened by
// one pixel and clamped to the image, and their largest NDC depth.  Returns
// false when a corner has w <= 1e-6 (no bound).
bool sphere_screen_bound(const LhMat4& m, F3 c, float r, uint32_t W, uint32_t H, int32_t* rect, float& zmax) {
    float xmin = 3.0e38f, xmax = -3.0e38f, ymin = 3.0e38f, ymax = -3.0e38f;
    zmax = -3.0e38f;
    const float hw = 0.5f * float(W), hh = 0.5f * float(H);
    for (int k = 0; k < 8; k++) {
        F3 p{c.x + ((k & 1) ? r : -r), c.y + ((k & 2) ? r : -r), c.z + ((k & 4) ? r : -r)};
        F4 q = mat4_mul_vec4(m, F4{p.x, p.y, p.z, 1.0f});
        if (!(q.w > 1e-6f) || !std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) return false;
        float xf = (q.x / q.w) * hw + hw, yf = (q.y / q.w) * hh + hh, z = q.z / q.w;
        xmin = std::min(xmin, xf);
        xmax = std::max(xmax, xf);
        ymin = std::min(ymin, yf);
        ymax = std::max(ymax, yf);
        zmax = std::max(zmax, z);
    }
    xmin = std::max(xmin, -1.0e6f);
    ymin = std::max(ymin, -1.0e6f);
    xmax = std::min(xmax, 1.0e6f);
    ymax = std::min(ymax, 1.0e6f);
    rect[0] = std::max(int32_t(std::floor(xmin)) - 1, 0);
    rect[1] = std::max(int32_t(std::floor(ymin)) - 1, 0);
    rect[2] = std::min(int32_t(std::ceil(xmax)) + 1, int32_t(W) - 1);
    rect[3] = std::min(int32_t(std::ceil(ymax)) + 1, int32_t(H) - 1);
    return true;
}

}  // namespace

extern "C" {

// One texel of a BC1 block (8 bytes), packed RGBA8 with red in the low byte.
uint32_t orc_bc1_texel(const uint8_t* block, uint32_t x, uint32_t y) {
    LhTexture t{};
    t.format = LH_TEXTURE_BC1;
    const orc_texel::RGBA8 c = orc_texel::fetch(t, block, 0, 4, x & 3u, y & 3u);
    return uint32_t(c.c[0]) | (uint32_t(c.c[1]) << 8) | (uint32_t(c.c[2]) << 16) | (uint32_t(c.c[3]) << 24);
}


int64_t orc_hiz_bound_violations(const lh_scene* sc, const lh_view* v, uint64_t* checked) {
    if (!sc || !v || !checked) return -1;
    *checked = 0;
    const Uniforms& u = *v->uniforms;
    const uint32_t count = std::min(v->counters[LH_CNT_DRAWS], v->draw_capacity);
    int64_t bad = 0;
    for (uint32_t d = 0; d < count; d++) {
        const MeshletReference& ref = v->draws[d];
        const Instance& in = sc->instances[ref.instance_index];
        const MeshInfo& mi = mesh_info_of(in);
        const Meshlet& ml = meshlet_of(mi, ref.meshlet_index);
        F3 c;
        float r;
        world_sphere(in, &ml.bounding_sphere.x, c, r);
        int32_t rect[4];
        float zmax;
        if (!sphere_screen_bound(u.combined_perspective_view, c, r, v->width, v->height, rect, zmax)) continue;
        for (uint32_t k = 0; k < ml.triangle_count; k++) {
            CV q[3];
            bool ok = true;
            for (uint32_t j = 0; j < 3; j++) {
                F3 w = mat4_point(in.transform, load_position(mi, corner_vertex(mi, ml, k, j)));
                F4 t = mat4_mul_vec4(u.combined_perspective_view, F4{w.x, w.y, w.z, 1.0f});
                q[j] = CV{t.x, t.y, t.z, t.w};
                ok = ok && q[j].w > 0.0f;
            }
            if (!ok) {
                bad++;  // a bounded meshlet cannot have a vertex at or behind w = 0
                continue;
            }
            SV s3[3];
            for (int j = 0; j < 3; j++) project(q[j], v->width, v->height, s3[j]);
            int32_t minX = std::min({s3[0].X, s3[1].X, s3[2].X}), maxX = std::max({s3[0].X, s3[1].X, s3[2].X});
            int32_t minY = std::min({s3[0].Y, s3[1].Y, s3[2].Y}), maxY = std::max({s3[0].Y, s3[1].Y, s3[2].Y});
            int32_t xlo = std::max((minX + 127) >> 8, 0), xhi = std::min((maxX - 128) >> 8, int32_t(v->width) - 1);
            int32_t ylo = std::max((minY + 127) >> 8, 0), yhi = std::min((maxY - 128) >> 8, int32_t(v->height) - 1);
            (*checked)++;
            if (xlo <= xhi && ylo <= yhi && (xlo < rect[0] || xhi > rect[2] || ylo < rect[1] || yhi > rect[3])) bad++;
            for (int j = 0; j < 3; j++)
                if (std::min(std::max(s3[j].z, 0.0f), 1.0f) > zmax * (1.0f + 1e-4f)) bad++;
        }
    }
    return bad;
}

lh_status orc_resolve_frozen(const lh_scene* sc, const lh_view* v, orc_frozen* out) {
    if (!sc || !v || !v->visbuffer || !v->uniforms || !v->misc || !out) return LH_ERR_INVALID;
    for (uint32_t y = 0; y < v->height; y++)
        for (uint32_t x = 0; x < v->width; x++) frozen_pixel(*sc, *v, x, y, out[size_t(y) * v->width + x]);
    return LH_OK;
}

lh_status orc_view_reset(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v)) return LH_ERR_INVALID;
    std::memset(v->counters, 0, LH_NUM_COUNTERS * 4);
    for (uint32_t p = 0; p < LH_NUM_PASSES; p++) sector(*v, p)[0] = 0;
    v->misc->min_depth = 0xFFFFFFFFu;  // reset_buffers_a (reset_buffers.glsl:9-13)
    v->misc->max_depth = 0;
    if (v->visbuffer) std::fill(v->visbuffer, v->visbuffer + size_t(v->width) * v->height, LH_VISBUFFER_CLEAR);
    if (v->shadow_maps)
        std::fill(v->shadow_maps, v->shadow_maps + size_t(LH_NUM_CASCADES) * v->shadow_size * v->shadow_size,
                  LH_SHADOW_CLEAR);
    return LH_OK;
}

// cull_instances (src/shaders/cull_instances.comp:7-27).
lh_status orc_cull_instances(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v)) return LH_ERR_INVALID;
    const Uniforms& u = *v->uniforms;
    uint32_t n = std::min(u.num_instances, sc->num_instances);
    for (uint32_t i = 0; i < n; i++) {
        const Instance& in = sc->instances[i];
        const MeshInfo& mi = mesh_info_of(in);
        if (cull_sphere(in, &mi.bounding_sphere.x, u)) continue;
        prefix_append(*v, 0, i, mi.num_meshlets);
    }
    return LH_OK;
}

// write_draw_calls / write_draw_calls_shadows (src/shaders/write_draw_calls.comp:7-117)
// with get_meshlet_reference (src/shaders/common/util.glsl:42-59).
lh_status orc_cull_meshlets(const lh_scene* sc, const lh_view* v, uint32_t pass) {
    if (!valid(sc, v) || pass >= LH_NUM_PASSES) return LH_ERR_INVALID;
    const Uniforms& u = *v->uniforms;
    uint32_t total = uint32_t(sector(*v, pass)[0]);
    MeshletReference* out = v->draws + size_t(pass) * v->draw_capacity;
    for (uint32_t id = 0; id < total; id++) {
        PrefixSumValue e;
        if (!prefix_search(*v, pass, id, e)) continue;
        const Instance& in = sc->instances[e.index];
        const MeshInfo& mi = mesh_info_of(in);
        uint32_t local = id - (e.sum - mi.num_meshlets);
        const Meshlet& ml = meshlet_of(mi, local);
        bool alpha = (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) != 0;
        if (pass == 0) {
            if (cull_sphere(in, &ml.bounding_sphere.x, u)) continue;
            if (!alpha && cull_cone(in, ml, u, false)) continue;
        } else {
            if (cull_sphere_shadow(in, &ml.bounding_sphere.x, *v->misc, pass - 1)) continue;
            if (!alpha && cull_cone(in, ml, u, true)) continue;
        }
        uint32_t slot = v->counters[LH_CNT_DRAWS + pass]++;
        if (slot >= v->draw_capacity) {
            v->counters[LH_CNT_OVERFLOW] |= LH_OVERFLOW_DRAWS;
            continue;
        }
        out[slot].instance_index = e.index;
        out[slot].meshlet_index = uint16_t(local);
    }
    return LH_OK;
}

lh_status orc_raster_view(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v) || !v->visbuffer) return LH_ERR_INVALID;
    raster_pass(*sc, *v, 0);
    return LH_OK;
}

lh_status orc_raster_shadow(const lh_scene* sc, const lh_view* v, uint32_t cascade) {
    if (!valid(sc, v) || !v->shadow_maps || cascade >= LH_NUM_CASCADES) return LH_ERR_INVALID;
    raster_pass(*sc, *v, 1 + cascade);
    return LH_OK;
}

// read_depth (src/shaders/compute/read_depth.glsl:3-83): min over non-zero
// depth bits, max over all; depth bits are ~(key >> 32).
lh_status orc_depth_bounds(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v) || !v->visbuffer) return LH_ERR_INVALID;
    uint32_t mn = v->misc->min_depth, mx = v->misc->max_depth;
    for (size_t p = 0; p < size_t(v->width) * v->height; p++) {
        uint32_t d = ~uint32_t(v->visbuffer[p] >> 32);
        if (d != 0) mn = std::min(mn, d);
        mx = std::max(mx, d);
    }
    v->misc->min_depth = mn;
    v->misc->max_depth = mx;
    return LH_OK;
}

// generate_shadow_matrices (src/shaders/compute/generate_shadow_matrices.glsl:15-110)
// with lookAt / OrthographicProjection (src/shaders/common/matrices.glsl:9-40).
lh_status orc_fit_cascades(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v)) return LH_ERR_INVALID;
    const Uniforms& u = *v->uniforms;
    MiscStorage& m = *v->misc;
    if (m.min_depth == 0xFFFFFFFFu) v->counters[LH_CNT_FLAGS] |= LH_FLAG_EMPTY_VIEW;
    float min_depth = bitsf(m.min_depth), max_depth = bitsf(m.max_depth);
    float nearp = near_plane(&u);
    float min_distance = nearp / max_depth, max_distance = nearp / min_depth;
    float splits[4] = {max_depth, nearp / mixf(min_distance, max_distance, split_pow(0.5f, u.cascade_split_pow)),
                       nearp / mixf(min_distance, max_distance, split_pow(0.75f, u.cascade_split_pow)), min_depth};
    for (uint32_t c = 0; c < 4; c++) {
        float last = c > 0 ? splits[c - 1] : max_depth, split = splits[c];
        F3 corners[8] = {{-1, 1, last}, {1, 1, last}, {1, -1, last}, {-1, -1, last},
                         {-1, 1, split}, {1, 1, split}, {1, -1, split}, {-1, -1, split}};
        for (auto& k : corners) {
            F4 q = mat4_mul_vec4(u.inv_perspective_view, F4{k.x, k.y, k.z, 1.0f});
            k = F3{q.x / q.w, q.y / q.w, q.z / q.w};
        }
        F3 centre{0, 0, 0};
        for (auto& k : corners) centre = add(centre, k);
        centre = F3{centre.x / 8.0f, centre.y / 8.0f, centre.z / 8.0f};
        float radius = 0.0f;
        for (auto& k : corners) radius = std::max(radius, length3(sub(k, centre)));
        radius = std::ceil(radius * 16.0f) / 16.0f;
        float dist = u.shadow_cam_distance;
        F3 eye = add(mul(f3(u.sun_dir), dist), centre);
        // lookAt(eye, centre, +Y)
        F3 f = normalize3(sub(centre, eye));
        F3 r = normalize3(cross3(f, F3{0, 1, 0}));
        F3 up = cross3(r, f);
        LhMat4 view;
        view.c[0] = LhVec4{r.x, up.x, -f.x, 0};
        view.c[1] = LhVec4{r.y, up.y, -f.y, 0};
        view.c[2] = LhVec4{r.z, up.z, -f.z, 0};
        view.c[3] = LhVec4{-dot3(r, eye), -dot3(up, eye), dot3(f, eye), 1};
        // OrthographicProjection(-r, r, -r, r, 0, 2 dist)
        float left = -radius, right = radius, bottom = -radius, top = radius, zn = 0.0f, zf = dist * 2.0f;
        float rml = right - left, rpl = right + left, tmb = top - bottom, tpb = top + bottom, fmn = zf - zn;
        LhMat4 proj;
        proj.c[0] = LhVec4{2.0f / rml, 0, 0, 0};
        proj.c[1] = LhVec4{0, -2.0f / tmb, 0, 0};
        proj.c[2] = LhVec4{0, 0, -1.0f / fmn, 0};
        proj.c[3] = LhVec4{-(rpl / rml), -(tpb / tmb), -(zn / fmn), 1};
        LhMat4 mat = mat4_mul(proj, view);
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
    return LH_OK;
}

// cull_instances_shadows (src/shaders/cull_instances.comp:31-56).
lh_status orc_cull_instances_shadow(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v)) return LH_ERR_INVALID;
    uint32_t n = std::min(v->uniforms->num_instances, sc->num_instances);
    for (uint32_t c = 0; c < LH_NUM_CASCADES; c++)
        for (uint32_t i = 0; i < n; i++) {
            const Instance& in = sc->instances[i];
            const MeshInfo& mi = mesh_info_of(in);
            if (cull_sphere_shadow(in, &mi.bounding_sphere.x, *v->misc, c)) continue;
            prefix_append(*v, 1 + c, i, mi.num_meshlets);
        }
    return LH_OK;
}

static void resolve_rows(const lh_scene* sc, const lh_view* v, uint32_t y0, uint32_t y1) {
    for (uint32_t y = y0; y < y1; y++)
        for (uint32_t x = 0; x < v->width; x++) {
            F3 c = resolve_pixel(*sc, *v, x, y);
            float* o = v->radiance + 4 * (size_t(y) * v->width + x);
            o[0] = c.x;
            o[1] = c.y;
            o[2] = c.z;
            o[3] = 1.0f;
            if (v->ldr) v->ldr[size_t(y) * v->width + x] = display_pixel(*v, x, y, c);
        }
}

lh_status orc_resolve(const lh_scene* sc, const lh_view* v) {
    if (!valid(sc, v) || !v->visbuffer || !v->radiance) return LH_ERR_INVALID;
    resolve_rows(sc, v, 0, v->height);
    return LH_OK;
}

lh_status orc_tonemap(const lh_view* v) {
    if (!v || !v->radiance || !v->ldr) return LH_ERR_INVALID;
    for (size_t p = 0; p < size_t(v->width) * v->height; p++) {
        const float* c = v->radiance + 4 * p;
        v->ldr[p] = display_pixel(*v, uint32_t(p % v->width), uint32_t(p / v->width), F3{c[0], c[1], c[2]});
    }
    return LH_OK;
}

// render() pass order (src/rendering.cpp:79-502).
lh_status orc_frame(const lh_scene* sc, const lh_view* v) {
    lh_status s;
    if ((s = orc_view_reset(sc, v))) return s;
    if ((s = orc_cull_instances(sc, v))) return s;
    if ((s = orc_cull_meshlets(sc, v, 0))) return s;
    if ((s = orc_raster_view(sc, v))) return s;
    if ((s = orc_depth_bounds(sc, v))) return s;
    if ((s = orc_fit_cascades(sc, v))) return s;
    if ((s = orc_cull_instances_shadow(sc, v))) return s;
    for (uint32_t c = 0; c < LH_NUM_CASCADES; c++) {
        if ((s = orc_cull_meshlets(sc, v, 1 + c))) return s;
        if ((s = orc_raster_shadow(sc, v, c))) return s;
    }
    if ((s = orc_resolve(sc, v))) return s;
    return v->counters[LH_CNT_OVERFLOW] ? LH_ERR_CAPACITY : LH_OK;
}

// orc_frame with the rasters (contiguous draw ranges, minimum by
// compare-exchange) and the resolve (row bands) on `threads` std::threads;
// identical outputs.  The CPU baseline of the benchmark.
lh_status orc_frame_threads(const lh_scene* sc, const lh_view* v, uint32_t threads) {
    lh_status s;
    if (!valid(sc, v) || !v->visbuffer || !v->shadow_maps || !v->radiance) return LH_ERR_INVALID;
    if ((s = orc_view_reset(sc, v))) return s;
    if ((s = orc_cull_instances(sc, v))) return s;
    if ((s = orc_cull_meshlets(sc, v, 0))) return s;
    raster_pass(*sc, *v, 0, threads);
    if ((s = orc_depth_bounds(sc, v))) return s;
    if ((s = orc_fit_cascades(sc, v))) return s;
    if ((s = orc_cull_instances_shadow(sc, v))) return s;
    for (uint32_t c = 0; c < LH_NUM_CASCADES; c++) {
        if ((s = orc_cull_meshlets(sc, v, 1 + c))) return s;
        raster_pass(*sc, *v, 1 + c, threads);
    }
    parallel_for(v->height, threads, [&](uint32_t a, uint32_t b, uint32_t) { resolve_rows(sc, v, a, b); });
    return v->counters[LH_CNT_OVERFLOW] ? LH_ERR_CAPACITY : LH_OK;
}

uint64_t orc_raster_clip_triangle(const float* clip, uint32_t width, uint32_t height, uint32_t kind, uint32_t id,
                                  uint64_t* vis64, uint32_t* depth32) {
    CV c[3];
    for (int i = 0; i < 3; i++) c[i] = CV{clip[4 * i], clip[4 * i + 1], clip[4 * i + 2], clip[4 * i + 3]};
    RasterTarget t{width, height, kind, vis64, depth32};
    return raster_clip_triangle(c, t, id);
}

int orc_cull_sphere(const Instance* in, const float* s, const Uniforms* u) { return cull_sphere(*in, s, *u) ? 1 : 0; }
int orc_cull_sphere_shadow(const Instance* in, const float* s, const MiscStorage* m, uint32_t c) {
    return cull_sphere_shadow(*in, s, *m, c) ? 1 : 0;
}
int orc_cull_cone(const Instance* in, const Meshlet* ml, const Uniforms* u, int ortho) {
    return cull_cone(*in, *ml, *u, ortho != 0) ? 1 : 0;
}
void orc_calc_full_bary(const float* a, const float* b, const float* c, const float* ndc, const float* win, float* o) {
    BaryDeriv d = calc_full_bary(F4{a[0], a[1], a[2], a[3]}, F4{b[0], b[1], b[2], b[3]}, F4{c[0], c[1], c[2], c[3]},
                                 F2{ndc[0], ndc[1]}, F2{win[0], win[1]});
    float r[9] = {d.lambda.x, d.lambda.y, d.lambda.z, d.ddx.x, d.ddx.y, d.ddx.z, d.ddy.x, d.ddy.y, d.ddy.z};
    std::memcpy(o, r, sizeof(r));
}
void orc_brdf(const float* v, const float* l, const float* n, float r, float m, const float* base, int thin, float* o) {
    F3 x = brdf(F3{v[0], v[1], v[2]}, F3{l[0], l[1], l[2]}, F3{n[0], n[1], n[2]}, r, m, F3{base[0], base[1], base[2]},
                thin != 0);
    o[0] = x.x;
    o[1] = x.y;
    o[2] = x.z;
}
float orc_srgb_oetf(float x) { return srgb_oetf(x); }

void orc_temperature(float t, float* rgb) {
    const F3 c = temperature(t);
    rgb[0] = c.x, rgb[1] = c.y, rgb[2] = c.z;
}

void orc_sky_radiance(const lh_view* v, const float* ndc, float* rgb) {
    const F3 c = sky_radiance(*v, F2{ndc[0], ndc[1]});
    rgb[0] = c.x, rgb[1] = c.y, rgb[2] = c.z;
}

void orc_bc7_block(const uint8_t* block, uint8_t* rgba64) {
    orc_texel::RGBA8 t[16];
    orc_texel::bc7::decode(block, t);
    for (int k = 0; k < 16; k++) std::memcpy(rgba64 + 4 * k, t[k].c, 4);
}
float orc_srgb_to_linear(uint32_t v) { return orc_texel::srgb_to_linear<float>(uint8_t(v)); }

// ---- shading building blocks (known-answer tests) ----
float orc_shadow_tap(const lh_view* v, uint32_t layer, const float* uv, float ref) {
    return shadow_tap(*v, layer, F2{uv[0], uv[1]}, ref);
}
float orc_shadow_factor(const lh_view* v, const MeshInfo* mi, const float* world, uint32_t* cascade) {
    uint32_t c = 0;
    float r = shadow_factor(*v, *mi, F3{world[0], world[1], world[2]}, &c);
    if (cascade) *cascade = c;
    return r;
}
static Interp2 uv_of(const float* uv6) { return {{uv6[0], uv6[1]}, {uv6[2], uv6[3]}, {uv6[4], uv6[5]}}; }
void orc_texture_lod(const LhTexture* t, const float* uv6, uint32_t* l0, uint32_t* l1, float* f) {
    Lod lod = texture_lod(*t, uv_of(uv6));
    *l0 = lod.l0;
    *l1 = lod.l1;
    *f = lod.f;
}
void orc_bilinear(const LhTexture* t, uint32_t level, const float* uv, float* rgba) {
    F4 c = bilinear(*t, level, F2{uv[0], uv[1]});
    rgba[0] = c.x, rgba[1] = c.y, rgba[2] = c.z, rgba[3] = c.w;
}
void orc_sample_texture(const lh_scene* sc, uint32_t index, const float* uv6, float* rgba) {
    F4 c = sample_texture(*sc, index, uv_of(uv6));
    rgba[0] = c.x, rgba[1] = c.y, rgba[2] = c.z, rgba[3] = c.w;
}
void orc_normal_map(const float* n, const float* world9, const float* uv6, const float* texel, float* out) {
    Interp3 w{{world9[0], world9[1], world9[2]}, {world9[3], world9[4], world9[5]}, {world9[6], world9[7], world9[8]}};
    F3 r = normal_map(F3{n[0], n[1], n[2]}, w, uv_of(uv6), F4{texel[0], texel[1], texel[2], texel[3]});
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}

}  // extern "C"
