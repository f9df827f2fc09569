// Host-side scene producer for the GPU-driven frame (C++20, CPU only).
//
// Replaces the reference's asset path for this workload:
//   - meshlet build: src/resources/meshlets.cpp:3-140 (meshoptimizer,
//     64 vertices / 124 triangles, u8 micro indices, per-meshlet triangle
//     block padded to 4 bytes as at :62-64);
//   - meshlet bounds: meshoptimizer's published cluster-bounds method
//     (sphere over triangle corners, normal cone from the bounding sphere of
//     triangle normals, apex behind every triangle plane), called at
//     src/resources/meshlets.cpp:71-78;
//   - mesh bounding sphere: src/resources/bounding_sphere.h:3-74
//     (extremal-points initial sphere, grown until every point fits);
//   - MeshInfo assembly: src/resources/mesh_loading.cpp:660-697;
//   - Instance construction: src/shared_cpu_gpu.h:92-97 (normal matrix fixed
//     to transpose(inverse(M)));
//   - uniforms: src/main.cpp:734-760 with src/projection.h:3-20.
// Geometry is synthetic (seeded); the meshlet grouping is this file's own
// greedy builder, so meshlet boundaries differ from meshoptimizer's.
#include "lh/lh_scene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------- utilities

struct V3 {
    float x = 0, y = 0, z = 0;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    float l = length(a);
    return l > 0 ? a * (1.0f / l) : V3{0, 0, 0};
}

// splitmix64-seeded xoshiro256** generator.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& v : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            v = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    uint32_t below(uint32_t n) { return uint32_t(uniform() * n) % n; }
};

inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Smooth 3D value noise in [-1, 1].
float value_noise3(float x, float y, float z, uint32_t seed) {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = int(fx), iy = int(fy), iz = int(fz);
    float tx = x - fx, ty = y - fy, tz = z - fz;
    tx = tx * tx * (3 - 2 * tx);
    ty = ty * ty * (3 - 2 * ty);
    tz = tz * tz * (3 - 2 * tz);
    auto h = [&](int a, int b, int c) {
        uint32_t k = hash32(uint32_t(a) * 73856093u ^ uint32_t(b) * 19349663u ^ uint32_t(c) * 83492791u ^ seed);
        return float(k & 0xFFFFFF) / float(0xFFFFFF) * 2.0f - 1.0f;
    };
    float v = 0;
    for (int c = 0; c < 8; c++) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        float w = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
        v += w * h(ix + dx, iy + dy, iz + dz);
    }
    return v;
}

LhVec3 lv3(V3 v) { return LhVec3{v.x, v.y, v.z}; }

// ------------------------------------------------------------ mesh sources

struct MeshSource {
    std::vector<float> pos, nrm, uv;  // 3, 3, 2 floats per vertex
    std::vector<uint32_t> idx;        // 3 per triangle
    uint32_t nv() const { return uint32_t(pos.size() / 3); }
    uint32_t nt() const { return uint32_t(idx.size() / 3); }
    V3 p(uint32_t i) const { return {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}; }
    void add_v(V3 p, V3 n, float u, float v) {
        pos.insert(pos.end(), {p.x, p.y, p.z});
        nrm.insert(nrm.end(), {n.x, n.y, n.z});
        uv.insert(uv.end(), {u, v});
    }
    void add_t(uint32_t a, uint32_t b, uint32_t c) { idx.insert(idx.end(), {a, b, c}); }
    // Area-weighted vertex normals from the faces.
    void compute_normals() {
        std::vector<V3> acc(nv());
        for (uint32_t t = 0; t < nt(); t++) {
            uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
            V3 n = cross(p(b) - p(a), p(c) - p(a));
            acc[a] = acc[a] + n;
            acc[b] = acc[b] + n;
            acc[c] = acc[c] + n;
        }
        for (uint32_t i = 0; i < nv(); i++) {
            V3 n = normalize(acc[i]);
            nrm[3 * i] = n.x;
            nrm[3 * i + 1] = n.y;
            nrm[3 * i + 2] = n.z;
        }
    }
};

V3 axis_vec(int a, float s) {
    V3 v;
    (a == 0 ? v.x : a == 1 ? v.y : v.z) = s;
    return v;
}

// Octahedron subdivided `level` times, projected to the unit sphere:
// 8 * 4^level triangles and 4 * 4^level + 2 vertices, CCW seen from outside.
MeshSource octa_sphere(uint32_t level) {
    MeshSource m;
    const int n = 1 << level;
    std::unordered_map<uint64_t, uint32_t> lattice;
    auto vid = [&](int x, int y, int z) {
        uint64_t key = (uint64_t(uint32_t(x + n)) << 42) | (uint64_t(uint32_t(y + n)) << 21) | uint64_t(uint32_t(z + n));
        auto it = lattice.find(key);
        if (it != lattice.end()) return it->second;
        V3 d = normalize(V3{float(x), float(y), float(z)});
        float u = std::atan2(d.z, d.x) * 0.15915494f + 0.5f;
        float v = std::asin(std::clamp(d.y, -1.0f, 1.0f)) * 0.31830989f + 0.5f;
        uint32_t id = m.nv();
        m.add_v(d, d, u, v);
        lattice.emplace(key, id);
        return id;
    };
    for (int oct = 0; oct < 8; oct++) {
        int sx = (oct & 1) ? -1 : 1, sy = (oct & 2) ? -1 : 1, sz = (oct & 4) ? -1 : 1;
        int A[3] = {sx * n, 0, 0}, B[3] = {0, sy * n, 0}, C[3] = {0, 0, sz * n};
        // Orientation: (B-A) x (C-A) must point away from the origin.
        V3 a{float(A[0]), 0, 0}, b{0, float(B[1]), 0}, c{0, 0, float(C[2])};
        bool flip = dot(cross(b - a, c - a), a + b + c) < 0;
        auto P = [&](int i, int j) {
            int k = n - i - j;  // weights: A k, B i, C j (all / n)
            return vid((A[0] * k) / n, (B[1] * i) / n, (C[2] * j) / n);
        };
        for (int i = 0; i < n; i++) {
            for (int j = 0; i + j < n; j++) {
                uint32_t p0 = P(i, j), p1 = P(i + 1, j), p2 = P(i, j + 1);
                flip ? m.add_t(p0, p2, p1) : m.add_t(p0, p1, p2);
                if (i + j < n - 1) {
                    uint32_t q0 = P(i + 1, j), q1 = P(i + 1, j + 1), q2 = P(i, j + 1);
                    flip ? m.add_t(q0, q2, q1) : m.add_t(q0, q1, q2);
                }
            }
        }
    }
    return m;
}

// Cube lattice of n x n quads per face with vertices shared along the seams,
// projected to a noise-displaced sphere: 12 n^2 triangles, closed.
MeshSource displaced_sphere(uint32_t n, float amp, float freq, uint32_t seed) {
    MeshSource m;
    std::unordered_map<uint64_t, uint32_t> lattice;
    auto vid = [&](int x, int y, int z) {
        uint64_t key = (uint64_t(uint32_t(x)) << 42) | (uint64_t(uint32_t(y)) << 21) | uint64_t(uint32_t(z));
        auto it = lattice.find(key);
        if (it != lattice.end()) return it->second;
        V3 c{2.0f * x / n - 1, 2.0f * y / n - 1, 2.0f * z / n - 1};
        V3 d = normalize(c);
        float r = 1.0f + amp * value_noise3(d.x * freq, d.y * freq, d.z * freq, seed);
        float u = std::atan2(d.z, d.x) * 0.15915494f + 0.5f;
        float v = std::asin(std::clamp(d.y, -1.0f, 1.0f)) * 0.31830989f + 0.5f;
        uint32_t id = m.nv();
        m.add_v(d * r, d, u, v);
        lattice.emplace(key, id);
        return id;
    };
    for (int f = 0; f < 6; f++) {
        int ax = f >> 1;
        bool pos = (f & 1) == 0;
        int u_ax = ax == 1 ? 2 : (ax == 2 ? 0 : 1);  // tangent axis
        V3 N = axis_vec(ax, pos ? 1.f : -1.f), U = axis_vec(u_ax, 1.f), V = cross(N, U);
        int v_ax = std::fabs(V.x) > 0.5f ? 0 : std::fabs(V.y) > 0.5f ? 1 : 2;
        float v_sign = (v_ax == 0 ? V.x : v_ax == 1 ? V.y : V.z);
        auto P = [&](int i, int j) {
            int c[3];
            c[ax] = pos ? int(n) : 0;
            c[u_ax] = i;
            c[v_ax] = v_sign > 0 ? j : int(n) - j;
            return vid(c[0], c[1], c[2]);
        };
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < n; j++) {
                uint32_t p00 = P(i, j), p10 = P(i + 1, j), p11 = P(i + 1, j + 1), p01 = P(i, j + 1);
                m.add_t(p00, p10, p11);
                m.add_t(p00, p11, p01);
            }
    }
    m.compute_normals();
    return m;
}

// Axis-aligned box [-1,1]^3, n x n quads per face, flat face normals: 12 n^2 triangles.
MeshSource box_mesh(uint32_t n) {
    MeshSource m;
    for (int f = 0; f < 6; f++) {
        int ax = f >> 1;
        bool pos = (f & 1) == 0;
        int u_ax = ax == 1 ? 2 : (ax == 2 ? 0 : 1);
        V3 N = axis_vec(ax, pos ? 1.f : -1.f), U = axis_vec(u_ax, 1.f), V = cross(N, U);
        uint32_t base = m.nv();
        for (uint32_t i = 0; i <= n; i++)
            for (uint32_t j = 0; j <= n; j++) {
                float a = 2.0f * i / n - 1, b = 2.0f * j / n - 1;
                m.add_v(N + U * a + V * b, N, float(i) / n, float(j) / n);
            }
        auto P = [&](uint32_t i, uint32_t j) { return base + i * (n + 1) + j; };
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < n; j++) {
                m.add_t(P(i, j), P(i + 1, j), P(i + 1, j + 1));
                m.add_t(P(i, j), P(i + 1, j + 1), P(i, j + 1));
            }
    }
    return m;
}

// Planar grid over [-1,1]^2 in xz with noise height: 2 nx nz triangles, facing +Y.
MeshSource grid_mesh(uint32_t nx, uint32_t nz, float amp, float freq, uint32_t seed) {
    MeshSource m;
    for (uint32_t i = 0; i <= nz; i++)
        for (uint32_t j = 0; j <= nx; j++) {
            float z = 2.0f * i / nz - 1, x = 2.0f * j / nx - 1;
            float h = amp * value_noise3(x * freq, 0.5f, z * freq, seed);
            m.add_v(V3{x, h, z}, V3{0, 1, 0}, float(j) / nx * 4.0f, float(i) / nz * 4.0f);
        }
    auto P = [&](uint32_t i, uint32_t j) { return i * (nx + 1) + j; };
    // U = +Z (i), V = +X (j): U x V = +Y.
    for (uint32_t i = 0; i < nz; i++)
        for (uint32_t j = 0; j < nx; j++) {
            m.add_t(P(i, j), P(i + 1, j), P(i + 1, j + 1));
            m.add_t(P(i, j), P(i + 1, j + 1), P(i, j + 1));
        }
    m.compute_normals();
    return m;
}

// ----------------------------------------------------------- bounding sphere

// Extremal points along the three axes give the initial diameter; the sphere
// is then grown to include every point (Ritter's method, the same steps as
// src/resources/bounding_sphere.h:3-74 and meshoptimizer's cluster bounds).
std::array<float, 4> bounding_sphere(const std::vector<V3>& pts) {
    if (pts.empty()) return {0, 0, 0, 0};
    size_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    auto comp = [](const V3& v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; };
    for (size_t i = 0; i < pts.size(); i++)
        for (int a = 0; a < 3; a++) {
            if (comp(pts[i], a) < comp(pts[lo[a]], a)) lo[a] = i;
            if (comp(pts[i], a) > comp(pts[hi[a]], a)) hi[a] = i;
        }
    int best = 0;
    float best_d2 = 0;
    for (int a = 0; a < 3; a++) {
        V3 d = pts[hi[a]] - pts[lo[a]];
        float d2 = dot(d, d);
        if (d2 > best_d2) {
            best_d2 = d2;
            best = a;
        }
    }
    V3 c = (pts[lo[best]] + pts[hi[best]]) * 0.5f;
    float r = std::sqrt(best_d2) * 0.5f;
    for (const V3& p : pts) {
        V3 d = p - c;
        float d2 = dot(d, d);
        if (d2 > r * r) {
            float dist = std::sqrt(d2);
            float k = 0.5f + (r / dist) * 0.5f;  // new centre moves towards p
            c = c * k + p * (1.0f - k);
            r = (r + dist) * 0.5f;
        }
    }
    return {c.x, c.y, c.z, r};
}

// --------------------------------------------------------------- meshlets

struct BuiltMesh {
    MeshSource src;
    std::vector<Meshlet> meshlets;
    std::vector<uint32_t> remap;
    std::vector<uint8_t> micro;
    std::array<float, 4> sphere{};
    uint64_t off_pos = 0, off_nrm = 0, off_uv = 0, off_remap = 0, off_micro = 0, off_meshlets = 0;
    // Quantized vertex format: src.pos holds u16 grid coordinates; a position
    // is grid_lo + step * q (folded into each instance's transform), a uv is
    // uv_lo + uv_step * q (folded into the mesh info's texture transform).
    bool quantized = false, index16 = false;
    float grid_lo[3] = {0, 0, 0}, step = 1.0f;
    float uv_lo[2] = {0, 0}, uv_step[2] = {1, 1};
};

// Grid step of n values spanning [lo, hi] in u16.
float u16_step(float lo, float hi) { return hi > lo ? (hi - lo) / 65535.0f : 1.0f; }
uint16_t to_u16(float x, float lo, float step) {
    return uint16_t(std::clamp(std::lround((x - lo) / step), 0l, 65535l));
}

// The reference's vertex format (KHR_mesh_quantization, read by
// src/shaders/common/util.glsl:11-16 and buffer_references.glsl:38-50):
// positions are u16x3 on a per-mesh grid whose origin and step go into each
// instance's transform, so the kernels see float(u16) mesh-space coordinates.
// One step for all three axes keeps directions (normal cones, bounding
// spheres) valid under the grid mapping.  The positions are replaced by their
// grid coordinates before meshlets and bounds are built, so every bound is in
// the space the kernels transform.
void quantize_positions(BuiltMesh& bm) {
    MeshSource& m = bm.src;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < m.nv(); v++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], m.pos[3 * v + k]);
            hi[k] = std::max(hi[k], m.pos[3 * v + k]);
        }
    float step = 0.0f;
    for (int k = 0; k < 3; k++) step = std::max(step, u16_step(lo[k], hi[k]));
    for (uint32_t v = 0; v < m.nv(); v++)
        for (int k = 0; k < 3; k++) m.pos[3 * v + k] = float(to_u16(m.pos[3 * v + k], lo[k], step));
    bm.quantized = true;
    for (int k = 0; k < 3; k++) bm.grid_lo[k] = lo[k];
    bm.step = step;
}

// Meshlet bounds (meshoptimizer's published cluster-bounds method).
void meshlet_bounds(const MeshSource& m, const uint32_t* remap, const uint8_t* micro, uint32_t ntri,
                    Meshlet& out) {
    std::vector<V3> corners, normals;
    std::vector<std::array<V3, 3>> tris;
    for (uint32_t t = 0; t < ntri; t++) {
        V3 p0 = m.p(remap[micro[3 * t]]), p1 = m.p(remap[micro[3 * t + 1]]), p2 = m.p(remap[micro[3 * t + 2]]);
        V3 n = cross(p1 - p0, p2 - p0);
        float area = length(n);
        if (area == 0.0f) continue;
        corners.insert(corners.end(), {p0, p1, p2});
        normals.push_back(n * (1.0f / area));
        tris.push_back({p0, p1, p2});
    }
    if (corners.empty()) {
        for (uint32_t t = 0; t < ntri; t++)
            for (int k = 0; k < 3; k++) corners.push_back(m.p(remap[micro[3 * t + k]]));
    }
    auto ps = bounding_sphere(corners);
    out.bounding_sphere = LhVec4{ps[0], ps[1], ps[2], ps[3]};
    out.cone_apex = LhVec3{0, 0, 0};
    out.cone_axis = LhVec3{0, 0, 0};
    out.cone_cutoff = 1.0f;  // never culls (normalize of a zero axis is NaN)
    if (normals.empty()) return;
    auto ns = bounding_sphere(normals);
    V3 axis{ns[0], ns[1], ns[2]};
    float al = length(axis);
    axis = al == 0.0f ? V3{0, 0, 0} : axis * (1.0f / al);
    float mindp = 1.0f;
    for (const V3& n : normals) mindp = std::min(mindp, dot(n, axis));
    if (mindp <= 0.1f) return;  // cone wider than ~168 degrees: not useful
    V3 centre{ps[0], ps[1], ps[2]};
    float maxt = 0.0f;
    for (size_t i = 0; i < tris.size(); i++) {
        // Find t with dot(centre - t*axis - corner, n) = 0; the apex lies on
        // the negative side of every triangle plane.
        float dc = dot(centre - tris[i][0], normals[i]);
        float dn = dot(axis, normals[i]);
        float t = dc / dn;
        maxt = std::max(maxt, t);
    }
    out.cone_apex = lv3(centre - axis * maxt);
    out.cone_axis = lv3(axis);
    out.cone_cutoff = std::sqrt(1.0f - mindp * mindp);
}

// Renumbers the vertices in first-use order over the meshlets (then any
// unreferenced ones), so that each meshlet's vertices are nearly contiguous in
// the position / normal / uv arrays: the raster's and the resolve's vertex
// gathers then touch few cache lines per meshlet instead of one per vertex.
// Geometry and meshlet contents are unchanged; only vertex numbering moves.
void reorder_vertices_for_fetch(BuiltMesh& bm) {
    MeshSource& m = bm.src;
    const uint32_t nv = m.nv();
    std::vector<uint32_t> new_of(nv, UINT32_MAX);
    uint32_t next = 0;
    for (uint32_t v : bm.remap)
        if (new_of[v] == UINT32_MAX) new_of[v] = next++;
    for (uint32_t v = 0; v < nv; v++)
        if (new_of[v] == UINT32_MAX) new_of[v] = next++;
    std::vector<float> pos(m.pos.size()), nrm(m.nrm.size()), uv(m.uv.size());
    for (uint32_t v = 0; v < nv; v++) {
        const uint32_t w = new_of[v];
        for (int k = 0; k < 3; k++) {
            pos[3 * w + k] = m.pos[3 * v + k];
            nrm[3 * w + k] = m.nrm[3 * v + k];
        }
        uv[2 * w] = m.uv[2 * v];
        uv[2 * w + 1] = m.uv[2 * v + 1];
    }
    m.pos.swap(pos);
    m.nrm.swap(nrm);
    m.uv.swap(uv);
    for (uint32_t& i : m.idx) i = new_of[i];
    for (uint32_t& r : bm.remap) r = new_of[r];
}

// Greedy meshlet builder: grows a meshlet from a seed triangle by repeatedly
// adding the adjacent triangle that introduces the fewest new vertices, within
// 64 vertices / 124 triangles.  The next seed is taken from the frontier of the
// previous meshlet so consecutive meshlets stay spatially coherent.
void build_meshlets(BuiltMesh& bm) {
    const MeshSource& m = bm.src;
    const uint32_t nv = m.nv(), nt = m.nt();
    std::vector<uint32_t> adj_off(nv + 1, 0), adj;
    for (uint32_t i : m.idx) adj_off[i + 1]++;
    for (uint32_t i = 0; i < nv; i++) adj_off[i + 1] += adj_off[i];
    adj.resize(adj_off[nv]);
    {
        std::vector<uint32_t> fill(adj_off.begin(), adj_off.end() - 1);
        for (uint32_t t = 0; t < nt; t++)
            for (int k = 0; k < 3; k++) adj[fill[m.idx[3 * t + k]]++] = t;
    }
    std::vector<uint8_t> used(nt, 0);
    std::vector<int32_t> local(nv, -1);
    std::vector<uint32_t> cand_stamp(nt, 0);
    std::vector<uint32_t> verts, tris, cand;
    uint32_t scan = 0, stamp = 0;
    int64_t next_seed = -1;
    while (true) {
        uint32_t seed;
        if (next_seed >= 0 && !used[next_seed]) {
            seed = uint32_t(next_seed);
        } else {
            while (scan < nt && used[scan]) scan++;
            if (scan == nt) break;
            seed = scan;
        }
        stamp++;
        verts.clear();
        tris.clear();
        cand.clear();
        auto add = [&](uint32_t t) {
            used[t] = 1;
            tris.push_back(t);
            for (int k = 0; k < 3; k++) {
                uint32_t v = m.idx[3 * t + k];
                if (local[v] < 0) {
                    local[v] = int32_t(verts.size());
                    verts.push_back(v);
                    for (uint32_t a = adj_off[v]; a < adj_off[v + 1]; a++) {
                        uint32_t u = adj[a];
                        if (!used[u] && cand_stamp[u] != stamp) {
                            cand_stamp[u] = stamp;
                            cand.push_back(u);
                        }
                    }
                }
            }
        };
        add(seed);
        while (tris.size() < MAX_MESHLET_TRIANGLES) {
            int64_t best = -1;
            uint32_t best_new = 4;
            size_t w = 0;
            for (size_t i = 0; i < cand.size(); i++) {
                uint32_t c = cand[i];
                if (used[c]) continue;
                cand[w++] = c;
                if (best_new == 0) continue;
                uint32_t nnew = 0;
                for (int k = 0; k < 3; k++) nnew += local[m.idx[3 * c + k]] < 0 ? 1u : 0u;
                if (nnew < best_new) {
                    best_new = nnew;
                    best = c;
                }
            }
            cand.resize(w);
            if (best < 0 || verts.size() + best_new > MAX_MESHLET_UNIQUE_VERTICES) break;
            add(uint32_t(best));
        }
        // Emit.
        Meshlet ml{};
        ml.index_offset = uint32_t(bm.remap.size());
        ml.triangle_offset = uint32_t(bm.micro.size());
        ml.triangle_count = uint8_t(tris.size());
        ml.index_count = uint8_t(verts.size());
        bm.remap.insert(bm.remap.end(), verts.begin(), verts.end());
        for (uint32_t t : tris)
            for (int k = 0; k < 3; k++) bm.micro.push_back(uint8_t(local[m.idx[3 * t + k]]));
        while (bm.micro.size() & 3) bm.micro.push_back(0);
        meshlet_bounds(m, bm.remap.data() + ml.index_offset, bm.micro.data() + ml.triangle_offset,
                       ml.triangle_count, ml);
        bm.meshlets.push_back(ml);
        for (uint32_t v : verts) local[v] = -1;
        next_seed = -1;
        for (uint32_t c : cand)
            if (!used[c] && (next_seed < 0 || c < next_seed)) next_seed = c;
    }
    std::vector<V3> all(nv);
    for (uint32_t i = 0; i < nv; i++) all[i] = m.p(i);
    bm.sphere = bounding_sphere(all);
    reorder_vertices_for_fetch(bm);
}

// ----------------------------------------------------------------- textures

struct TextureData {
    uint32_t w = 0, h = 0, levels = 0;
    std::vector<uint32_t> level_offset;  // texels (or blocks of an encoded texture)
    std::vector<uint8_t> rgba;           // all levels (the stored data of an encoded texture)
    bool encoded = false;                // stored as given: `format`, `rgba` holds the levels
    uint32_t format = LH_TEXTURE_RGBA8;
};

// BC1 (S3TC) encoder of the producer.  Per 4x4 block the endpoints are the
// texels of largest and smallest luma (r + 2g + b), rounded to RGB565 and
// ordered c0 > c1 (four-colour mode); every texel takes the palette entry, as
// the decoder forms it (endpoints widened by bit replication, thirds rounded
// to nearest), nearest in squared RGB distance.  Texels beyond the level's
// edge repeat the edge; alpha is dropped (opaque).
uint16_t to565(const uint8_t* p) {
    return uint16_t((((p[0] * 31u + 127u) / 255u) << 11) | (((p[1] * 63u + 127u) / 255u) << 5) |
                    ((p[2] * 31u + 127u) / 255u));
}
void bc1_palette(uint16_t c0, uint16_t c1, int pal[4][3]) {
    auto widen = [](uint32_t v, int o[3]) {
        uint32_t r = (v >> 11) & 31u, g = (v >> 5) & 63u, b = v & 31u;
        o[0] = int((r << 3) | (r >> 2));
        o[1] = int((g << 2) | (g >> 4));
        o[2] = int((b << 3) | (b >> 2));
    };
    widen(c0, pal[0]);
    widen(c1, pal[1]);
    for (int k = 0; k < 3; k++) {
        pal[2][k] = (2 * pal[0][k] + pal[1][k] + 1) / 3;
        pal[3][k] = (pal[0][k] + 2 * pal[1][k] + 1) / 3;
    }
}
void encode_bc1_level(const uint8_t* rgba, uint32_t w, uint32_t h, std::vector<uint8_t>& out) {
    for (uint32_t by = 0; by < (h + 3) / 4; by++)
        for (uint32_t bx = 0; bx < (w + 3) / 4; bx++) {
            const uint8_t* px[16];
            for (uint32_t k = 0; k < 16; k++) {
                uint32_t x = std::min(w - 1, bx * 4 + (k & 3)), y = std::min(h - 1, by * 4 + (k >> 2));
                px[k] = rgba + 4 * (size_t(y) * w + x);
            }
            int lo = 0, hi = 0;
            auto luma = [&](int k) { return int(px[k][0]) + 2 * int(px[k][1]) + int(px[k][2]); };
            for (int k = 1; k < 16; k++) {
                if (luma(k) < luma(lo)) lo = k;
                if (luma(k) > luma(hi)) hi = k;
            }
            uint16_t c0 = to565(px[hi]), c1 = to565(px[lo]);
            if (c0 < c1) std::swap(c0, c1);
            uint32_t idx = 0;
            if (c0 != c1) {
                int pal[4][3];
                bc1_palette(c0, c1, pal);
                for (int k = 0; k < 16; k++) {
                    int best = 0, bd = 1 << 30;
                    for (int s = 0; s < 4; s++) {
                        int d = 0;
                        for (int c = 0; c < 3; c++) d += (int(px[k][c]) - pal[s][c]) * (int(px[k][c]) - pal[s][c]);
                        if (d < bd) {
                            bd = d;
                            best = s;
                        }
                    }
                    idx |= uint32_t(best) << (2 * k);
                }
            }
            const uint8_t blk[8] = {uint8_t(c0), uint8_t(c0 >> 8), uint8_t(c1), uint8_t(c1 >> 8),
                                    uint8_t(idx), uint8_t(idx >> 8), uint8_t(idx >> 16), uint8_t(idx >> 24)};
            out.insert(out.end(), blk, blk + 8);
        }
}

// BC7 encoder of the producer: mode 6 only (one subset; RGBA endpoints of 7
// bits plus one p-bit per endpoint, 4-bit indices).  Per 4x4 block the
// endpoints are the extremes of the texels projected on the principal axis
// of their RGBA covariance (power iteration), each channel rounded to the
// nearest value whose lowest bit is the endpoint's p-bit (the majority of its
// channels' lowest bits); every texel takes the index whose interpolated RGBA,
// as the decoder forms it, is nearest in squared distance.  Texel 0 is the
// anchor and stores 3 index bits, so when its index exceeds 7 the endpoints
// swap and every index mirrors to 15 - i (the weights are symmetric).  Texels
// beyond the level's edge repeat the edge.
struct Bits128 {
    uint64_t w[2] = {0, 0};
    uint32_t pos = 0;
    void put(uint32_t v, uint32_t n) {
        for (uint32_t k = 0; k < n; k++, pos++)
            if ((v >> k) & 1u) w[pos >> 6] |= 1ull << (pos & 63);
    }
};
void encode_bc7_level(const uint8_t* rgba, uint32_t w, uint32_t h, std::vector<uint8_t>& out) {
    static const int kW4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};
    for (uint32_t by = 0; by < (h + 3) / 4; by++)
        for (uint32_t bx = 0; bx < (w + 3) / 4; bx++) {
            const uint8_t* px[16];
            for (uint32_t k = 0; k < 16; k++) {
                uint32_t x = std::min(w - 1, bx * 4 + (k & 3)), y = std::min(h - 1, by * 4 + (k >> 2));
                px[k] = rgba + 4 * (size_t(y) * w + x);
            }
            double mean[4] = {0, 0, 0, 0}, cov[4][4] = {};
            for (int k = 0; k < 16; k++)
                for (int c = 0; c < 4; c++) mean[c] += px[k][c] / 16.0;
            for (int k = 0; k < 16; k++)
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++) cov[a][b] += (px[k][a] - mean[a]) * (px[k][b] - mean[b]);
            double axis[4] = {0, 0, 0, 0};
            int widest = 0;
            for (int c = 1; c < 4; c++)
                if (cov[c][c] > cov[widest][widest]) widest = c;
            axis[widest] = 1.0;
            for (int it = 0; it < 16; it++) {
                double nx[4] = {0, 0, 0, 0}, len = 0;
                for (int a = 0; a < 4; a++) {
                    for (int b = 0; b < 4; b++) nx[a] += cov[a][b] * axis[b];
                    len += nx[a] * nx[a];
                }
                if (len <= 1e-12) break;
                for (int a = 0; a < 4; a++) axis[a] = nx[a] / std::sqrt(len);
            }
            double tmin = 0, tmax = 0;
            for (int k = 0; k < 16; k++) {
                double t = 0;
                for (int c = 0; c < 4; c++) t += (px[k][c] - mean[c]) * axis[c];
                tmin = std::min(tmin, t);
                tmax = std::max(tmax, t);
            }
            int ep[2][4];
            for (int c = 0; c < 4; c++) {
                ep[0][c] = int(std::lround(std::clamp(mean[c] + tmin * axis[c], 0.0, 255.0)));
                ep[1][c] = int(std::lround(std::clamp(mean[c] + tmax * axis[c], 0.0, 255.0)));
            }
            uint32_t pbit[2];
            for (int e = 0; e < 2; e++) {
                int odd = 0;
                for (int c = 0; c < 4; c++) odd += ep[e][c] & 1;
                pbit[e] = odd >= 2 ? 1u : 0u;
                for (int c = 0; c < 4; c++)
                    if (uint32_t(ep[e][c] & 1) != pbit[e]) ep[e][c] += ep[e][c] == 255 ? -1 : 1;
            }
            int pal[16][4];
            for (int i = 0; i < 16; i++)
                for (int c = 0; c < 4; c++) pal[i][c] = ((64 - kW4[i]) * ep[0][c] + kW4[i] * ep[1][c] + 32) >> 6;
            uint32_t idx[16];
            for (int k = 0; k < 16; k++) {
                int best = 0, bd = 1 << 30;
                for (int i = 0; i < 16; i++) {
                    int d = 0;
                    for (int c = 0; c < 4; c++) d += (int(px[k][c]) - pal[i][c]) * (int(px[k][c]) - pal[i][c]);
                    if (d < bd) {
                        bd = d;
                        best = i;
                    }
                }
                idx[k] = uint32_t(best);
            }
            if (idx[0] >= 8) {
                for (int c = 0; c < 4; c++) std::swap(ep[0][c], ep[1][c]);
                std::swap(pbit[0], pbit[1]);
                for (int k = 0; k < 16; k++) idx[k] = 15u - idx[k];
            }
            Bits128 b;
            b.put(1u << 6, 7);  // mode 6
            for (int c = 0; c < 4; c++)
                for (int e = 0; e < 2; e++) b.put(uint32_t(ep[e][c]) >> 1, 7);
            b.put(pbit[0], 1);
            b.put(pbit[1], 1);
            b.put(idx[0], 3);
            for (int k = 1; k < 16; k++) b.put(idx[k], 4);
            uint8_t blk[16];
            std::memcpy(blk, b.w, 16);
            out.insert(out.end(), blk, blk + 16);
        }
}

void build_mips(TextureData& t) {
    uint32_t w = t.w, h = t.h;
    t.level_offset = {0};
    t.levels = 1;
    while ((w > 1 || h > 1) && t.levels < LH_MAX_TEXTURE_LEVELS) {
        uint32_t nw = std::max(1u, w / 2), nh = std::max(1u, h / 2);
        uint32_t src = t.level_offset.back();
        uint32_t dst = src + w * h;
        t.rgba.resize(size_t(dst + nw * nh) * 4);
        for (uint32_t y = 0; y < nh; y++)
            for (uint32_t x = 0; x < nw; x++)
                for (int c = 0; c < 4; c++) {
                    uint32_t acc = 0;
                    for (int k = 0; k < 4; k++) {
                        uint32_t sx = std::min(w - 1, 2 * x + (k & 1)), sy = std::min(h - 1, 2 * y + (k >> 1));
                        acc += t.rgba[size_t(src + sy * w + sx) * 4 + c];
                    }
                    t.rgba[size_t(dst + y * nw + x) * 4 + c] = uint8_t((acc + 2) / 4);
                }
        t.level_offset.push_back(dst);
        w = nw;
        h = nh;
        t.levels++;
    }
}

// Tileable value noise, three octaves, tinted by a per-texture colour pair.
void value_noise_texture(TextureData& t, uint32_t size, uint32_t seed) {
    t.w = t.h = size;
    t.rgba.assign(size_t(size) * size * 4, 255);
    Rng rng(seed);
    float ca[3], cb[3];
    for (int c = 0; c < 3; c++) {
        ca[c] = float(rng.uniform(0.05, 0.95));
        cb[c] = float(rng.uniform(0.05, 0.95));
    }
    const uint32_t cells[3] = {8, 32, 128};
    const float amps[3] = {0.6f, 0.3f, 0.1f};
    auto lattice = [&](uint32_t o, uint32_t x, uint32_t y) {
        uint32_t c = cells[o];
        return float(hash32((x % c) * 2654435761u ^ (y % c) * 40503u ^ (seed * 97u + o)) & 0xFFFF) / 65535.0f;
    };
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++) {
            float v = 0;
            for (uint32_t o = 0; o < 3; o++) {
                float fx = float(x) * cells[o] / size, fy = float(y) * cells[o] / size;
                uint32_t ix = uint32_t(fx), iy = uint32_t(fy);
                float tx = fx - ix, ty = fy - iy;
                tx = tx * tx * (3 - 2 * tx);
                ty = ty * ty * (3 - 2 * ty);
                float a = lattice(o, ix, iy), b = lattice(o, ix + 1, iy), c = lattice(o, ix, iy + 1),
                      d = lattice(o, ix + 1, iy + 1);
                v += amps[o] * ((a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty);
            }
            uint8_t* px = &t.rgba[(size_t(y) * size + x) * 4];
            for (int c = 0; c < 3; c++) px[c] = uint8_t(std::lround(255.0f * (ca[c] * (1 - v) + cb[c] * v)));
            px[3] = 255;
        }
}

struct MaterialSpec {
    LhVec3 base_color;
    float roughness, metallic;
    uint32_t tex_base, tex_mr, tex_normal;
    LhVec2 tscale, toffset;
    bool alpha_clip = false;  // MESH_INFO_FLAGS_ALPHA_CLIP: double-sided, base-colour alpha < 0.5 discarded
};

struct InstanceSpec {
    uint32_t mesh, material;
    LhMat4 transform;
};

// Column-major inverse-transpose of the upper 3x3 (normal matrix).
LhMat3 normal_matrix(const LhMat4& m) {
    double a[3][3];  // a[row][col]
    for (int c = 0; c < 3; c++) {
        a[0][c] = m.c[c].x;
        a[1][c] = m.c[c].y;
        a[2][c] = m.c[c].z;
    }
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    // transpose(inverse): element (row r, col c) = inv[c][r]
    LhMat3 out;
    for (int c = 0; c < 3; c++) out.c[c] = LhVec3{float(inv[c][0]), float(inv[c][1]), float(inv[c][2])};
    return out;
}

LhMat4 inverse4(const LhMat4& m) {
    double a[16];  // column major
    for (int c = 0; c < 4; c++) {
        a[4 * c + 0] = m.c[c].x;
        a[4 * c + 1] = m.c[c].y;
        a[4 * c + 2] = m.c[c].z;
        a[4 * c + 3] = m.c[c].w;
    }
    double inv[16];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    LhMat4 out;
    for (int c = 0; c < 4; c++)
        out.c[c] = LhVec4{float(inv[4 * c] / det), float(inv[4 * c + 1] / det), float(inv[4 * c + 2] / det),
                          float(inv[4 * c + 3] / det)};
    return out;
}

LhMat4 mul4(const LhMat4& a, const LhMat4& b) {
    LhMat4 r;
    for (int c = 0; c < 4; c++) {
        const LhVec4& v = b.c[c];
        double o[4];
        for (int k = 0; k < 4; k++) {
            auto row = [&](int col) {
                const LhVec4& cc = a.c[col];
                return k == 0 ? cc.x : k == 1 ? cc.y : k == 2 ? cc.z : cc.w;
            };
            o[k] = double(row(0)) * v.x + double(row(1)) * v.y + double(row(2)) * v.z + double(row(3)) * v.w;
        }
        r.c[c] = LhVec4{float(o[0]), float(o[1]), float(o[2]), float(o[3])};
    }
    return r;
}

}  // namespace

// ------------------------------------------------------------------ scene

struct lh_scene_host {
    std::string cache_dir;  // on-disk cache of built meshes; empty = none
    uint32_t vertex_format = LH_VERTEX_FORMAT_FP32;
    uint32_t texture_format = LH_TEXTURE_RGBA8;
    std::vector<std::unique_ptr<BuiltMesh>> meshes;
    std::vector<TextureData> textures;
    std::vector<MaterialSpec> materials;
    std::vector<InstanceSpec> instances;
    bool finished = false;
    // finished arrays
    std::vector<uint8_t> geometry;
    std::vector<MeshInfo> mesh_infos;
    std::vector<LhMaterialFactors> factors;
    std::vector<Instance> inst_out;
    std::vector<uint32_t> meshlet_base;
    std::vector<MeshletReference> meshlet_table;
    std::vector<uint8_t> texels;
    std::vector<LhTexture> tex_out;
    uint64_t total_tris = 0, unique_tris = 0;
};

namespace {

template <class T>
uint64_t append_bytes(std::vector<uint8_t>& arena, const std::vector<T>& v) {
    uint64_t off = (arena.size() + 255) & ~uint64_t(255);  // 256-byte aligned sections
    arena.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(arena.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// ---- on-disk cache of built meshes (the reference's FsCache,
// src/resources/fs_cache.h:3-25, which stores built meshlets per asset).
// Key: FNV-1a 64 of the source arrays, the vertex format, the file version
// and the builder's identity (its limits and revision); file:
// <dir>/lhmesh-<key>.bin holding the built mesh (reordered vertices,
// meshlets, remap, micro indices, bounds, quantization grid).  A file whose
// checksum, sizes or meshlet limits do not hold is ignored (the mesh is
// rebuilt).  kBuilderRevision changes with the meshlet builder's output.
constexpr uint32_t kCacheMagic = 0x434D484Cu;  // "LHMC"
constexpr uint32_t kCacheVersion = 1;
constexpr uint32_t kBuilderRevision = 1;
constexpr uint32_t kBuilderIdentity[3] = {kBuilderRevision, MAX_MESHLET_UNIQUE_VERTICES, MAX_MESHLET_TRIANGLES};

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    for (size_t i = 0; i < n; i++) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}
template <class T>
uint64_t fnv1a_vec(const std::vector<T>& v, uint64_t h) {
    const uint64_t n = v.size();
    h = fnv1a(&n, sizeof n, h);
    return v.empty() ? h : fnv1a(v.data(), v.size() * sizeof(T), h);
}
std::string cache_path(const lh_scene_host* s, const MeshSource& m) {
    uint64_t h = fnv1a(&kCacheVersion, sizeof kCacheVersion);
    h = fnv1a(kBuilderIdentity, sizeof kBuilderIdentity, h);
    h = fnv1a(&s->vertex_format, sizeof s->vertex_format, h);
    h = fnv1a_vec(m.pos, h);
    h = fnv1a_vec(m.nrm, h);
    h = fnv1a_vec(m.uv, h);
    h = fnv1a_vec(m.idx, h);
    char name[64];
    std::snprintf(name, sizeof name, "/lhmesh-%016llx.bin", static_cast<unsigned long long>(h));
    return s->cache_dir + name;
}
// Files: magic, version, the arrays (u64 count + elements), bounds and grid,
// then an FNV-1a 64 of everything before it.
template <class T>
void put(std::vector<uint8_t>& out, const T* p, size_t n) {
    const uint8_t* b = reinterpret_cast<const uint8_t*>(p);
    out.insert(out.end(), b, b + n * sizeof(T));
}
template <class T>
void put_vec(std::vector<uint8_t>& out, const std::vector<T>& v) {
    const uint64_t n = v.size();
    put(out, &n, 1);
    put(out, v.data(), v.size());
}
struct Reader {
    const std::vector<uint8_t>& b;
    size_t at = 0;
    template <class T>
    bool get(T* p, size_t n) {
        if (n * sizeof(T) > b.size() - at) return false;
        std::memcpy(p, b.data() + at, n * sizeof(T));
        at += n * sizeof(T);
        return true;
    }
    template <class T>
    bool get_vec(std::vector<T>& v, uint64_t max_n) {
        uint64_t n = 0;
        if (!get(&n, 1) || n > max_n) return false;
        v.resize(n);
        return get(v.data(), n);
    }
};
void save_cached(const std::string& path, const BuiltMesh& bm) {
    std::vector<uint8_t> out;
    const uint32_t head[2] = {kCacheMagic, kCacheVersion};
    put(out, head, 2);
    put_vec(out, bm.src.pos);
    put_vec(out, bm.src.nrm);
    put_vec(out, bm.src.uv);
    put_vec(out, bm.src.idx);
    put_vec(out, bm.meshlets);
    put_vec(out, bm.remap);
    put_vec(out, bm.micro);
    const uint8_t q = bm.quantized ? 1 : 0;
    put(out, bm.sphere.data(), 4);
    put(out, &q, 1);
    put(out, bm.grid_lo, 3);
    put(out, &bm.step, 1);
    const uint64_t h = fnv1a(out.data(), out.size());
    put(out, &h, 1);
    const std::string tmp = path + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) == 0 && ok)
        std::rename(tmp.c_str(), path.c_str());
    else
        std::remove(tmp.c_str());
}
bool load_cached(const std::string& path, BuiltMesh& bm, const MeshSource& src) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<uint8_t> buf;
    uint8_t chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    std::fclose(f);
    if (buf.size() < 16) return false;
    uint64_t h = 0;
    std::memcpy(&h, buf.data() + buf.size() - 8, 8);
    if (h != fnv1a(buf.data(), buf.size() - 8)) return false;  // damaged or truncated
    buf.resize(buf.size() - 8);
    BuiltMesh c;
    Reader r{buf};
    uint32_t head[2] = {0, 0};
    uint8_t q = 0;
    const uint64_t nv = src.nv(), nt = src.nt();
    bool ok = r.get(head, 2) && head[0] == kCacheMagic && head[1] == kCacheVersion && r.get_vec(c.src.pos, 3 * nv) &&
              r.get_vec(c.src.nrm, 3 * nv) && r.get_vec(c.src.uv, 2 * nv) && r.get_vec(c.src.idx, 3 * nt) &&
              r.get_vec(c.meshlets, nt) && r.get_vec(c.remap, 64 * nt + 64) && r.get_vec(c.micro, 8 * nt) &&
              r.get(c.sphere.data(), 4) && r.get(&q, 1) && r.get(c.grid_lo, 3) && r.get(&c.step, 1) &&
              r.at == buf.size();
    ok = ok && c.src.pos.size() == 3 * nv && c.src.nrm.size() == 3 * nv && c.src.uv.size() == 2 * nv &&
         c.src.idx.size() == 3 * nt && !c.meshlets.empty();
    // every meshlet within the limits the raster assumes (<= 64 vertices,
    // <= 124 triangles), its ranges within the stored arrays, every micro
    // index below its vertex count, every remap entry a vertex
    for (size_t k = 0; ok && k < c.meshlets.size(); k++) {
        const Meshlet& ml = c.meshlets[k];
        ok = ml.index_count > 0 && ml.index_count <= MAX_MESHLET_UNIQUE_VERTICES && ml.triangle_count > 0 &&
             ml.triangle_count <= MAX_MESHLET_TRIANGLES && size_t(ml.index_offset) + ml.index_count <= c.remap.size() &&
             size_t(ml.triangle_offset) + 3u * ml.triangle_count <= c.micro.size();
        for (uint32_t t = 0; ok && t < 3u * ml.triangle_count; t++) ok = c.micro[ml.triangle_offset + t] < ml.index_count;
    }
    for (size_t k = 0; ok && k < c.remap.size(); k++) ok = c.remap[k] < nv;
    if (!ok) return false;
    c.quantized = q != 0;
    bm = std::move(c);
    return true;
}

int32_t add_mesh(lh_scene_host* s, MeshSource&& src) {
    if (s->finished || src.nt() == 0) return -1;
    auto bm = std::make_unique<BuiltMesh>();
    const std::string path = s->cache_dir.empty() ? std::string() : cache_path(s, src);
    if (path.empty() || !load_cached(path, *bm, src)) {
        bm->src = std::move(src);
        if (s->vertex_format == LH_VERTEX_FORMAT_QUANTIZED) quantize_positions(*bm);
        build_meshlets(*bm);
        if (!path.empty()) save_cached(path, *bm);
    }
    if (bm->meshlets.size() >= (1u << 16)) return -1;  // MeshInfo.num_meshlets is u16
    s->meshes.push_back(std::move(bm));
    return int32_t(s->meshes.size() - 1);
}

}  // namespace

extern "C" {

lh_scene_host* lh_scene_create(void) { return new lh_scene_host(); }

int32_t lh_scene_set_cache_dir(lh_scene_host* s, const char* dir) {
    if (!s || s->finished || !s->meshes.empty()) return -1;
    s->cache_dir = dir ? dir : "";
    return 0;
}

int32_t lh_scene_set_texture_format(lh_scene_host* s, uint32_t format) {
    if (!s || s->finished || (format & ~(LH_TEXTURE_FORMAT_MASK | LH_TEXTURE_SRGB)) ||
        (format & LH_TEXTURE_FORMAT_MASK) > LH_TEXTURE_BC7)
        return -1;
    s->texture_format = format;
    return 0;
}

int32_t lh_scene_set_vertex_format(lh_scene_host* s, uint32_t format) {
    if (!s || s->finished || !s->meshes.empty() || format > LH_VERTEX_FORMAT_QUANTIZED) return -1;
    s->vertex_format = format;
    return 0;
}

void lh_scene_destroy(lh_scene_host* s) { delete s; }

int32_t lh_scene_add_mesh(lh_scene_host* s, const float* positions, const float* normals, const float* uvs,
                          uint32_t num_vertices, const uint32_t* indices, uint32_t num_triangles) {
    if (!s || !positions || !indices || num_vertices == 0) return -1;
    MeshSource m;
    m.pos.assign(positions, positions + 3 * size_t(num_vertices));
    m.nrm = normals ? std::vector<float>(normals, normals + 3 * size_t(num_vertices))
                    : std::vector<float>(3 * size_t(num_vertices), 0.0f);
    m.uv = uvs ? std::vector<float>(uvs, uvs + 2 * size_t(num_vertices)) : std::vector<float>(2 * size_t(num_vertices), 0.0f);
    m.idx.assign(indices, indices + 3 * size_t(num_triangles));
    for (uint32_t i : m.idx)
        if (i >= num_vertices) return -1;
    if (!normals) m.compute_normals();
    return add_mesh(s, std::move(m));
}

int32_t lh_scene_add_texture(lh_scene_host* s, const uint8_t* rgba, uint32_t width, uint32_t height) {
    if (!s || s->finished || !rgba || width == 0 || height == 0) return -1;
    TextureData t;
    t.w = width;
    t.h = height;
    t.rgba.assign(rgba, rgba + size_t(width) * height * 4);
    build_mips(t);
    s->textures.push_back(std::move(t));
    return int32_t(s->textures.size() - 1);
}

int32_t lh_scene_add_texture_encoded(lh_scene_host* s, uint32_t format, uint32_t width, uint32_t height,
                                     uint32_t levels, const uint8_t* data, uint64_t bytes) {
    if (!s || s->finished || !data || width == 0 || height == 0 || levels == 0 || levels > LH_MAX_TEXTURE_LEVELS ||
        (format & ~(LH_TEXTURE_FORMAT_MASK | LH_TEXTURE_SRGB)) || (format & LH_TEXTURE_FORMAT_MASK) > LH_TEXTURE_BC7)
        return -1;
    // A chain may stop early, but not beyond 1 x 1.
    if (levels > 1 && (std::max(width, height) >> (levels - 1)) == 0) return -1;
    const uint32_t fmt = format & LH_TEXTURE_FORMAT_MASK;
    const uint64_t unit = fmt == LH_TEXTURE_BC1 ? 8 : fmt == LH_TEXTURE_BC7 ? 16 : 4;
    TextureData t;
    t.w = width;
    t.h = height;
    t.levels = levels;
    t.encoded = true;
    t.format = format;
    uint64_t n = 0;
    for (uint32_t l = 0; l < levels; l++) {
        const uint64_t w = std::max(1u, width >> l), h = std::max(1u, height >> l);
        if (n > 0xFFFFFFFFull) return -1;
        t.level_offset.push_back(uint32_t(n));
        n += fmt == LH_TEXTURE_RGBA8 ? w * h : ((w + 3) / 4) * ((h + 3) / 4);
    }
    if (n * unit != bytes) return -1;
    t.rgba.assign(data, data + bytes);
    s->textures.push_back(std::move(t));
    return int32_t(s->textures.size() - 1);
}

int32_t lh_scene_add_material(lh_scene_host* s, LhVec3 base_color_factor, float roughness_factor,
                              float metallic_factor, uint32_t base_color_texture, uint32_t metallic_roughness_texture,
                              uint32_t normal_texture, LhVec2 texture_scale, LhVec2 texture_offset) {
    if (!s || s->finished) return -1;
    s->materials.push_back(MaterialSpec{base_color_factor, roughness_factor, metallic_factor, base_color_texture,
                                        metallic_roughness_texture, normal_texture, texture_scale, texture_offset});
    return int32_t(s->materials.size() - 1);
}

int32_t lh_scene_set_material_alpha_clip(lh_scene_host* s, uint32_t material, uint32_t enabled) {
    if (!s || s->finished || material >= s->materials.size()) return -1;
    s->materials[material].alpha_clip = enabled != 0;
    return 0;
}

int32_t lh_scene_add_instance(lh_scene_host* s, uint32_t mesh, uint32_t material, const LhMat4* transform) {
    if (!s || s->finished || mesh >= s->meshes.size() || material >= s->materials.size() || !transform) return -1;
    s->instances.push_back(InstanceSpec{mesh, material, *transform});
    return int32_t(s->instances.size() - 1);
}

int32_t lh_scene_finish(lh_scene_host* s) {
    if (!s || s->finished) return -1;
    for (const auto& m : s->materials) {
        for (uint32_t t : {m.tex_base, m.tex_mr, m.tex_normal})
            if (t != UNUSED_TEXTURE_INDEX && t >= s->textures.size()) return -1;
    }
    // Geometry arena.
    for (auto& bm : s->meshes) {
        const MeshSource& m = bm->src;
        if (bm->quantized) {
            std::vector<uint16_t> qp(m.pos.size()), qu(m.uv.size());
            std::vector<int8_t> qn(m.nrm.size());
            for (size_t i = 0; i < m.pos.size(); i++) qp[i] = uint16_t(m.pos[i]);  // grid coordinates already
            for (size_t i = 0; i < m.nrm.size(); i++)
                qn[i] = int8_t(std::lround(std::clamp(m.nrm[i], -1.0f, 1.0f) * 127.0f));
            for (int k = 0; k < 2; k++) {
                float lo = INFINITY, hi = -INFINITY;
                for (uint32_t v = 0; v < m.nv(); v++) {
                    lo = std::min(lo, m.uv[2 * v + k]);
                    hi = std::max(hi, m.uv[2 * v + k]);
                }
                bm->uv_lo[k] = lo;
                bm->uv_step[k] = u16_step(lo, hi);
                for (uint32_t v = 0; v < m.nv(); v++) qu[2 * v + k] = to_u16(m.uv[2 * v + k], lo, bm->uv_step[k]);
            }
            bm->off_pos = append_bytes(s->geometry, qp);
            bm->off_nrm = append_bytes(s->geometry, qn);
            bm->off_uv = append_bytes(s->geometry, qu);
            bm->index16 = m.nv() <= 65536;
            if (bm->index16) {
                std::vector<uint16_t> r16(bm->remap.begin(), bm->remap.end());
                bm->off_remap = append_bytes(s->geometry, r16);
            } else {
                bm->off_remap = append_bytes(s->geometry, bm->remap);
            }
        } else {
            bm->off_pos = append_bytes(s->geometry, m.pos);
            bm->off_nrm = append_bytes(s->geometry, m.nrm);
            bm->off_uv = append_bytes(s->geometry, m.uv);
            bm->off_remap = append_bytes(s->geometry, bm->remap);
        }
        bm->off_micro = append_bytes(s->geometry, bm->micro);
        bm->off_meshlets = append_bytes(s->geometry, bm->meshlets);
        s->unique_tris += bm->src.nt();
    }
    s->geometry.resize((s->geometry.size() + 255) & ~size_t(255));
    // One MeshInfo per used (mesh, material) pair, in first-use order.
    std::unordered_map<uint64_t, uint32_t> combo;
    std::vector<uint32_t> inst_mi(s->instances.size());
    for (size_t i = 0; i < s->instances.size(); i++) {
        const auto& is = s->instances[i];
        uint64_t key = (uint64_t(is.mesh) << 32) | is.material;
        auto it = combo.find(key);
        if (it == combo.end()) {
            const BuiltMesh& bm = *s->meshes[is.mesh];
            const MaterialSpec& ms = s->materials[is.material];
            MeshInfo mi{};
            mi.positions = bm.off_pos;
            mi.indices = bm.off_remap;
            mi.normals = bm.off_nrm;
            mi.uvs = bm.off_uv;
            mi.micro_indices = bm.off_micro;
            mi.meshlets = bm.off_meshlets;
            mi.num_meshlets = uint16_t(bm.meshlets.size());
            mi.flags = bm.quantized ? (bm.index16 ? 0u : MESH_INFO_FLAGS_32_BIT_INDICES)
                                    : MESH_INFO_FLAGS_32_BIT_INDICES | MESH_INFO_FLAGS_FP32_VERTICES;
            if (ms.alpha_clip) mi.flags |= MESH_INFO_FLAGS_ALPHA_CLIP;
            mi.bounding_sphere = LhVec4{bm.sphere[0], bm.sphere[1], bm.sphere[2], bm.sphere[3]};
            if (bm.quantized) {  // uv = (uv_lo + uv_step q) * scale + offset
                mi.texture_scale = LhVec2{bm.uv_step[0] * ms.tscale.x, bm.uv_step[1] * ms.tscale.y};
                mi.texture_offset = LhVec2{bm.uv_lo[0] * ms.tscale.x + ms.toffset.x,
                                           bm.uv_lo[1] * ms.tscale.y + ms.toffset.y};
            } else {
                mi.texture_scale = ms.tscale;
                mi.texture_offset = ms.toffset;
            }
            mi.base_color_texture_index = uint16_t(ms.tex_base);
            mi.metallic_roughness_texture_index = uint16_t(ms.tex_mr);
            mi.normal_texture_index = uint16_t(ms.tex_normal);
            mi.base_color_factor = ms.base_color;
            it = combo.emplace(key, uint32_t(s->mesh_infos.size())).first;
            s->mesh_infos.push_back(mi);
            s->factors.push_back(LhMaterialFactors{ms.roughness, ms.metallic});
        }
        inst_mi[i] = it->second;
    }
    // Instances, stable meshlet ids.
    uint64_t base = 0;
    for (size_t i = 0; i < s->instances.size(); i++) {
        const auto& is = s->instances[i];
        Instance in{};
        in.transform = is.transform;
        const BuiltMesh& bm = *s->meshes[is.mesh];
        if (bm.quantized) {  // transform * translate(grid_lo) * scale(step)
            LhMat4 g{};
            g.c[0] = LhVec4{bm.step, 0, 0, 0};
            g.c[1] = LhVec4{0, bm.step, 0, 0};
            g.c[2] = LhVec4{0, 0, bm.step, 0};
            g.c[3] = LhVec4{bm.grid_lo[0], bm.grid_lo[1], bm.grid_lo[2], 1};
            in.transform = mul4(is.transform, g);
        }
        in.mesh_info_address = inst_mi[i];
        in.normal_transform = normal_matrix(in.transform);
        s->inst_out.push_back(in);
        s->meshlet_base.push_back(uint32_t(base));
        uint32_t n = s->mesh_infos[inst_mi[i]].num_meshlets;
        for (uint32_t k = 0; k < n; k++) s->meshlet_table.push_back(MeshletReference{uint32_t(i), uint16_t(k)});
        base += n;
        s->total_tris += s->meshes[is.mesh]->src.nt();
        if (base > LH_MAX_SCENE_MESHLETS) return -2;
    }
    // Texture arena.
    for (const auto& t : s->textures) {
        LhTexture lt{};
        lt.width = t.w;
        lt.height = t.h;
        lt.levels = t.levels;
        lt.format = t.encoded ? t.format : s->texture_format;
        const uint32_t fmt = lt.format & LH_TEXTURE_FORMAT_MASK;
        if (t.encoded) {
            lt.texels = append_bytes(s->texels, t.rgba);
            for (uint32_t l = 0; l < t.levels; l++) lt.level_offset[l] = t.level_offset[l];
        } else if (fmt == LH_TEXTURE_BC1 || fmt == LH_TEXTURE_BC7) {
            const size_t unit = fmt == LH_TEXTURE_BC1 ? 8 : 16;
            std::vector<uint8_t> blocks;
            for (uint32_t l = 0; l < t.levels; l++) {
                lt.level_offset[l] = uint32_t(blocks.size() / unit);
                const uint32_t w = std::max(1u, t.w >> l), h = std::max(1u, t.h >> l);
                const uint8_t* src = t.rgba.data() + size_t(t.level_offset[l]) * 4;
                if (fmt == LH_TEXTURE_BC1)
                    encode_bc1_level(src, w, h, blocks);
                else
                    encode_bc7_level(src, w, h, blocks);
            }
            lt.texels = append_bytes(s->texels, blocks);
        } else {
            lt.texels = append_bytes(s->texels, t.rgba);
            for (uint32_t l = 0; l < t.levels; l++) lt.level_offset[l] = t.level_offset[l];
        }
        s->tex_out.push_back(lt);
    }
    if (s->texels.empty()) s->texels.resize(256);
    s->finished = true;
    return 0;
}

int32_t lh_scene_get(const lh_scene_host* s, lh_scene_arrays* o) {
    if (!s || !s->finished || !o) return -1;
    o->geometry = s->geometry.data();
    o->geometry_bytes = s->geometry.size();
    o->mesh_infos = s->mesh_infos.data();
    o->num_mesh_infos = uint32_t(s->mesh_infos.size());
    o->factors = s->factors.data();
    o->instances = s->inst_out.data();
    o->num_instances = uint32_t(s->inst_out.size());
    o->instance_meshlet_base = s->meshlet_base.data();
    o->meshlet_table = s->meshlet_table.data();
    o->total_meshlets = uint32_t(s->meshlet_table.size());
    o->texels = s->texels.data();
    o->texel_bytes = s->texels.size();
    o->textures = s->tex_out.data();
    o->num_textures = uint32_t(s->tex_out.size());
    o->num_meshes = uint32_t(s->meshes.size());
    o->total_triangles = s->total_tris;
    o->unique_triangles = s->unique_tris;
    return 0;
}

int32_t lh_scene_relocate(const lh_scene_host* s, uint64_t geometry_base, uint64_t mesh_info_base, uint64_t texel_base,
                          MeshInfo* mi_out, Instance* in_out, LhTexture* tex_out) {
    if (!s || !s->finished) return -1;
    if (mi_out)
        for (size_t i = 0; i < s->mesh_infos.size(); i++) {
            MeshInfo m = s->mesh_infos[i];
            m.positions += geometry_base;
            m.indices += geometry_base;
            m.normals += geometry_base;
            m.uvs += geometry_base;
            m.micro_indices += geometry_base;
            m.meshlets += geometry_base;
            mi_out[i] = m;
        }
    if (in_out)
        for (size_t i = 0; i < s->inst_out.size(); i++) {
            Instance in = s->inst_out[i];
            in.mesh_info_address = mesh_info_base + in.mesh_info_address * sizeof(MeshInfo);
            in_out[i] = in;
        }
    if (tex_out)
        for (size_t i = 0; i < s->tex_out.size(); i++) {
            LhTexture t = s->tex_out[i];
            t.texels += texel_base;
            tex_out[i] = t;
        }
    return 0;
}

void lh_make_transform(LhVec3 t, float yaw, float scale, LhMat4* out) {
    float c = std::cos(yaw), s = std::sin(yaw);
    // translate * rotate_y * scale, column major
    out->c[0] = LhVec4{c * scale, 0, -s * scale, 0};
    out->c[1] = LhVec4{0, scale, 0, 0};
    out->c[2] = LhVec4{s * scale, 0, c * scale, 0};
    out->c[3] = LhVec4{t.x, t.y, t.z, 1};
}

int32_t lh_make_uniforms(const lh_camera* cam, uint32_t num_instances, Uniforms* u) {
    if (!cam || !u || cam->width == 0 || cam->height == 0) return -1;
    std::memset(u, 0, sizeof(*u));
    V3 eye{cam->position.x, cam->position.y, cam->position.z};
    V3 at{cam->target.x, cam->target.y, cam->target.z};
    V3 f = normalize(at - eye);
    V3 r = normalize(cross(f, V3{0, 1, 0}));
    V3 up = cross(r, f);
    LhMat4 view;
    view.c[0] = LhVec4{r.x, up.x, -f.x, 0};
    view.c[1] = LhVec4{r.y, up.y, -f.y, 0};
    view.c[2] = LhVec4{r.z, up.z, -f.z, 0};
    view.c[3] = LhVec4{-dot(r, eye), -dot(up, eye), dot(f, eye), 1};
    // src/projection.h:3-20
    float focal = 1.0f / std::tan(cam->fov_y_rad / 2.0f);
    float aspect = float(cam->width) / float(cam->height);
    LhMat4 proj;
    proj.c[0] = LhVec4{focal / aspect, 0, 0, 0};
    proj.c[1] = LhVec4{0, -focal, 0, 0};
    proj.c[2] = LhVec4{0, 0, 0, -1};
    proj.c[3] = LhVec4{0, 0, cam->near_plane, 0};
    u->view = view;
    u->initial_view = view;
    u->perspective = proj;
    u->combined_perspective_view = mul4(proj, view);
    u->inv_perspective_view = inverse4(u->combined_perspective_view);
    u->perspective_inverse = inverse4(proj);
    u->view_inverse = inverse4(view);
    u->camera_pos = cam->position;
    u->sun_dir = cam->sun_dir;
    u->sun_intensity = cam->sun_intensity;
    u->num_instances = num_instances;
    u->window_size = LhUVec2{cam->width, cam->height};
    u->shadow_cam_distance = cam->shadow_cam_distance;
    u->cascade_split_pow = cam->cascade_split_pow;
    u->debug = UNIFORMS_DEBUG_OFF;
    u->debug_shadowmaps = 0;
    return 0;
}

// ----------------------------------------------------------------- presets

static LhVec3 default_sun() {
    V3 s = normalize(V3{0.35f, 0.85f, 0.40f});
    return LhVec3{s.x, s.y, s.z};
}

int32_t lh_preset_camera(uint32_t preset, lh_camera* c) {
    if (!c) return -1;
    std::memset(c, 0, sizeof(*c));
    c->fov_y_rad = 60.0f * 3.14159265358979f / 180.0f;
    c->sun_dir = default_sun();
    c->sun_intensity = LhVec3{1, 1, 1};
    c->shadow_cam_distance = 1024.0f;  // src/main.cpp:653
    c->cascade_split_pow = 3.0f;       // src/main.cpp:654
    switch (preset) {
        case LH_PRESET_SMALL:
            c->position = LhVec3{0, 20, 80};
            c->target = LhVec3{0, 0, 0};
            c->near_plane = 0.1f;
            c->width = 1280;
            c->height = 720;
            return 0;
        case LH_PRESET_LARGE:
        case LH_PRESET_STRESS:
            // Over the clustered field, 40 m up, looking down ~25 degrees: ~9.95M
            // triangles of instances survive the main-view instance cull for the
            // 2,000-instance scene (~10x that for 20,000 instances).
            c->position = LhVec3{150, 40, 150};
            c->target = LhVec3{79.3f, -5, 220.7f};
            c->near_plane = 0.1f;
            c->width = preset == LH_PRESET_LARGE ? 1920 : 3840;
            c->height = preset == LH_PRESET_LARGE ? 1080 : 2160;
            return 0;
        default:
            return -1;
    }
}

int32_t lh_batch_cameras(uint64_t seed, uint32_t count, uint32_t min_side, uint32_t max_side, lh_camera* out) {
    if (!out || min_side < 8 || max_side < min_side) return -1;
    Rng rng(seed * 7919u + 17u);
    lh_camera base;
    lh_preset_camera(LH_PRESET_LARGE, &base);
    for (uint32_t i = 0; i < count; i++) {
        lh_camera c = base;
        double lo = std::log(double(min_side)), hi = std::log(double(max_side));
        uint32_t side = uint32_t(std::exp(rng.uniform(lo, hi)));
        side = std::clamp((side + 4) / 8 * 8, (min_side + 7) / 8 * 8, max_side / 8 * 8);
        double ang = rng.uniform(0, 6.283185307179586), dist = rng.uniform(150, 380);
        c.position = LhVec3{float(dist * std::cos(ang)), float(rng.uniform(10, 80)), float(dist * std::sin(ang))};
        c.target = LhVec3{float(rng.uniform(-60, 60)), 0.0f, float(rng.uniform(-60, 60))};
        c.width = c.height = side;
        out[i] = c;
    }
    return 0;
}

lh_scene_host* lh_scene_build_preset(const lh_preset_config* cfg) {
    if (!cfg) return nullptr;
    uint64_t seed = cfg->seed ? cfg->seed : cfg->preset;
    Rng rng(seed);
    auto s = std::make_unique<lh_scene_host>();
    if (lh_scene_set_vertex_format(s.get(), cfg->vertex_format) != 0) return nullptr;
    if (lh_scene_set_texture_format(s.get(), cfg->texture_format) != 0) return nullptr;
    if (cfg->cache_dir && lh_scene_set_cache_dir(s.get(), cfg->cache_dir) != 0) return nullptr;
    const float tri_scale = (cfg->mesh_tri_scale_pct ? cfg->mesh_tri_scale_pct : 100) / 100.0f;
    if (cfg->preset == LH_PRESET_SMALL) {
        // One sphere of exactly 8,192 triangles (octahedron, level 5), 4 constant
        // materials, 64 instances uniformly in a 100 m cube.
        if (add_mesh(s.get(), octa_sphere(5)) < 0) return nullptr;
        for (int m = 0; m < 4; m++) {
            LhVec3 col{float(rng.uniform(0.1, 0.9)), float(rng.uniform(0.1, 0.9)), float(rng.uniform(0.1, 0.9))};
            float rough = float(rng.uniform(0.2, 0.9)), metal = (m & 1) ? 1.0f : 0.0f;
            lh_scene_add_material(s.get(), col, rough, metal, UNUSED_TEXTURE_INDEX, UNUSED_TEXTURE_INDEX,
                                  UNUSED_TEXTURE_INDEX, LhVec2{1, 1}, LhVec2{0, 0});
        }
        uint32_t n = cfg->num_instances ? cfg->num_instances : 64;
        for (uint32_t i = 0; i < n; i++) {
            LhVec3 p{float(rng.uniform(-50, 50)), float(rng.uniform(-50, 50)), float(rng.uniform(-50, 50))};
            float yaw = float(rng.uniform(0, 6.283185307179586));
            float sc = float(rng.uniform(0.5, 2.0));
            LhMat4 t;
            lh_make_transform(p, yaw, sc, &t);
            lh_scene_add_instance(s.get(), 0, i % 4, &t);
        }
    } else if (cfg->preset == LH_PRESET_LARGE || cfg->preset == LH_PRESET_STRESS) {
        // 50 meshes with triangle counts log-uniform in [5k, 500k]: displaced
        // spheres, boxes and planar grids.
        const int kMeshes = 50;
        std::vector<double> targets(kMeshes);
        for (int m = 0; m < kMeshes; m++) targets[m] = 5000.0 * std::pow(100.0, (m + rng.uniform()) / kMeshes);
        for (int m = kMeshes - 1; m > 0; m--) std::swap(targets[m], targets[rng.below(m + 1)]);
        std::vector<int> kind(kMeshes);
        std::vector<double> tri_count(kMeshes);
        for (int m = 0; m < kMeshes; m++) {
            double t = targets[m] * tri_scale;
            kind[m] = m % 3;
            MeshSource src;
            uint32_t mseed = uint32_t(seed * 1000 + m);
            if (kind[m] == 0) {
                uint32_t n = std::max(2u, uint32_t(std::lround(std::sqrt(t / 12.0))));
                src = displaced_sphere(n, float(rng.uniform(0.05, 0.25)), float(rng.uniform(1.5, 4.0)), mseed);
            } else if (kind[m] == 1) {
                uint32_t n = std::max(1u, uint32_t(std::lround(std::sqrt(t / 12.0))));
                src = box_mesh(n);
            } else {
                uint32_t n = std::max(2u, uint32_t(std::lround(std::sqrt(t / 2.0))));
                src = grid_mesh(n, n, float(rng.uniform(0.02, 0.15)), float(rng.uniform(2.0, 6.0)), mseed);
            }
            tri_count[m] = double(src.nt());
            if (add_mesh(s.get(), std::move(src)) < 0) return nullptr;
        }
        // 64 materials, one 1024^2 value-noise base-colour texture each.
        uint32_t tsize = cfg->texture_size ? cfg->texture_size : 1024;
        const int kMaterials = 64;
        std::vector<TextureData> texs(kMaterials);
        {
            unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (unsigned w = 0; w < nth; w++)
                pool.emplace_back([&, w] {
                    for (int k = int(w); k < kMaterials; k += int(nth)) {
                        value_noise_texture(texs[k], tsize, uint32_t(seed * 131 + k));
                        build_mips(texs[k]);
                    }
                });
            for (auto& t : pool) t.join();
        }
        for (int k = 0; k < kMaterials; k++) {
            s->textures.push_back(std::move(texs[k]));
            LhVec3 col{float(rng.uniform(0.6, 1.0)), float(rng.uniform(0.6, 1.0)), float(rng.uniform(0.6, 1.0))};
            float rough = float(rng.uniform(0.25, 0.95)), metal = rng.uniform() < 0.25 ? 1.0f : 0.0f;
            float ts = float(rng.uniform(1.0, 4.0));
            lh_scene_add_material(s.get(), col, rough, metal, uint32_t(k), UNUSED_TEXTURE_INDEX, UNUSED_TEXTURE_INDEX,
                                  LhVec2{ts, ts}, LhVec2{float(rng.uniform()), float(rng.uniform())});
        }
        // Instances: mesh choice weighted by 1/triangles (each mesh contributes
        // about the same number of triangles), lognormal-clustered positions.
        uint32_t n = cfg->num_instances ? cfg->num_instances : (cfg->preset == LH_PRESET_LARGE ? 2000 : 20000);
        std::vector<double> cdf(kMeshes);
        double acc = 0;
        for (int m = 0; m < kMeshes; m++) cdf[m] = (acc += 1.0 / tri_count[m]);
        const int kClusters = 24;
        std::vector<V3> centres(kClusters);
        std::vector<double> cw(kClusters), spread(kClusters);
        double cacc = 0;
        for (int k = 0; k < kClusters; k++) {
            centres[k] = V3{float(rng.uniform(-300, 300)), 0.0f, float(rng.uniform(-300, 300))};
            cw[k] = (cacc += std::exp(rng.normal()));
            spread[k] = std::exp(std::log(25.0) + 0.6 * rng.normal());
        }
        for (uint32_t i = 0; i < n; i++) {
            double u = rng.uniform() * acc;
            int m = int(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            m = std::min(m, kMeshes - 1);
            double cu = rng.uniform() * cacc;
            int k = std::min(int(std::lower_bound(cw.begin(), cw.end(), cu) - cw.begin()), kClusters - 1);
            // Planar grids are large terrain tiles (they overlap into a ground);
            // spheres and boxes are objects of a few metres.
            float sc = kind[m] == 2 ? float(std::exp(std::log(28.0) + 0.3 * rng.normal()))
                                    : float(std::exp(std::log(2.0) + 0.5 * rng.normal()));
            V3 p = centres[k] + V3{float(rng.normal() * spread[k]), 0.0f, float(rng.normal() * spread[k])};
            p.y = kind[m] == 2 ? float(rng.uniform(-1.0, 1.0)) : sc + float(std::fabs(rng.normal()) * 4.0);
            float yaw = float(rng.uniform(0, 6.283185307179586));
            LhMat4 t;
            lh_make_transform(LhVec3{p.x, p.y, p.z}, yaw, sc, &t);
            lh_scene_add_instance(s.get(), uint32_t(m), rng.below(kMaterials), &t);
        }
    } else {
        return nullptr;
    }
    if (lh_scene_finish(s.get()) != 0) return nullptr;
    return s.release();
}

}  // extern "C"
