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
// Work split for one pass (all persistent, grid-stride, stream ordered):
//   k_raster_meshlets  one wave per surviving meshlet: its <= 64 vertices are
//                      transformed once (one per lane) into LDS, then each lane
//                      sets up <= 2 of its <= 124 triangles.  Triangles whose
//                      vertex bbox fits in `small_extent` pixels are rasterized
//                      in the lane with int32 edge functions; larger ones go to
//                      the big queue, clip-crossing ones to the clip queue
//                      (both appended with one atomic per wave).
//   k_raster_clip      clips queued triangles and queues the fan pieces.
//   k_raster_big       one 256-thread block per queued triangle, 16x16 tiles.
#include "lh_device.h"
#include "kernels.h"

using namespace lhd;

namespace {

constexpr float kGuard = 4.0f;
constexpr uint32_t kBadVertex = 64u;  // non-finite, or w <= 0 inside every plane

struct BigTri {
    int32_t X[3];
    int32_t Y[3];
    float z[3];
    uint32_t id;
    uint16_t xlo, ylo, xhi, yhi;
};
static_assert(sizeof(BigTri) == LH_BIG_ENTRY_BYTES, "big queue entry");

struct SV {
    int32_t X, Y;
    float z;
};

struct Pass {
    uint32_t pass, W, H;
    uint64_t* vis;
    uint32_t* depth;
    const MeshletReference* draws;
    uint32_t* clip_queue;
    BigTri* big_queue;
    uint32_t clip_cap, big_cap;
    int32_t small_extent;  // subpixels
};

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

// Culls by facing, orients to positive area and computes the pixel bbox.
template <int KIND>
LH_DEV bool setup(SV& v0, SV& v1, SV& v2, uint32_t W, uint32_t H, int64_t& area, int32_t& xlo, int32_t& xhi,
                  int32_t& ylo, int32_t& yhi, int32_t& extx, int32_t& exty) {
    area = edge64(v0.X, v0.Y, v1.X, v1.Y, v2.X, v2.Y);
    if (KIND == 0 ? area >= 0 : area <= 0) return false;
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

template <int KIND>
LH_DEV void fragment(const Pass& p, int32_t px, int32_t py, float z, uint32_t id) {
    z = fminf(fmaxf(z, 0.0f), 1.0f);
    size_t o = size_t(py) * p.W + uint32_t(px);
    if (KIND == 0) {
        if (!(z > 0.0f)) return;  // depth GREATER against the 0.0 clear
        unsigned long long key = ((unsigned long long)(~__float_as_uint(z)) << 32) | id;
        unsigned long long* a = reinterpret_cast<unsigned long long*>(p.vis + o);
        if (key < *a) atomicMin(a, key);  // a stale value is never smaller: safe early-out
    } else {
        uint32_t d = __float_as_uint(z);
        if (d < p.depth[o]) atomicMin(p.depth + o, d);
    }
}

// In-lane raster of a triangle whose vertex bbox is within small_extent
// (<= 32 px): every edge value fits in int32 and the edge functions are
// stepped exactly in integers.
template <int KIND>
LH_DEV void raster_small(const Pass& p, SV v0, SV v1, SV v2, int32_t area, int32_t xlo, int32_t xhi, int32_t ylo,
                         int32_t yhi, uint32_t id) {
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
            if (in) {
                float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
                fragment<KIND>(p, px, py, (v0.z + l1 * d10) + l2 * d20, id);
            }
            e0 += sx0;
            e1 += sx1;
            e2 += sx2;
        }
    }
}

LH_DEV void store_big(BigTri* q, const SV& v0, const SV& v1, const SV& v2, uint32_t id, int32_t xlo, int32_t xhi,
                      int32_t ylo, int32_t yhi) {
    BigTri b;
    b.X[0] = v0.X, b.X[1] = v1.X, b.X[2] = v2.X;
    b.Y[0] = v0.Y, b.Y[1] = v1.Y, b.Y[2] = v2.Y;
    b.z[0] = v0.z, b.z[1] = v1.z, b.z[2] = v2.z;
    b.id = id;
    b.xlo = uint16_t(xlo), b.ylo = uint16_t(ylo), b.xhi = uint16_t(xhi), b.yhi = uint16_t(yhi);
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

template <int KIND>
LH_DEV const LhMat4& view_matrix(const lh_view& v, uint32_t pass) {
    return KIND == 0 ? v.uniforms->combined_perspective_view : v.misc->shadow_matrices[pass - 1];
}

template <int KIND>
__global__ void __launch_bounds__(64) k_raster_meshlets(lh_scene s, lh_view v, Pass p) {
    __shared__ int32_t sX[64], sY[64];
    __shared__ float sZ[64];
    __shared__ uint32_t sO[64];
    const uint32_t lane = threadIdx.x;
    const LhMat4 M = view_matrix<KIND>(v, p.pass);
    const uint32_t count = min(v.counters[LH_CNT_DRAWS + p.pass], v.draw_capacity);
    uint32_t* clip_counter = &v.counters[LH_CNT_CLIP + p.pass];
    uint32_t* big_counter = &v.counters[LH_CNT_BIG + p.pass];
    for (uint32_t d = blockIdx.x; d < count; d += gridDim.x) {
        const MeshletReference ref = p.draws[d];
        const Instance& in = s.instances[ref.instance_index];
        const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
        const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
        const uint32_t mid = s.instance_meshlet_base[ref.instance_index] + ref.meshlet_index;
        const uint32_t nvert = ml.index_count, ntri = ml.triangle_count;
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
        const uint8_t* micro = ptr<uint8_t>(mi.micro_indices) + ml.triangle_offset;
        for (uint32_t round = 0; round < 2; round++) {
            const uint32_t t = lane + 64 * round;
            // 0 none, 1 clip queue, 2 big queue, 3 in-lane
            uint32_t cls = 0;
            SV v0{}, v1{}, v2{};
            int64_t area = 0;
            int32_t xlo = 0, xhi = 0, ylo = 0, yhi = 0, ex = 0, ey = 0;
            if (t < ntri) {
                uint32_t i0 = micro[3 * t], i1 = micro[3 * t + 1], i2 = micro[3 * t + 2];
                uint32_t o0 = sO[i0], o1 = sO[i1], o2 = sO[i2];
                if (((o0 | o1 | o2) & kBadVertex) == 0 && (o0 & o1 & o2) == 0) {
                    if (o0 | o1 | o2) {
                        cls = 1;
                    } else {
                        v0 = SV{sX[i0], sY[i0], sZ[i0]};
                        v1 = SV{sX[i1], sY[i1], sZ[i1]};
                        v2 = SV{sX[i2], sY[i2], sZ[i2]};
                        if (setup<KIND>(v0, v1, v2, p.W, p.H, area, xlo, xhi, ylo, yhi, ex, ey))
                            cls = (ex <= p.small_extent && ey <= p.small_extent) ? 3u : 2u;
                    }
                }
            }
            if (__ballot(t < ntri) == 0) break;
            const uint32_t id = (mid << LH_KEY_TRI_BITS) | t;
            uint32_t cslot = wave_append(clip_counter, cls == 1);
            if (cls == 1) {
                if (cslot < p.clip_cap)
                    p.clip_queue[cslot] = (d << LH_KEY_TRI_BITS) | t;
                else
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_CLIP);
            }
            uint32_t bslot = wave_append(big_counter, cls == 2);
            if (cls == 2) {
                if (bslot < p.big_cap)
                    store_big(p.big_queue + bslot, v0, v1, v2, id, xlo, xhi, ylo, yhi);
                else
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_BIG);
            }
            if (cls == 3) raster_small<KIND>(p, v0, v1, v2, int32_t(area), xlo, xhi, ylo, yhi, id);
        }
        __syncthreads();
    }
}

// Clip polygon in homogeneous coordinates (Sutherland-Hodgman).
LH_DEV f4 clip_point(f4 a, f4 b, float da, float db) {
    float t = da / (da - db);
    return {a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t};
}

template <int KIND>
__global__ void __launch_bounds__(64) k_raster_clip(lh_scene s, lh_view v, Pass p) {
    const LhMat4 M = view_matrix<KIND>(v, p.pass);
    const uint32_t count = min(v.counters[LH_CNT_CLIP + p.pass], p.clip_cap);
    uint32_t* big_counter = &v.counters[LH_CNT_BIG + p.pass];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        const uint32_t e = p.clip_queue[q];
        const uint32_t d = e >> LH_KEY_TRI_BITS, t = e & ((1u << LH_KEY_TRI_BITS) - 1);
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
            if (!setup<KIND>(a, b, c, p.W, p.H, area, xlo, xhi, ylo, yhi, ex, ey)) continue;
            uint32_t slot = atomicAdd(big_counter, 1u);
            if (slot < p.big_cap)
                store_big(p.big_queue + slot, a, b, c, id, xlo, xhi, ylo, yhi);
            else
                atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_BIG);
        }
    }
}

// One 256-thread block per queued triangle; the bbox is walked in 16x16 tiles
// and each pixel evaluates the int64 edge functions directly.
template <int KIND>
__global__ void __launch_bounds__(256) k_raster_big(lh_view v, Pass p) {
    const uint32_t count = min(v.counters[LH_CNT_BIG + p.pass], p.big_cap);
    const int32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const BigTri b = p.big_queue[e];
        const int32_t X0 = b.X[0], X1 = b.X[1], X2 = b.X[2], Y0 = b.Y[0], Y1 = b.Y[1], Y2 = b.Y[2];
        const int64_t area = edge64(X0, Y0, X1, Y1, X2, Y2);
        const float inv_area = 1.0f / float(area);
        const float z0 = b.z[0], d10 = b.z[1] - b.z[0], d20 = b.z[2] - b.z[0];
        const bool tl0 = top_left(X1, Y1, X2, Y2), tl1 = top_left(X2, Y2, X0, Y0), tl2 = top_left(X0, Y0, X1, Y1);
        const int32_t xlo = b.xlo, xhi = b.xhi, ylo = b.ylo, yhi = b.yhi;
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
                const float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
                fragment<KIND>(p, px, py, (z0 + l1 * d10) + l2 * d20, b.id);
            }
        }
    }
}

}  // namespace

namespace lhk {

hipError_t raster(const lh_scene& s, const lh_view& v, uint32_t pass, const RasterTuning& t, hipStream_t st) {
    Pass p{};
    p.pass = pass;
    if (pass == 0) {
        p.W = v.width;
        p.H = v.height;
        p.vis = v.visbuffer;
    } else {
        p.W = p.H = v.shadow_size;
        p.depth = v.shadow_maps + size_t(pass - 1) * v.shadow_size * v.shadow_size;
    }
    p.draws = v.draws + size_t(pass) * v.draw_capacity;
    p.clip_queue = static_cast<uint32_t*>(v.scratch);
    uint64_t clip_bytes = (uint64_t(v.clip_capacity) * LH_CLIP_ENTRY_BYTES + 255) & ~uint64_t(255);
    p.big_queue = reinterpret_cast<BigTri*>(static_cast<char*>(v.scratch) + clip_bytes);
    p.clip_cap = v.clip_capacity;
    p.big_cap = v.big_capacity;
    p.small_extent = int32_t(std::min<uint32_t>(std::max<uint32_t>(t.small_extent_px, 1), 32) * 256);
    if (pass == 0) {
        hipLaunchKernelGGL(k_raster_meshlets<0>, dim3(t.meshlet_blocks), dim3(64), 0, st, s, v, p);
        hipLaunchKernelGGL(k_raster_clip<0>, dim3(t.clip_blocks), dim3(64), 0, st, s, v, p);
        hipLaunchKernelGGL(k_raster_big<0>, dim3(t.big_blocks), dim3(256), 0, st, v, p);
    } else {
        hipLaunchKernelGGL(k_raster_meshlets<1>, dim3(t.meshlet_blocks), dim3(64), 0, st, s, v, p);
        hipLaunchKernelGGL(k_raster_clip<1>, dim3(t.clip_blocks), dim3(64), 0, st, s, v, p);
        hipLaunchKernelGGL(k_raster_big<1>, dim3(t.big_blocks), dim3(256), 0, st, v, p);
    }
    return hipGetLastError();
}

}  // namespace lhk
