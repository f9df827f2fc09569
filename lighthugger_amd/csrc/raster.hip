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
//   k_raster_big       queued triangles, one wave each (whole block for
//                      boxes over kWaveBoxPx pixels).
#include "raster.h"

using namespace lhd;
using namespace lhr;

namespace {

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
    auto body = [&](uint32_t view, uint32_t pi, uint32_t d, bool clip_queued) {
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
            // clipped and invalid vertices store zeros, so that the triangle
            // setup below may read any corner
            SV sv{0, 0, 0.0f};
            if (oc == 0 && !project(c, p.W, p.H, sv)) oc = kBadVertex;
            sX[lane] = sv.X;
            sY[lane] = sv.Y;
            sZ[lane] = sv.z;
            sO[lane] = oc;
        }
        __syncthreads();
        for (uint32_t round = 0; round < 2; round++) {
            const uint32_t t = lane + 64 * round;
            if (__ballot(t < ntri) == 0) break;
            // Every lane sets up its triangle (lanes past the meshlet read
            // corner 0 three times: zero area) and the class is computed
            // without nested branches, whose exits would each re-zero the
            // setup's registers.
            const uint32_t packed = corners[round];
            const uint32_t i0 = packed & 255u, i1 = (packed >> 8) & 255u, i2 = packed >> 16;
            const uint32_t o0 = sO[i0], o1 = sO[i1], o2 = sO[i2];
            SV v0{sX[i0], sY[i0], sZ[i0]}, v1{sX[i1], sY[i1], sZ[i1]}, v2{sX[i2], sY[i2], sZ[i2]};
            int64_t area;
            int32_t xlo, xhi, ylo, yhi, ex, ey;
            const bool set = setup<KIND>(v0, v1, v2, p.W, p.H, area, xlo, xhi, ylo, yhi, ex, ey, ALPHA);
            const bool drawn = t < ntri && ((o0 | o1 | o2) & kBadVertex) == 0 && (o0 & o1 & o2) == 0;
            const int32_t e = max(ex, ey);
            // 0 none, 1 clip queue (unless the tiled raster's binning queued
            // it already), 2 big queue, 3 in-lane, 4 wave-cooperative
            const uint32_t cls = !drawn                 ? 0u
                                 : (o0 | o1 | o2) != 0  ? (clip_queued ? 0u : 1u)
                                 : !set                 ? 0u
                                 : e <= ps.small_extent ? 3u
                                 : e <= ps.medium_extent ? 4u
                                                         : 2u;
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
        if (ps.direct) {  // the draws the tiled raster left to this kernel
            const lh_view& v = vs.get(0);
            const lhs::ScratchLayout L = lhs::scratch_layout(v);
            const uint32_t* list = reinterpret_cast<const uint32_t*>(static_cast<const char*>(v.scratch) + L.direct);
            const uint32_t n = min(v.counters[LH_CNT_TILE_DIRECT], L.direct_capacity);
            for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
                const uint32_t e = list[k];
                uint32_t view, pi, d;
                seg.find(e & ~lhs::kBinDirect, view, pi, d);
                body(view, pi, d, (e & lhs::kBinDirect) != 0);
            }
            return;
        }
        for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
            uint32_t view, pi, d;
            seg.find(g, view, pi, d);
            body(view, pi, d, false);
        }
    } else {
        for (uint32_t c0 = blockIdx.x * C; c0 < total; c0 += gridDim.x * C) {
            auto cur = seg.cursor(c0);
            const uint32_t c1 = min(c0 + C, total);
            for (uint32_t g = c0; g < c1; g++) {
                uint32_t view, pi, d;
                seg.at(cur, g, view, pi, d);
                body(view, pi, d, false);
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

// Triangles of the large queue whose pixel box holds at most this many pixels
// are rasterized by one wave alone; larger ones by the whole block.
constexpr int32_t kWaveBoxPx = 4096;

// Rasterizes the pixels of one queued triangle's box that fall on lanes
// first, first + stride, ... of the box's row-major pixel order.  Coverage
// and depth of each pixel are evaluated exactly as in the other paths
// (64-bit edge functions at the pixel centre, top-left rule).
template <int KIND, bool ST, bool ALPHA>
LH_DEV void raster_box(const lh_scene& s, const lh_view& v, const PassSet& ps, const BigTri& b, int32_t first,
                       int32_t stride, Tally<ST>* tally) {
    const uint32_t pi = uint32_t(b.xlo) >> 13;
    const Pass p = make_pass(v, ps.pass0 + pi);
    const int32_t X0 = b.X[0], X1 = b.X[1], X2 = b.X[2], Y0 = b.Y[0], Y1 = b.Y[1], Y2 = b.Y[2];
    const int64_t area = edge64(X0, Y0, X1, Y1, X2, Y2);
    const float inv_area = 1.0f / float(area);
    const float z0 = b.z[0], d10 = b.z[1] - b.z[0], d20 = b.z[2] - b.z[0];
    const bool tl0 = top_left(X1, Y1, X2, Y2), tl1 = top_left(X2, Y2, X0, Y0), tl2 = top_left(X0, Y0, X1, Y1);
    const int32_t xlo = b.xlo & 0x1FFF, xhi = b.xhi, ylo = b.ylo & ~kBigAlpha, yhi = b.yhi;
    const int32_t bw = xhi - xlo + 1, n = bw * (yhi - ylo + 1);
    if (first >= n) return;
    // Step by `stride` pixels in row-major order: sy rows and sx columns,
    // plus one row when the column passes the box's right edge.
    const int32_t sy = stride / bw, sx = stride - sy * bw;
    int32_t py = ylo + first / bw, px = xlo + (first - (first / bw) * bw);
    for (int32_t i = first; i < n; i += stride) {
        const int32_t Px = px * 256 + 128, Py = py * 256 + 128;
        const int64_t e0 = edge64(X1, Y1, X2, Y2, Px, Py);
        const int64_t e1 = edge64(X2, Y2, X0, Y0, Px, Py);
        const int64_t e2 = edge64(X0, Y0, X1, Y1, Px, Py);
        const bool in = (e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) && (e2 > 0 || (e2 == 0 && tl2));
        if (in && (!ALPHA || alpha_pass<KIND>(s, v, p, b.id, px, py))) {
            const float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
            const bool issued = fragment<KIND>(p, px, py, (z0 + l1 * d10) + l2 * d20, b.id);
#pragma unroll
            for (uint32_t q = 0; q < LH_NUM_CASCADES; q++)
                if (q == pi) tally[q].hit(issued);
        }
        px += sx;
        py += sy;
        if (px > xhi) {
            px -= bw;
            py++;
        }
    }
}

// View blockIdx.z.  Each round the block takes four queue entries, one per
// wave.  A wave rasterizes its own triangle when the box is at most
// kWaveBoxPx pixels (most of the queue: slivers just over the medium extent),
// so a 54-pixel box costs one wave pass instead of a 256-thread block; the
// larger boxes of the round are rasterized by all four waves together first.
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
    const uint32_t wave = threadIdx.x >> 6;
    // wide[r & 1][w]: wave w's triangle of round r goes to the whole block.
    // Two buffers, because a wave may write round r + 1's flag while another
    // still reads round r's (no wave gets two rounds ahead: each round has a
    // barrier).
    __shared__ uint32_t wide[2][4];
    // Statistics: each lane counts per pass of the set over all of the
    // block's triangles, and the wave adds them to the view's counters once
    // at the end (a flush per triangle also measured wrong depths on the
    // large frame's sliver triangles with statistics on).
    Tally<ST> tally[LH_NUM_CASCADES];
    uint32_t round = 0;
    for (uint32_t e0 = blockIdx.x * 4; e0 < count; e0 += gridDim.x * 4, round ^= 1) {  // block-uniform
        const uint32_t e = e0 + wave;
        bool own = false, shared = false;
        BigTri b;
        if (e < count) {
            b = qs.big[e];
            if (((b.ylo & kBigAlpha) != 0) == ALPHA) {
                const int32_t bw = b.xhi - (b.xlo & 0x1FFF) + 1, bh = b.yhi - (b.ylo & ~kBigAlpha) + 1;
                own = bw * bh <= kWaveBoxPx;
                shared = !own;
            }
        }
        if (lane_id() == 0) wide[round][wave] = shared ? 1u : 0u;
        __syncthreads();
        for (uint32_t w = 0; w < 4; w++)
            if (wide[round][w]) raster_box<KIND, ST, ALPHA>(s, v, ps, qs.big[e0 + w], threadIdx.x, 256, tally);
        if (own) raster_box<KIND, ST, ALPHA>(s, v, ps, b, int32_t(lane_id()), 64, tally);
    }
#pragma unroll
    for (uint32_t q = 0; q < LH_NUM_CASCADES; q++)
        if (q < ps.n) tally[q].flush(v, ps.pass0 + q);
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

// The per-meshlet raster of a pass set and its queue launches.  A single
// view's main pass with bin_capacity > 0 goes through the tiled raster first
// (tile.hip) when its tuning class asks for it; the per-meshlet kernel then
// covers only the draws it left.
template <int KIND, bool ST, class VS>
hipError_t launch_kernels(const lh_scene& s, const VS& vs, PassSet ps, const RasterTuning& t, uint32_t clip_blocks,
                          uint32_t big_blocks, hipStream_t st) {
    const uint32_t nv = vs.count();
    bool tiled = false;
    if constexpr (std::is_same<VS, OneView>::value)
        tiled = KIND == 0 && vs.v.bin_capacity > 0 && t.tile_raster != 0;
    if (tiled) {
        if constexpr (std::is_same<VS, OneView>::value) {
            hipError_t e = raster_tiles<ST>(s, vs, ps, t, st);
            if (e != hipSuccess) return e;
        }
        ps.direct = 1;
        mark("raster_direct_view", st);
    } else {
        mark(KIND == 0 ? "raster_meshlets_view" : "raster_meshlets_shadow", st);
    }
    hipLaunchKernelGGL((k_raster_meshlets<KIND, VS, ST, false>), dim3(t.meshlet_blocks), dim3(64), 0, st, s, vs, ps);
    ps.direct = 0;
    const bool alpha = !(s.flags & LH_SCENE_OPAQUE_ONLY);
    if (alpha) {
        mark(KIND == 0 ? "raster_alpha_view" : "raster_alpha_shadow", st);
        hipLaunchKernelGGL((k_raster_meshlets<KIND, VS, ST, true>), dim3(t.meshlet_blocks), dim3(64), 0, st, s, vs,
                           ps);
    }
    mark(KIND == 0 ? "raster_clip_view" : "raster_clip_shadow", st);
    hipLaunchKernelGGL((k_raster_clip<KIND, VS>), dim3(clip_blocks, 1, nv), dim3(64), 0, st, s, vs, ps);
    mark(KIND == 0 ? "raster_big_view" : "raster_big_shadow", st);
    hipLaunchKernelGGL((k_raster_big<KIND, VS, ST, false>), dim3(big_blocks, 1, nv), dim3(256), 0, st, s, vs, ps);
    if (alpha) {
        mark(KIND == 0 ? "raster_big_alpha_view" : "raster_big_alpha_shadow", st);
        hipLaunchKernelGGL((k_raster_big<KIND, VS, ST, true>), dim3(big_blocks, 1, nv), dim3(256), 0, st, s, vs, ps);
    }
    return hipGetLastError();
}

// Statistics are a single-view feature (LH_VIEW_STATISTICS).
template <int KIND>
hipError_t launch_set(const lh_scene& s, const OneView& vs, const PassSet& ps, const RasterTuning& t, uint32_t cb,
                      uint32_t bb, hipStream_t st) {
    if (vs.v.flags & LH_VIEW_STATISTICS) return launch_kernels<KIND, true>(s, vs, ps, t, cb, bb, st);
    return launch_kernels<KIND, false>(s, vs, ps, t, cb, bb, st);
}
template <int KIND>
hipError_t launch_set(const lh_scene& s, const ViewTable& vs, const PassSet& ps, const RasterTuning& t, uint32_t cb,
                      uint32_t bb, hipStream_t st) {
    return launch_kernels<KIND, false>(s, vs, ps, t, cb, bb, st);
}

template <class VS>
hipError_t raster(const lh_scene& s, const VS& vs, uint32_t pass0, uint32_t npasses, const RasterTuning& t,
                  hipStream_t st, bool phase2) {
    if (npasses == 0 || npasses > LH_NUM_CASCADES || pass0 + npasses > LH_NUM_PASSES || (pass0 == 0 && npasses != 1))
        return hipErrorInvalidValue;
    // the cascades may take their own extents (their triangles are smaller
    // and their raster depth-only)
    const bool cascades = pass0 > 0;
    const uint32_t small_sel = cascades && t.shadow_small_extent_px ? t.shadow_small_extent_px : t.small_extent_px;
    const uint32_t medium_sel = cascades && t.shadow_medium_extent_px ? t.shadow_medium_extent_px : t.medium_extent_px;
    const uint32_t small_px = std::min<uint32_t>(std::max<uint32_t>(small_sel, 1), 32);
    const uint32_t medium_px = std::min<uint32_t>(std::max(medium_sel, small_px), 64);
    PassSet ps{pass0, npasses, int32_t(small_px * 256), int32_t(medium_px * 256), phase2 ? 1u : 0u, 0u};
    const uint32_t nv = vs.count();
    // per-view persistent grids of the queue kernels shrink with the batch size
    const uint32_t clip_blocks = nv > 1 ? std::max(t.clip_blocks / 16, 4u) : t.clip_blocks;
    const uint32_t big_blocks = nv > 1 ? std::max(t.big_blocks / 16, 32u) : t.big_blocks;
    if (pass0 == 0) return launch_set<0>(s, vs, ps, t, clip_blocks, big_blocks, st);
    return launch_set<1>(s, vs, ps, t, clip_blocks, big_blocks, st);
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
