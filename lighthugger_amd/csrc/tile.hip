// Tiled visibility-buffer raster of the main view (LDS-binned 64x64 screen
// tiles).  Same coverage and depth rules as the per-meshlet raster
// (raster.h; the reference's visbuffer_opaque_vertex/pixel,
// src/shaders/rasterization.glsl:38-58, with the pipeline state of
// src/pipelines.cpp:10-77), so the keys are identical; what changes is where
// the depth test happens: in LDS, one 64x64 tile of 64-bit keys per work item,
// and each pixel of a tile reaches the visibility buffer once per item (a
// plain store when the item is the tile's only one), instead of one
// device-scope atomicMin per covered sample.
//
// Launches of one main-view raster (single view):
//   k_bin_meshlets  one wave per surviving meshlet (the first bin_capacity):
//                   vertices fetched, transformed and snapped once; the snapped
//                   vertices and the front-facing, unclipped triangles (oriented,
//                   compacted) go to the bin cache, the tile rectangle of their
//                   pixel bounding boxes to the bin header; clip-crossing
//                   triangles to the clip queue and triangles over the medium
//                   extent to the large-triangle queue, as in the per-meshlet
//                   raster.
//   k_bin_count     one thread per draw: reserves its (meshlet, tile) entries,
//                   counts them per tile; draws that cannot be binned (over
//                   bin_capacity, too many tiles, no entry left) go to the
//                   direct list for the per-meshlet raster.
//   k_bin_scan      one block: tile offsets and work items of at most
//                   `item_entries` entries per tile.
//   k_bin_fill      one thread per draw: writes its entries.
//   k_tile_raster   persistent blocks of 8 waves over the items: LDS tile
//                   cleared, each wave rasterizes one binned meshlet at a time
//                   (two when their triangles fit one round) from the cache
//                   (ds_min_u64 per covered sample), then the tile's covered
//                   pixels are written out.
// The per-meshlet raster then runs over the direct list, followed by the
// alpha-clip, clip and large-triangle launches (all device-scope atomicMin on
// top of what the tiles wrote).
#include "raster.h"

using namespace lhd;
using namespace lhr;
using lhs::BinCache;
using lhs::BinHeader;
using lhs::ScratchLayout;
using lhs::TileItem;

namespace {

constexpr uint32_t kTile = LH_TILE_SIDE;
constexpr uint32_t kTileWaves = 8;
constexpr uint32_t kTileThreads = 64 * kTileWaves;
constexpr uint32_t kScanThreads = 1024;

struct Bins {
    BinHeader* hdr;
    BinCache* cache;
    uint32_t* tile_count;
    uint32_t* tile_cursor;
    TileItem* items;
    uint32_t* entries;
    uint32_t* direct;
    ScratchLayout L;
};
LH_DEV Bins bins(const lh_view& v) {
    const ScratchLayout L = lhs::scratch_layout(v);
    char* b = static_cast<char*>(v.scratch);
    return {reinterpret_cast<BinHeader*>(b + L.hdr),
            reinterpret_cast<BinCache*>(b + L.cache),
            reinterpret_cast<uint32_t*>(b + L.tile_count),
            reinterpret_cast<uint32_t*>(b + L.tile_cursor),
            reinterpret_cast<TileItem*>(b + L.items),
            reinterpret_cast<uint32_t*>(b + L.entries),
            reinterpret_cast<uint32_t*>(b + L.direct),
            L};
}

LH_DEV uint32_t rect_tiles(uint32_t r) {
    return (((r >> 16) & 255u) - (r & 255u) + 1u) * ((r >> 24) - ((r >> 8) & 255u) + 1u);
}

// Draws [0, n) of the launch binned by k_bin_meshlets.
LH_DEV uint32_t binned_draws(const Segments<OneView>& seg, const lh_view& v) { return min(seg.total(), v.bin_capacity); }

// ---------------------------------------------------------------- binning

// One wave per draw g < min(total, bin_capacity).  Alpha-clip meshlets are
// left to the alpha-clip raster (header with no triangles).
template <int KIND>
__global__ void __launch_bounds__(64) LH_RASTER_OCCUPANCY k_bin_meshlets(lh_scene s, OneView vs, PassSet ps) {
    __shared__ int32_t sX[64], sY[64];
    __shared__ float sZ[64];
    __shared__ uint32_t sO[64];
    const uint32_t lane = threadIdx.x;
    const lh_view& v = vs.v;
    const Segments<OneView> seg(vs, ps);
    const uint32_t total = binned_draws(seg, v);
    const Bins B = bins(v);
    const Queues qs = queues(v);
    uint32_t* clip_counter = &v.counters[LH_CNT_CLIP + ps.pass0];
    uint32_t* big_counter = &v.counters[LH_CNT_BIG + ps.pass0];
    // this launch's tile counts and reservations start from zero (k_bin_count runs after it)
    for (uint32_t t = blockIdx.x * 64 + lane; t < B.L.tiles; t += gridDim.x * 64) B.tile_count[t] = 0;
    if (blockIdx.x == 0 && lane == 0) {
        v.counters[LH_CNT_TILE_DIRECT] = 0;
        v.counters[LH_CNT_TILE_ITEMS] = 0;
    }
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        uint32_t view, pi, d;
        seg.find(g, view, pi, d);
        const Pass p = make_pass(v, ps.pass0 + pi);
        const LhMat4& M = view_matrix<KIND>(v, p.pass);
        const MeshletReference ref = p.draws[d];
        const Instance& in = s.instances[ref.instance_index];
        const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
        const uint32_t mid = s.instance_meshlet_base[ref.instance_index] + ref.meshlet_index;
        if (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) {  // wave-uniform
            if (lane == 0) B.hdr[g] = BinHeader{0u, 0u, mid, 0u};
            continue;
        }
        const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
        const uint32_t nvert = ml.index_count, ntri = ml.triangle_count;
        const uint8_t* micro = ptr<uint8_t>(mi.micro_indices) + ml.triangle_offset;
        uint32_t corners[2] = {0u, 0u};
        for (uint32_t r = 0; r < 2; r++) {
            const uint32_t t = lane + 64 * r;
            if (t < ntri)
                corners[r] = uint32_t(micro[3 * t]) | (uint32_t(micro[3 * t + 1]) << 8) |
                             (uint32_t(micro[3 * t + 2]) << 16);
        }
        BinCache& c = B.cache[g];
        if (lane < nvert) {
            f3 w = mat4_point(in.transform, load_position(mi, load_index(mi, ml.index_offset + lane)));
            f4 q = mat4_mul_vec4(M, f4{w.x, w.y, w.z, 1.0f});
            uint32_t oc = finite4(q) ? outcode(q) : kBadVertex;
            SV sv{0, 0, 0.0f};
            if (oc == 0 && !project(q, p.W, p.H, sv)) oc = kBadVertex;
            sX[lane] = sv.X;
            sY[lane] = sv.Y;
            sZ[lane] = sv.z;
            sO[lane] = oc;
            c.X[lane] = sv.X;
            c.Y[lane] = sv.Y;
            c.Z[lane] = sv.z;
        }
        __syncthreads();
        int32_t xlo_m = 0x7FFFFFFF, ylo_m = 0x7FFFFFFF, xhi_m = -1, yhi_m = -1;
        uint32_t nfront = 0, work = 0;  // work: pixels of the front triangles' boxes
        for (uint32_t round = 0; round < 2; round++) {
            const uint32_t t = lane + 64 * round;
            if (__ballot(t < ntri) == 0) break;
            // Every lane sets up its triangle (lanes past the meshlet read
            // corner 0 three times: zero area; clipped corners read as zero)
            // and the outcome is flat predicates, not nested branches whose
            // exits would each re-zero the setup's registers.
            uint32_t packed = corners[round];
            const uint32_t i0 = packed & 255u, i1 = (packed >> 8) & 255u, i2 = packed >> 16;
            const uint32_t o0 = sO[i0], o1 = sO[i1], o2 = sO[i2];
            SV v0{sX[i0], sY[i0], sZ[i0]}, v1{sX[i1], sY[i1], sZ[i1]}, v2{sX[i2], sY[i2], sZ[i2]};
            int64_t area;
            int32_t xlo, xhi, ylo, yhi, ex, ey;
            bool swapped;
            const bool set = setup<KIND>(v0, v1, v2, p.W, p.H, area, xlo, xhi, ylo, yhi, ex, ey, false, &swapped);
            const bool drawn = t < ntri && ((o0 | o1 | o2) & kBadVertex) == 0 && (o0 & o1 & o2) == 0;
            const bool clip = drawn && (o0 | o1 | o2) != 0;
            const bool inside = drawn && (o0 | o1 | o2) == 0 && set;
            // triangles over the medium extent: spread over many blocks by the large-triangle raster
            const bool big = inside && max(ex, ey) > ps.medium_extent;
            const bool front = inside && !big;
            if (front) {
                if (swapped) packed = i0 | (i2 << 8) | (i1 << 16);
                xlo_m = min(xlo_m, xlo);
                ylo_m = min(ylo_m, ylo);
                xhi_m = max(xhi_m, xhi);
                yhi_m = max(yhi_m, yhi);
                work += uint32_t((xhi - xlo + 1) * (yhi - ylo + 1));
            }
            const uint32_t cslot = wave_append(clip_counter, clip);
            if (clip) {
                if (cslot < v.clip_capacity)
                    qs.clip[cslot] = make_uint2((d << LH_KEY_TRI_BITS) | t, pi);
                else
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_CLIP);
            }
            const uint32_t bslot = wave_append(big_counter, big);
            if (big) {
                if (bslot < v.big_capacity)
                    store_big(qs.big + bslot, v0, v1, v2, (mid << LH_KEY_TRI_BITS) | t, xlo, xhi, ylo, yhi, pi, false);
                else
                    atomicOr(&v.counters[LH_CNT_OVERFLOW], LH_OVERFLOW_BIG);
            }
            const uint64_t fm = ballot(front);
            if (front) c.tri[nfront + lanes_below(fm)] = packed | (t << 24);
            nfront += uint32_t(__popcll(fm));
        }
        const int32_t x0 = int32_t(wave_min_u32(uint32_t(xlo_m))), y0 = int32_t(wave_min_u32(uint32_t(ylo_m)));
        const int32_t x1 = int32_t(wave_max_u32(uint32_t(xhi_m + 1))) - 1,
                      y1 = int32_t(wave_max_u32(uint32_t(yhi_m + 1))) - 1;
        const uint32_t total_work = wave_sum_u32(work);
        if (lane == 0) {
            BinHeader h{0u, 0u, mid, 0u};
            if (nfront) {
                h.rect = (uint32_t(x0) >> lhs::kTileShift) | ((uint32_t(y0) >> lhs::kTileShift) << 8) |
                         ((uint32_t(x1) >> lhs::kTileShift) << 16) | ((uint32_t(y1) >> lhs::kTileShift) << 24);
                // a meshlet with much pixel work is split by triangle range into 2^lp parts, each its own
                // entry (so its work spreads over several waves), within LH_TILE_MAX_RECT entries in all
                const uint32_t tiles = rect_tiles(h.rect);
                uint32_t lp = 0;
                while (lp < lhs::kMaxPartsLog2 && (total_work >> (lhs::kPartWorkLog2 + lp)) != 0 &&
                       (tiles << (lp + 1)) <= LH_TILE_MAX_RECT && (nfront >> (lp + 1)) >= 8)
                    lp++;
                h.ntri = nfront | (lp << lhs::kPartShift);
            }
            B.hdr[g] = h;
        }
        __syncthreads();
    }
}

// Per-tile counters are aggregated in LDS per block before they reach the
// device (same-address device atomics serialize): a block handles runs of
// kBinChunk consecutive draws and tiles up to kHistTiles; views with more
// tiles count in device memory directly.  One draw per thread and chunk: the
// grid is sized for the draw capacity, and longer chunks left most blocks
// without a chunk when far fewer draws survive (the large frame: 66k draws,
// 33 busy blocks at 2,048 draws a chunk; the fill pass took 0.024 ms, 0.011
// at 256).
constexpr uint32_t kBinThreads = 256;
constexpr uint32_t kBinChunk = kBinThreads;
constexpr uint32_t kHistTiles = 16384;  // 128 x 128 tiles: an 8192 x 8192 view

LH_DEV void for_each_tile(uint32_t rect, uint32_t tiles_x, auto&& f) {
    const uint32_t tx0 = rect & 255u, ty0 = (rect >> 8) & 255u, tx1 = (rect >> 16) & 255u, ty1 = rect >> 24;
    for (uint32_t ty = ty0; ty <= ty1; ty++)
        for (uint32_t tx = tx0; tx <= tx1; tx++) f(ty * tiles_x + tx);
}

// One thread per draw of the launch.  Draws past bin_capacity and those whose
// rectangle exceeds LH_TILE_MAX_RECT tiles go to the direct list (flattened
// index, kBinDirect when k_bin_meshlets queued their clip triangles); the
// entry list holds LH_TILE_MAX_RECT entries per binned draw, so every other
// draw finds room.
__global__ void __launch_bounds__(kBinThreads) k_bin_count(OneView vs, PassSet ps) {
    __shared__ uint32_t hist[kHistTiles];
    const lh_view& v = vs.v;
    const Segments<OneView> seg(vs, ps);
    const uint32_t total = seg.total();
    const uint32_t nb = binned_draws(seg, v);
    const Bins B = bins(v);
    const bool lds = B.L.tiles <= kHistTiles;
    if (lds)
        for (uint32_t t = threadIdx.x; t < B.L.tiles; t += kBinThreads) hist[t] = 0;
    __syncthreads();
    for (uint32_t base = blockIdx.x * kBinThreads; base < total; base += gridDim.x * kBinThreads) {
        const uint32_t g = base + threadIdx.x;
        bool direct = false;
        uint32_t rect = 0, n = 0, tag = 0, parts = 1;
        if (g < nb) {
            const BinHeader h = B.hdr[g];
            if (h.ntri) {
                rect = h.rect;
                n = rect_tiles(rect);
                parts = 1u << lhs::parts_log2(h.ntri);
                if (n > LH_TILE_MAX_RECT) {
                    direct = true;
                    n = 0;
                }
                tag = lhs::kBinDirect;
            }
        } else if (g < total) {
            direct = true;
        }
        if (direct && g < nb) B.hdr[g].ntri |= lhs::kBinDirect;
        const uint32_t slot = wave_append(&v.counters[LH_CNT_TILE_DIRECT], direct);
        if (direct && slot < B.L.direct_capacity) B.direct[slot] = g | tag;
        if (n) {
            if (lds)
                for_each_tile(rect, B.L.tiles_x, [&](uint32_t t) { atomicAdd(&hist[t], parts); });
            else
                for_each_tile(rect, B.L.tiles_x, [&](uint32_t t) { atomicAdd(&B.tile_count[t], parts); });
        }
    }
    __syncthreads();
    if (lds)
        for (uint32_t t = threadIdx.x; t < B.L.tiles; t += kBinThreads)
            if (hist[t]) atomicAdd(&B.tile_count[t], hist[t]);
}

// One block: exclusive scan of the tile counts into entry offsets (the fill
// cursors) and the item list (runs of at most `item_entries` entries per tile).
__global__ void __launch_bounds__(kScanThreads) k_bin_scan(OneView vs, uint32_t item_entries) {
    __shared__ uint32_t wsum[2][kScanThreads / 64];
    __shared__ uint32_t carry[2];
    const lh_view& v = vs.v;
    const Bins B = bins(v);
    const uint32_t tiles = B.L.tiles, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry[0] = carry[1] = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < tiles; t0 += kScanThreads) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t c = t < tiles ? B.tile_count[t] : 0u;
        const uint32_t ni = (c + item_entries - 1) / item_entries;
        const uint32_t ic = wave_inclusive_scan(c), ii = wave_inclusive_scan(ni);
        if (lane == 63) {
            wsum[0][wave] = ic;
            wsum[1][wave] = ii;
        }
        __syncthreads();
        uint32_t bc = carry[0], bi = carry[1];
        for (uint32_t w = 0; w < wave; w++) {
            bc += wsum[0][w];
            bi += wsum[1][w];
        }
        const uint32_t off = bc + ic - c, ioff = bi + ii - ni;
        if (t < tiles) {
            B.tile_cursor[t] = off;
            for (uint32_t j = 0; j < ni; j++) {
                const uint32_t it = ioff + j;
                if (it < B.L.item_capacity) {
                    const uint32_t b = off + j * item_entries, e = min(b + item_entries, off + c);
                    B.items[it] = TileItem{t, b, e | (ni == 1 ? lhs::kItemOnly : 0u)};
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) {
            carry[0] = off + c;
            carry[1] = ioff + ni;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        v.counters[LH_CNT_TILE_ITEMS] = min(carry[1], B.L.item_capacity);
        v.counters[LH_CNT_TILE_NEXT] = 0;  // the tile raster's work queue starts empty
    }
}

// One thread per binned draw: its entries at the tiles' cursors.  Per run of
// kBinChunk draws a block counts its entries per tile in LDS, reserves each
// tile's range with one device atomic, then hands out the slots in LDS.
__global__ void __launch_bounds__(kBinThreads) k_bin_fill(OneView vs, PassSet ps) {
    __shared__ uint32_t cnt[kHistTiles];
    const lh_view& v = vs.v;
    const Segments<OneView> seg(vs, ps);
    const uint32_t nb = binned_draws(seg, v);
    const Bins B = bins(v);
    auto binned = [&](uint32_t g, uint32_t& rect, uint32_t& parts) {
        if (g >= nb) return false;
        const BinHeader h = B.hdr[g];
        rect = h.rect;
        parts = 1u << lhs::parts_log2(h.ntri);
        return h.ntri != 0 && (h.ntri & lhs::kBinDirect) == 0;
    };
    if (B.L.tiles > kHistTiles) {
        for (uint32_t g = blockIdx.x * kBinThreads + threadIdx.x; g < nb; g += gridDim.x * kBinThreads) {
            uint32_t rect, parts;
            if (binned(g, rect, parts))
                for_each_tile(rect, B.L.tiles_x, [&](uint32_t t) {
                    const uint32_t at = atomicAdd(&B.tile_cursor[t], parts);
                    for (uint32_t q = 0; q < parts; q++)
                        if (at + q < B.L.entry_capacity) B.entries[at + q] = g | (q << lhs::kEntryPartShift);
                });
        }
        return;
    }
    for (uint32_t c0 = blockIdx.x * kBinChunk; c0 < nb; c0 += gridDim.x * kBinChunk) {
        for (uint32_t t = threadIdx.x; t < B.L.tiles; t += kBinThreads) cnt[t] = 0;
        __syncthreads();
        for (uint32_t g = c0 + threadIdx.x; g < min(c0 + kBinChunk, nb); g += kBinThreads) {
            uint32_t rect, parts;
            if (binned(g, rect, parts))
                for_each_tile(rect, B.L.tiles_x, [&](uint32_t t) { atomicAdd(&cnt[t], parts); });
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < B.L.tiles; t += kBinThreads)
            if (cnt[t]) cnt[t] = atomicAdd(&B.tile_cursor[t], cnt[t]);
        __syncthreads();
        for (uint32_t g = c0 + threadIdx.x; g < min(c0 + kBinChunk, nb); g += kBinThreads) {
            uint32_t rect, parts;
            if (binned(g, rect, parts))
                for_each_tile(rect, B.L.tiles_x, [&](uint32_t t) {
                    const uint32_t at = atomicAdd(&cnt[t], parts);
                    for (uint32_t q = 0; q < parts; q++)
                        if (at + q < B.L.entry_capacity) B.entries[at + q] = g | (q << lhs::kEntryPartShift);
                });
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- tile raster

// LDS slot of pixel (lx, ly) of the tile.  A row of 64 keys spans the LDS
// banks twice, so every row starts on the same bank; the column is XORed
// with the row so that the rows of a box handed to one wave spread over them.
LH_DEV int32_t tile_slot(int32_t lx, int32_t ly) { return ly * int32_t(kTile) + (lx ^ ((ly & 7) << 3)); }

// Depth test of one covered sample in the LDS tile (lx, ly: pixel in the tile).
LH_DEV bool tile_fragment(uint64_t* tile, int32_t lx, int32_t ly, float z, uint32_t id) {
    z = fminf(fmaxf(z, 0.0f), 1.0f);
    if (!(z > 0.0f)) return false;  // depth GREATER against the 0.0 clear
    const unsigned long long key = ((unsigned long long)(~__float_as_uint(z)) << 32) | id;
    __hip_atomic_fetch_min(reinterpret_cast<unsigned long long*>(tile) + tile_slot(lx, ly), key, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    return true;
}

// Materializes v in a register at this point.  The triangle's values come
// from LDS reads; left to itself the compiler recomputes them inside a pixel
// loop, and the LDS-counter wait those reads need then also waits for every
// depth atomic the loop has issued.
template <class T>
LH_DEV void pin(T& v) {
    asm volatile("" : "+v"(v));
}

// Coverage of one pixel centre by an oriented triangle (exact integer edges).
template <class E>
LH_DEV bool inside(E e0, E e1, E e2, bool tl0, bool tl1, bool tl2) {
    return (e0 > 0 || (e0 == 0 && tl0)) && (e1 > 0 || (e1 == 0 && tl1)) && (e2 > 0 || (e2 == 0 && tl2));
}

// In-lane raster of a triangle whose box inside the tile is small (at most
// small_extent pixels a side): the lane walks the box itself.
template <bool ST>
LH_DEV void tile_lane(uint64_t* tile, int32_t ox, int32_t oy, SV a, SV b, SV c, int32_t xlo, int32_t xhi, int32_t ylo,
                      int32_t yhi, float inv_area, uint32_t id, Tally<ST>& tally) {
    pin(a.X), pin(a.Y), pin(b.X), pin(b.Y), pin(c.X), pin(c.Y);
    float z0 = a.z, d10 = b.z - a.z, d20 = c.z - a.z;
    pin(z0), pin(d10), pin(d20), pin(inv_area);
    const bool tl0 = top_left(b.X, b.Y, c.X, c.Y), tl1 = top_left(c.X, c.Y, a.X, a.Y), tl2 = top_left(a.X, a.Y, b.X, b.Y);
    for (int32_t py = ylo; py <= yhi; py++) {
        const int32_t Py = py * 256 + 128;
        for (int32_t px = xlo; px <= xhi; px++) {
            const int32_t Px = px * 256 + 128;
            const int32_t e0 = (c.X - b.X) * (Py - b.Y) - (c.Y - b.Y) * (Px - b.X);
            const int32_t e1 = (a.X - c.X) * (Py - c.Y) - (a.Y - c.Y) * (Px - c.X);
            const int32_t e2 = (b.X - a.X) * (Py - a.Y) - (b.Y - a.Y) * (Px - a.X);
            if (inside(e0, e1, e2, tl0, tl1, tl2)) {
                const float l1 = float(e1) * inv_area, l2 = float(e2) * inv_area;
                tally.hit(tile_fragment(tile, px - ox, py - oy, (z0 + l1 * d10) + l2 * d20, id));
            }
        }
    }
}

// Load-balanced raster of one round of triangles (one per lane) over the
// wave.  A lane's clamped box is cut into runs of up to kRun pixels along its
// rows (`n` runs, 0 for none); the runs are numbered in lane order (exclusive
// scan) and handed out 64 at a time, so a wave spends ceil(sum n / 64) passes
// on a round instead of the largest box.  A lane finds the triangle of its
// run as the last one starting at or before it (start lanes mark their first
// run in `own`, an inclusive max-scan spreads the marks), fetches the
// triangle's corners, box and 1/area from that lane and its snapped vertices
// from the wave's LDS copy, and steps the edge functions along the run.
constexpr int32_t kRun = 4;
LH_DEV int32_t runs_per_row(int32_t w) { return (w + kRun - 1) / kRun; }

template <bool ST>
LH_DEV void tile_expand(uint64_t* tile, int32_t ox, int32_t oy, int32_t* own, const int32_t* sX, const int32_t* sY,
                        const float* sZ, uint32_t tri, uint32_t n, int32_t xlo, int32_t ylo, int32_t w, float inv_area,
                        uint32_t mid, uint32_t lane, Tally<ST>& tally) {
    const uint32_t incl = wave_scan_add(n), excl = incl - n;
    const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
    const int32_t box = (xlo - ox) | ((ylo - oy) << 8) | (w << 16);  // tile-relative origin and width
    int32_t carry = 0;
    for (uint32_t base = 0; base < total; base += 64) {
        own[lane] = -1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (n && excl >= base && excl < base + 64) own[excl - base] = int32_t(lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int32_t o = max(wave_scan_max(own[lane]), carry);
        carry = __builtin_amdgcn_readlane(o, 63);
        // every lane takes part in the fetches (the owner lanes are active)
        const uint32_t otri = __shfl(tri, o, 64);
        const int32_t obox = __shfl(box, o, 64);
        const uint32_t oex = __shfl(excl, o, 64);
        const float oinv = __shfl(inv_area, o, 64);
        const uint32_t omid = uint32_t(__shfl(int(mid), o, 64));
        const uint32_t k = base + lane;
        if (k < total) {
            // run (local % wr, local / wr) of the owner's box; local < 4096, wr <= 16
            const int32_t ow = obox >> 16, wr = runs_per_row(ow);
            const int32_t local = int32_t(k - oex);
            int32_t q = int32_t(float(local) * __builtin_amdgcn_rcpf(float(wr)));
            int32_t r = local - q * wr;
            if (r < 0) q--, r += wr;
            if (r >= wr) q++, r -= wr;
            const int32_t lx0 = (obox & 255) + kRun * r, ly = ((obox >> 8) & 255) + q;
            const int32_t len = min(kRun, ow - kRun * r);
            const uint32_t i0 = otri & 255u, i1 = (otri >> 8) & 255u, i2 = (otri >> 16) & 255u;
            const SV a{sX[i0], sY[i0], sZ[i0]}, b{sX[i1], sY[i1], sZ[i1]}, c{sX[i2], sY[i2], sZ[i2]};
            const int32_t Px = (ox + lx0) * 256 + 128, Py = (oy + ly) * 256 + 128;
            int32_t e0 = (c.X - b.X) * (Py - b.Y) - (c.Y - b.Y) * (Px - b.X);
            int32_t e1 = (a.X - c.X) * (Py - c.Y) - (a.Y - c.Y) * (Px - c.X);
            int32_t e2 = (b.X - a.X) * (Py - a.Y) - (b.Y - a.Y) * (Px - a.X);
            int32_t s0 = -(c.Y - b.Y) * 256, s1 = -(a.Y - c.Y) * 256, s2 = -(b.Y - a.Y) * 256;
            bool tl0 = top_left(b.X, b.Y, c.X, c.Y), tl1 = top_left(c.X, c.Y, a.X, a.Y),
                 tl2 = top_left(a.X, a.Y, b.X, b.Y);
            float z0 = a.z, d10 = b.z - a.z, d20 = c.z - a.z, inv = oinv;
            pin(z0), pin(d10), pin(d20), pin(inv), pin(s0), pin(s1), pin(s2);
            const uint32_t id = (omid << LH_KEY_TRI_BITS) | (otri >> 24);
            for (int32_t j = 0; j < len; j++) {
                if (inside(e0, e1, e2, tl0, tl1, tl2)) {
                    const float l1 = float(e1) * inv, l2 = float(e2) * inv;
                    tally.hit(tile_fragment(tile, lx0 + j, ly, (z0 + l1 * d10) + l2 * d20, id));
                }
                e0 += s0;
                e1 += s1;
                e2 += s2;
            }
        }
    }
}

// Persistent blocks of kTileWaves waves over the work items.  Item: clear the
// LDS tile; each wave takes every kTileWaves-th entry (one binned meshlet),
// with the entry list of the item held one per lane and the next meshlet's
// cache lines requested while the current one is rasterized.  When an entry
// and the next hold at most 64 triangles together (an entry averages about
// 35), both go through one round: the second's vertices at LDS slots
// 64..127, its triangles on the lanes after the first's, each lane with its
// own meshlet id.  Then the covered pixels go out: a plain store when the
// item is the tile's only one (the visibility buffer holds the clear value
// there, as nothing else has drawn yet), an atomicMin otherwise and in the
// second occlusion phase.  Six waves per SIMD (80 VGPRs; the pairing's
// registers would otherwise take the kernel to five).
struct CacheLanes {
    int32_t X, Y;
    float Z;
    uint32_t t0, t1;
};
// The triangles [first, first + count) of an entry: the whole list, or one
// of the 2^lp parts it is split into.
LH_DEV void entry_range(uint32_t ntri_field, uint32_t part, uint32_t& first, uint32_t& count) {
    const uint32_t ntri = ntri_field & 255u, lp = lhs::parts_log2(ntri_field);
    first = (part * ntri) >> lp;
    count = (((part + 1) * ntri) >> lp) - first;
}
LH_DEV CacheLanes load_cache(const BinCache& c, uint32_t lane, uint32_t first) {
    // every lane reads its word whatever the meshlet's vertex and triangle
    // counts: the entry is allocated whole, and words past the counts are
    // never referenced
    return {c.X[lane], c.Y[lane], c.Z[lane], c.tri[min(first + lane, 127u)], c.tri[min(first + lane + 64, 127u)]};
}

template <bool ST>
__global__ void __launch_bounds__(kTileThreads) LH_WAVES_ATTR(6) k_tile_raster(OneView vs, PassSet ps) {
    __shared__ uint64_t tile[kTile * kTile];
    __shared__ int32_t sX[kTileWaves][128], sY[kTileWaves][128];
    __shared__ float sZ[kTileWaves][128];
    __shared__ int32_t own[kTileWaves][64];
    const lh_view& v = vs.v;
    const Bins B = bins(v);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t n_items = min(v.counters[LH_CNT_TILE_ITEMS], B.L.item_capacity);
    const int32_t W = int32_t(v.width), H = int32_t(v.height);
    Tally<ST> tally;
    // Block b starts on item b; each further item comes from a queue, so a
    // block that drew light items takes more of them instead of idling while
    // the heavy ones finish.
    __shared__ uint32_t next_item;
    for (uint32_t it = blockIdx.x; it < n_items;) {
        const TileItem item = B.items[it];
        const int32_t ox = int32_t(item.tile % B.L.tiles_x) * int32_t(kTile);
        const int32_t oy = int32_t(item.tile / B.L.tiles_x) * int32_t(kTile);
        const int32_t ex = min(ox + int32_t(kTile), W) - 1, ey = min(oy + int32_t(kTile), H) - 1;
        const uint32_t end = min(item.end & ~lhs::kItemOnly, B.L.entry_capacity);
        // this wave's entries: begin + wave + kTileWaves k, k < mine (<= 64: item_entries <= 512)
        const uint32_t first = item.begin + wave;
        const uint32_t mine = end > first ? min((end - first + kTileWaves - 1) / kTileWaves, 64u) : 0u;
        const uint32_t my_e = lane < mine ? B.entries[first + kTileWaves * lane] : 0u;
        const uint32_t my_g = my_e & lhs::kEntryDrawMask;
        uint32_t my_first = 0, my_count = 0, my_mid = 0;
        if (lane < mine) {
            const uint2 h = *reinterpret_cast<const uint2*>(&B.hdr[my_g].ntri);
            entry_range(h.x, my_e >> lhs::kEntryPartShift, my_first, my_count);
            my_mid = h.y;
        }
        for (uint32_t k = threadIdx.x; k < kTile * kTile; k += kTileThreads) tile[k] = LH_VISBUFFER_CLEAR;
        CacheLanes cur{};
        if (mine)
            cur = load_cache(B.cache[uint32_t(readlane(int32_t(my_g), 0))], lane,
                             uint32_t(readlane(int32_t(my_first), 0)));
        __syncthreads();
        for (uint32_t k = 0; k < mine;) {
            const uint32_t ntri = uint32_t(readlane(int32_t(my_count), k)), mid = uint32_t(readlane(int32_t(my_mid), k));
            const uint32_t ntri2 = k + 1 < mine ? uint32_t(readlane(int32_t(my_count), k + 1)) : 65u;
            const bool pair = ntri + ntri2 <= 64u;  // wave-uniform
            const uint32_t mid2 = pair ? uint32_t(readlane(int32_t(my_mid), k + 1)) : 0u;
            CacheLanes next{};
            if (k + 1 < mine)
                next = load_cache(B.cache[uint32_t(readlane(int32_t(my_g), k + 1))], lane,
                                  uint32_t(readlane(int32_t(my_first), k + 1)));
            sX[wave][lane] = cur.X;
            sY[wave][lane] = cur.Y;
            sZ[wave][lane] = cur.Z;
            uint32_t tri_pair = cur.t0, lane_mid = mid;
            if (pair) {
                sX[wave][64 + lane] = next.X;
                sY[wave][64 + lane] = next.Y;
                sZ[wave][64 + lane] = next.Z;
                const uint32_t from = uint32_t(__shfl(int(next.t0), int((lane - ntri) & 63u), 64));
                if (lane >= ntri) {
                    tri_pair = from + 0x404040u;  // corners of the second entry at slots 64..127
                    lane_mid = mid2;
                }
                // the second entry is consumed: its registers take the request
                // of the entry after the pair
                if (k + 2 < mine)
                    next = load_cache(B.cache[uint32_t(readlane(int32_t(my_g), k + 2))], lane,
                                      uint32_t(readlane(int32_t(my_first), k + 2)));
            }
            const uint32_t total = pair ? ntri + ntri2 : ntri;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t round = 0; round < 2 && round * 64 < total; round++) {
                const uint32_t tri = pair ? tri_pair : (round ? cur.t1 : cur.t0);
                uint32_t n = 0;
                int32_t xlo = 0, ylo = 0, w = 1;
                float inv_area = 0.0f;
                if (lane + 64 * round < total) {
                    const uint32_t i0 = tri & 255u, i1 = (tri >> 8) & 255u, i2 = (tri >> 16) & 255u;
                    const SV a{sX[wave][i0], sY[wave][i0], sZ[wave][i0]};
                    const SV b{sX[wave][i1], sY[wave][i1], sZ[wave][i1]};
                    const SV c{sX[wave][i2], sY[wave][i2], sZ[wave][i2]};
                    const int32_t minX = min(a.X, min(b.X, c.X)), maxX = max(a.X, max(b.X, c.X));
                    const int32_t minY = min(a.Y, min(b.Y, c.Y)), maxY = max(a.Y, max(b.Y, c.Y));
                    xlo = max((minX + 127) >> 8, ox);
                    const int32_t xhi = min((maxX - 128) >> 8, ex);
                    ylo = max((minY + 127) >> 8, oy);
                    const int32_t yhi = min((maxY - 128) >> 8, ey);
                    if (xlo <= xhi && ylo <= yhi) {
                        w = xhi - xlo + 1;
                        const int32_t h = yhi - ylo + 1;
                        // positive area (oriented by the binning), every value within int32 (<= 64 px)
                        inv_area = rcp_exact(float((b.X - a.X) * (c.Y - a.Y) - (b.Y - a.Y) * (c.X - a.X)));
                        if (max(w, h) * 256 <= ps.small_extent)
                            tile_lane<ST>(tile, ox, oy, a, b, c, xlo, xhi, ylo, yhi, inv_area,
                                          (lane_mid << LH_KEY_TRI_BITS) | (tri >> 24), tally);
                        else
                            n = uint32_t(runs_per_row(w) * h);
                    }
                }
                tile_expand<ST>(tile, ox, oy, own[wave], sX[wave], sY[wave], sZ[wave], tri, n, xlo, ylo, w, inv_area,
                                lane_mid, lane, tally);
            }
            k += pair ? 2u : 1u;
            cur = next;
        }
        tally.flush(v, 0);
        __syncthreads();
        const bool only = (item.end & lhs::kItemOnly) && !ps.phase2;
        for (uint32_t k = threadIdx.x; k < kTile * kTile; k += kTileThreads) {
            const int32_t lx = int32_t(k & (kTile - 1)), ly = int32_t(k >> lhs::kTileShift);
            const int32_t px = ox + lx, py = oy + ly;
            const uint64_t key = tile[tile_slot(lx, ly)];
            if (px > ex || py > ey || key == LH_VISBUFFER_CLEAR) continue;
            unsigned long long* dst = reinterpret_cast<unsigned long long*>(v.visbuffer + size_t(py) * v.width + px);
            if (only)
                *dst = key;
            else
                atomicMin(dst, (unsigned long long)key);
        }
        if (threadIdx.x == 0) next_item = gridDim.x + atomicAdd(&v.counters[LH_CNT_TILE_NEXT], 1u);
        __syncthreads();  // the item's LDS tile is free; next_item is written
        it = next_item;   // rewritten only after the next item's barriers
    }
}

}  // namespace

namespace lhk {

template <bool ST>
hipError_t raster_tiles(const lh_scene& s, const OneView& vs, const PassSet& ps, const RasterTuning& t,
                        hipStream_t st) {
    const uint32_t item_entries = std::min(std::max(t.tile_item_entries, lhs::kMinItemEntries), 64u * kTileWaves);
    mark("bin_meshlets_view", st);
    hipLaunchKernelGGL(k_bin_meshlets<0>, dim3(t.meshlet_blocks), dim3(64), 0, st, s, vs, ps);
    // the draw count is on the device: grids sized for the view's draw capacity
    const uint32_t chunks = std::max<uint32_t>(1, (std::min(vs.v.draw_capacity, 1u << 22) + kBinChunk - 1) / kBinChunk);
    const uint32_t bin_blocks = std::min<uint32_t>(chunks, 256);
    mark("bin_count_view", st);
    hipLaunchKernelGGL(k_bin_count, dim3(bin_blocks), dim3(kBinThreads), 0, st, vs, ps);
    mark("bin_scan_view", st);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kScanThreads), 0, st, vs, item_entries);
    mark("bin_fill_view", st);
    hipLaunchKernelGGL(k_bin_fill, dim3(bin_blocks), dim3(kBinThreads), 0, st, vs, ps);
    mark("tile_raster_view", st);
    hipLaunchKernelGGL((k_tile_raster<ST>), dim3(t.tile_blocks), dim3(kTileThreads), 0, st, vs, ps);
    return hipGetLastError();
}

template hipError_t raster_tiles<false>(const lh_scene&, const OneView&, const PassSet&, const RasterTuning&,
                                        hipStream_t);
template hipError_t raster_tiles<true>(const lh_scene&, const OneView&, const PassSet&, const RasterTuning&,
                                       hipStream_t);

}  // namespace lhk
