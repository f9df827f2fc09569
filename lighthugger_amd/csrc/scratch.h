// Layout of lh_view.scratch (the raster's queues and the tiled raster's bins),
// shared by the host (lh_scratch_bytes, argument checks) and the kernels.
#pragma once

#include <stdint.h>

#include "lh/lh.h"

#ifndef LH_HD
#define LH_HD __host__ __device__ __forceinline__
#endif

namespace lhs {

constexpr uint32_t kTileShift = 6;  // log2(LH_TILE_SIDE)
static_assert((1u << kTileShift) == LH_TILE_SIDE, "tile side");
constexpr uint32_t kMinItemEntries = 8;  // smallest run of entries per work item (bounds the item list)

struct BinHeader {
    uint32_t rect;  // pass-local tile rectangle: x0 | y0 << 8 | x1 << 16 | y1 << 24 (inclusive)
    uint32_t ntri;  // front-facing triangles in the cache; bit 31: left to the per-meshlet raster
    uint32_t mid;   // stable meshlet id (key bits above the triangle index)
    uint32_t pad;
};
static_assert(sizeof(BinHeader) == 16, "bin header");
constexpr uint32_t kBinDirect = 0x80000000u;
// BinHeader.ntri bits 8..10: log2 of the parts a meshlet with much pixel work
// is split into (by triangle range); entries carry the part above the draw.
constexpr uint32_t kPartShift = 8;
constexpr uint32_t kMaxPartsLog2 = 3;
constexpr uint32_t kPartWorkLog2 = 10;  // about 2^10 box pixels per part
constexpr uint32_t kEntryPartShift = 25;
constexpr uint32_t kEntryDrawMask = (1u << kEntryPartShift) - 1;
LH_HD uint32_t parts_log2(uint32_t ntri_field) { return (ntri_field >> kPartShift) & 7u; }

// Snapped vertices of one draw (vertex i of the meshlet at index i) and its
// front-facing triangles, oriented to positive area:
// tri = i0 | i1 << 8 | i2 << 16 | triangle index << 24.
struct BinCache {
    int32_t X[64];
    int32_t Y[64];
    float Z[64];
    uint32_t tri[128];
};
static_assert(sizeof(BinCache) == LH_BIN_CACHE_BYTES, "bin cache entry");

struct TileItem {
    uint32_t tile;   // tile index (row-major over the view)
    uint32_t begin;  // entries [begin, end & ~kItemOnly)
    uint32_t end;    // bit 31: the only item of its tile
};
constexpr uint32_t kItemOnly = 0x80000000u;

struct ScratchLayout {
    uint64_t clip, big, hdr, cache, tile_count, tile_cursor, items, entries, direct, end;
    uint32_t tiles_x, tiles_y, tiles, item_capacity, entry_capacity, direct_capacity;
};

LH_HD uint64_t align256(uint64_t n) { return (n + 255) & ~uint64_t(255); }

LH_HD ScratchLayout scratch_layout(uint32_t width, uint32_t height, uint32_t draw_capacity, uint32_t big_capacity,
                                   uint32_t clip_capacity, uint32_t bin_capacity) {
    ScratchLayout L{};
    uint64_t o = 0;
    L.clip = o;
    o += align256(uint64_t(clip_capacity) * LH_CLIP_ENTRY_BYTES);
    L.big = o;
    o += align256(uint64_t(big_capacity) * LH_BIG_ENTRY_BYTES);
    if (bin_capacity) {
        L.tiles_x = (width + LH_TILE_SIDE - 1) >> kTileShift;
        L.tiles_y = (height + LH_TILE_SIDE - 1) >> kTileShift;
        L.tiles = L.tiles_x * L.tiles_y;
        L.entry_capacity = LH_TILE_MAX_RECT * bin_capacity;  // every binned draw fits
        L.item_capacity = L.tiles + L.entry_capacity / kMinItemEntries;
        L.direct_capacity = draw_capacity;
        L.hdr = o;
        o += align256(uint64_t(bin_capacity) * sizeof(BinHeader));
        L.cache = o;
        o += align256(uint64_t(bin_capacity) * sizeof(BinCache));
        L.tile_count = o;
        o += align256(uint64_t(L.tiles) * 4);
        L.tile_cursor = o;
        o += align256(uint64_t(L.tiles) * 4);
        L.items = o;
        o += align256(uint64_t(L.item_capacity) * sizeof(TileItem));
        L.entries = o;
        o += align256(uint64_t(L.entry_capacity) * 4);
        L.direct = o;
        o += align256(uint64_t(L.direct_capacity) * 4);
    }
    L.end = o;
    return L;
}

LH_HD ScratchLayout scratch_layout(const lh_view& v) {
    return scratch_layout(v.width, v.height, v.draw_capacity, v.big_capacity, v.clip_capacity, v.bin_capacity);
}

}  // namespace lhs
