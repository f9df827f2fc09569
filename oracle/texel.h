// Texel decoding shared by the oracle's FP32 frame (oracle.cpp) and its FP64
// adjoint (oracle64.cpp).  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
//
// LhTexture formats (include/lh/shared_cpu_gpu.h): RGBA8 texels, or BC1
// (S3TC) blocks of 4x4 texels: two little-endian RGB565 endpoints c0, c1 and
// 32 bits of 2-bit indices, texel (x, y) of the block at bits 2 (4 y + x).
// Endpoints are widened to 8 bits by bit replication; when c0 > c1 the
// palette is c0, c1, (2 c0 + c1 + 1) / 3, (c0 + 2 c1 + 1) / 3 (opaque), else
// c0, c1, (c0 + c1 + 1) / 2 and transparent black.
#pragma once

#include <cstdint>
#include <cstring>

#include "lh/shared_cpu_gpu.h"

namespace orc_texel {

struct RGBA8 {
    uint8_t c[4];
};

inline void widen(uint32_t v, uint32_t out[3]) {
    uint32_t r = (v >> 11) & 31u, g = (v >> 5) & 63u, b = v & 31u;
    out[0] = (r << 3) | (r >> 2);
    out[1] = (g << 2) | (g >> 4);
    out[2] = (b << 3) | (b >> 2);
}

// Colour `sel` of a BC1 block with endpoints c0, c1.
inline RGBA8 bc1(uint32_t c0, uint32_t c1, uint32_t sel) {
    uint32_t a[3], b[3];
    widen(c0, a);
    widen(c1, b);
    RGBA8 o{{0, 0, 0, 255}};
    for (int k = 0; k < 3; k++) {
        uint32_t m;
        if (sel == 0) m = a[k];
        else if (sel == 1) m = b[k];
        else if (c0 > c1) m = sel == 2 ? (2 * a[k] + b[k] + 1) / 3 : (a[k] + 2 * b[k] + 1) / 3;
        else m = sel == 2 ? (a[k] + b[k] + 1) / 2 : 0;
        o.c[k] = uint8_t(m);
    }
    if (c0 <= c1 && sel == 3) o.c[3] = 0;
    return o;
}

// Texel (i, j), already wrapped, of a level of width w whose data starts at
// `base` (the texture's level-0 address).
inline RGBA8 fetch(const LhTexture& t, const uint8_t* base, uint32_t level, uint32_t w, uint32_t i, uint32_t j) {
    RGBA8 o;
    if (t.format != LH_TEXTURE_BC1) {
        std::memcpy(o.c, base + 4 * (size_t(t.level_offset[level]) + size_t(j) * w + i), 4);
        return o;
    }
    const uint8_t* blk = base + 8 * (size_t(t.level_offset[level]) + size_t(j >> 2) * ((w + 3) >> 2) + (i >> 2));
    uint32_t c0 = uint32_t(blk[0]) | (uint32_t(blk[1]) << 8), c1 = uint32_t(blk[2]) | (uint32_t(blk[3]) << 8);
    uint32_t idx = uint32_t(blk[4]) | (uint32_t(blk[5]) << 8) | (uint32_t(blk[6]) << 16) | (uint32_t(blk[7]) << 24);
    return bc1(c0, c1, (idx >> (2 * (4 * (j & 3) + (i & 3)))) & 3u);
}

}  // namespace orc_texel
