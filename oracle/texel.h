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

#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>

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

// ---- BC7 (BPTC) ----
// Decoded as the published format describes it (Khronos Data Format
// Specification, "BC7"; the D3D11 functional specification): the block read
// field after field from bit 0 (LSB first): the mode as a unary prefix, the
// partition, rotation and index-selection fields, colour endpoints channel
// after channel, alpha endpoints, p-bits, then the index planes, where the
// anchor texel of every subset stores one bit less.  Endpoints with a p-bit
// append it; every endpoint is widened to 8 bits by replicating its top bits;
// texels are ((64 - w) e0 + w e1 + 32) >> 6 with the 2-, 3- or 4-bit weight
// tables; rotation 1..3 swaps alpha with R, G or B.  A block without a mode
// bit in its first byte decodes to transparent black.
namespace bc7 {
struct Mode {
    int ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2;
};
// subsets, partition bits, rotation bits, index-selection bits, colour bits,
// alpha bits, per-endpoint p-bits, shared p-bits, index bits, second index bits
constexpr Mode kModes[8] = {
    {3, 4, 0, 0, 4, 0, 1, 0, 3, 0}, {2, 6, 0, 0, 6, 0, 0, 1, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 0, 2, 0},
    {2, 6, 0, 0, 7, 0, 1, 0, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 0, 2, 2},
    {1, 0, 0, 0, 7, 7, 1, 0, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 0, 2, 0},
};
constexpr int kW2[4] = {0, 21, 43, 64};
constexpr int kW3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
constexpr int kW4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};
// Two-subset partitions, texels in row order (x fastest).
constexpr uint8_t kPart2[64][16] = {
    {0,0,1,1,0,0,1,1,0,0,1,1,0,0,1,1}, {0,0,0,1,0,0,0,1,0,0,0,1,0,0,0,1}, {0,1,1,1,0,1,1,1,0,1,1,1,0,1,1,1},
    {0,0,0,1,0,0,1,1,0,0,1,1,0,1,1,1}, {0,0,0,0,0,0,0,1,0,0,0,1,0,0,1,1}, {0,0,1,1,0,1,1,1,0,1,1,1,1,1,1,1},
    {0,0,0,1,0,0,1,1,0,1,1,1,1,1,1,1}, {0,0,0,0,0,0,0,1,0,0,1,1,0,1,1,1}, {0,0,0,0,0,0,0,0,0,0,0,1,0,0,1,1},
    {0,0,1,1,0,1,1,1,1,1,1,1,1,1,1,1}, {0,0,0,0,0,0,0,1,0,1,1,1,1,1,1,1}, {0,0,0,0,0,0,0,0,0,0,0,1,0,1,1,1},
    {0,0,0,1,0,1,1,1,1,1,1,1,1,1,1,1}, {0,0,0,0,0,0,0,0,1,1,1,1,1,1,1,1}, {0,0,0,0,1,1,1,1,1,1,1,1,1,1,1,1},
    {0,0,0,0,0,0,0,0,0,0,0,0,1,1,1,1}, {0,0,0,0,1,0,0,0,1,1,1,0,1,1,1,1}, {0,1,1,1,0,0,0,1,0,0,0,0,0,0,0,0},
    {0,0,0,0,0,0,0,0,1,0,0,0,1,1,1,0}, {0,1,1,1,0,0,1,1,0,0,0,1,0,0,0,0}, {0,0,1,1,0,0,0,1,0,0,0,0,0,0,0,0},
    {0,0,0,0,1,0,0,0,1,1,0,0,1,1,1,0}, {0,0,0,0,0,0,0,0,1,0,0,0,1,1,0,0}, {0,1,1,1,0,0,1,1,0,0,1,1,0,0,0,1},
    {0,0,1,1,0,0,0,1,0,0,0,1,0,0,0,0}, {0,0,0,0,1,0,0,0,1,0,0,0,1,1,0,0}, {0,1,1,0,0,1,1,0,0,1,1,0,0,1,1,0},
    {0,0,1,1,0,1,1,0,0,1,1,0,1,1,0,0}, {0,0,0,1,0,1,1,1,1,1,1,0,1,0,0,0}, {0,0,0,0,1,1,1,1,1,1,1,1,0,0,0,0},
    {0,1,1,1,0,0,0,1,1,0,0,0,1,1,1,0}, {0,0,1,1,1,0,0,1,1,0,0,1,1,1,0,0}, {0,1,0,1,0,1,0,1,0,1,0,1,0,1,0,1},
    {0,0,0,0,1,1,1,1,0,0,0,0,1,1,1,1}, {0,1,0,1,1,0,1,0,0,1,0,1,1,0,1,0}, {0,0,1,1,0,0,1,1,1,1,0,0,1,1,0,0},
    {0,0,1,1,1,1,0,0,0,0,1,1,1,1,0,0}, {0,1,0,1,0,1,0,1,1,0,1,0,1,0,1,0}, {0,1,1,0,1,0,0,1,0,1,1,0,1,0,0,1},
    {0,1,0,1,1,0,1,0,1,0,1,0,0,1,0,1}, {0,1,1,1,0,0,1,1,1,1,0,0,1,1,1,0}, {0,0,0,1,0,0,1,1,1,1,0,0,1,0,0,0},
    {0,0,1,1,0,0,1,0,0,1,0,0,1,1,0,0}, {0,0,1,1,1,0,1,1,1,1,0,1,1,1,0,0}, {0,1,1,0,1,0,0,1,1,0,0,1,0,1,1,0},
    {0,0,1,1,1,1,0,0,1,1,0,0,0,0,1,1}, {0,1,1,0,0,1,1,0,1,0,0,1,1,0,0,1}, {0,0,0,0,0,1,1,0,0,1,1,0,0,0,0,0},
    {0,1,0,0,1,1,1,0,0,1,0,0,0,0,0,0}, {0,0,1,0,0,1,1,1,0,0,1,0,0,0,0,0}, {0,0,0,0,0,0,1,0,0,1,1,1,0,0,1,0},
    {0,0,0,0,0,1,0,0,1,1,1,0,0,1,0,0}, {0,1,1,0,1,1,0,0,1,0,0,1,0,0,1,1}, {0,0,1,1,0,1,1,0,1,1,0,0,1,0,0,1},
    {0,1,1,0,0,0,1,1,1,0,0,1,1,1,0,0}, {0,0,1,1,1,0,0,1,1,1,0,0,0,1,1,0}, {0,1,1,0,1,1,0,0,1,1,0,0,1,0,0,1},
    {0,1,1,0,0,0,1,1,0,0,1,1,1,0,0,1}, {0,1,1,1,1,1,1,0,1,0,0,0,0,0,0,1}, {0,0,0,1,1,0,0,0,1,1,1,0,0,1,1,1},
    {0,0,0,0,1,1,1,1,0,0,1,1,0,0,1,1}, {0,0,1,1,0,0,1,1,1,1,1,1,0,0,0,0}, {0,0,1,0,0,0,1,0,1,1,1,0,1,1,1,0},
    {0,1,0,0,0,1,0,0,0,1,1,1,0,1,1,1},
};
// Three-subset partitions.
constexpr uint8_t kPart3[64][16] = {
    {0,0,1,1,0,0,1,1,0,2,2,1,2,2,2,2}, {0,0,0,1,0,0,1,1,2,2,1,1,2,2,2,1}, {0,0,0,0,2,0,0,1,2,2,1,1,2,2,1,1},
    {0,2,2,2,0,0,2,2,0,0,1,1,0,1,1,1}, {0,0,0,0,0,0,0,0,1,1,2,2,1,1,2,2}, {0,0,1,1,0,0,1,1,0,0,2,2,0,0,2,2},
    {0,0,2,2,0,0,2,2,1,1,1,1,1,1,1,1}, {0,0,1,1,0,0,1,1,2,2,1,1,2,2,1,1}, {0,0,0,0,0,0,0,0,1,1,1,1,2,2,2,2},
    {0,0,0,0,1,1,1,1,1,1,1,1,2,2,2,2}, {0,0,0,0,1,1,1,1,2,2,2,2,2,2,2,2}, {0,0,1,2,0,0,1,2,0,0,1,2,0,0,1,2},
    {0,1,1,2,0,1,1,2,0,1,1,2,0,1,1,2}, {0,1,2,2,0,1,2,2,0,1,2,2,0,1,2,2}, {0,0,1,1,0,1,1,2,1,1,2,2,1,2,2,2},
    {0,0,1,1,2,0,0,1,2,2,0,0,2,2,2,0}, {0,0,0,1,0,0,1,1,0,1,1,2,1,1,2,2}, {0,1,1,1,0,0,1,1,2,0,0,1,2,2,0,0},
    {0,0,0,0,1,1,2,2,1,1,2,2,1,1,2,2}, {0,0,2,2,0,0,2,2,0,0,2,2,1,1,1,1}, {0,1,1,1,0,1,1,1,0,2,2,2,0,2,2,2},
    {0,0,0,1,0,0,0,1,2,2,2,1,2,2,2,1}, {0,0,0,0,0,0,1,1,0,1,2,2,0,1,2,2}, {0,0,0,0,1,1,0,0,2,2,1,0,2,2,1,0},
    {0,1,2,2,0,1,2,2,0,0,1,1,0,0,0,0}, {0,0,1,2,0,0,1,2,1,1,2,2,2,2,2,2}, {0,1,1,0,1,2,2,1,1,2,2,1,0,1,1,0},
    {0,0,0,0,0,1,1,0,1,2,2,1,1,2,2,1}, {0,0,2,2,1,1,0,2,1,1,0,2,0,0,2,2}, {0,1,1,0,0,1,1,0,2,0,0,2,2,2,2,2},
    {0,0,1,1,0,1,2,2,0,1,2,2,0,0,1,1}, {0,0,0,0,2,0,0,0,2,2,1,1,2,2,2,1}, {0,0,0,0,0,0,0,2,1,1,2,2,1,2,2,2},
    {0,2,2,2,0,0,2,2,0,0,1,2,0,0,1,1}, {0,0,1,1,0,0,1,2,0,0,2,2,0,2,2,2}, {0,1,2,0,0,1,2,0,0,1,2,0,0,1,2,0},
    {0,0,0,0,1,1,1,1,2,2,2,2,0,0,0,0}, {0,1,2,0,1,2,0,1,2,0,1,2,0,1,2,0}, {0,1,2,0,2,0,1,2,1,2,0,1,0,1,2,0},
    {0,0,1,1,2,2,0,0,1,1,2,2,0,0,1,1}, {0,0,1,1,1,1,2,2,2,2,0,0,0,0,1,1}, {0,1,0,1,0,1,0,1,2,2,2,2,2,2,2,2},
    {0,0,0,0,0,0,0,0,2,1,2,1,2,1,2,1}, {0,0,2,2,1,1,2,2,0,0,2,2,1,1,2,2}, {0,0,2,2,0,0,1,1,0,0,2,2,0,0,1,1},
    {0,2,2,0,1,2,2,1,0,2,2,0,1,2,2,1}, {0,1,0,1,2,2,2,2,2,2,2,2,0,1,0,1}, {0,0,0,0,2,1,2,1,2,1,2,1,2,1,2,1},
    {0,1,0,1,0,1,0,1,0,1,0,1,2,2,2,2}, {0,2,2,2,0,1,1,1,0,2,2,2,0,1,1,1}, {0,0,0,2,1,1,1,2,0,0,0,2,1,1,1,2},
    {0,0,0,0,2,1,1,2,2,1,1,2,2,1,1,2}, {0,2,2,2,0,1,1,1,0,1,1,1,0,2,2,2}, {0,0,0,2,1,1,1,2,1,1,1,2,0,0,0,2},
    {0,1,1,0,0,1,1,0,0,1,1,0,2,2,2,2}, {0,0,0,0,0,0,0,0,2,1,1,2,2,1,1,2}, {0,1,1,0,0,1,1,0,2,2,2,2,2,2,2,2},
    {0,0,2,2,0,0,1,1,0,0,1,1,0,0,2,2}, {0,0,2,2,1,1,2,2,1,1,2,2,0,0,2,2}, {0,0,0,0,0,0,0,0,0,0,0,0,2,1,1,2},
    {0,0,0,2,0,0,0,1,0,0,0,2,0,0,0,1}, {0,2,2,2,1,2,2,2,0,2,2,2,1,2,2,2}, {0,1,0,1,2,2,2,2,2,2,2,2,2,2,2,2},
    {0,1,1,1,2,0,1,1,2,2,0,1,2,2,2,0},
};
// Anchor texel of subset 1 (two subsets), of subsets 1 and 2 (three subsets).
constexpr uint8_t kAnchor2[64] = {15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15,
                                  15, 2,  8,  2,  2,  8,  8,  15, 2,  8,  2,  2,  8,  8,  2,  2,
                                  15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2,  2,  2,  15, 15, 6,
                                  6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15};
constexpr uint8_t kAnchor3a[64] = {3,  3,  15, 15, 8,  3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,
                                   3,  3,  8,  15, 3,  3,  6,  10, 5,  8,  8,  6,  8,  5,  15, 15,
                                   8,  15, 3,  5,  6,  10, 8,  15, 15, 3,  15, 5,  15, 15, 15, 15,
                                   3,  15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3};
constexpr uint8_t kAnchor3b[64] = {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,
                                   15, 8,  15, 3,  15, 8,  15, 8,  3,  15, 6,  10, 15, 15, 10, 8,
                                   15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15, 3,  6,  6,  8,
                                   15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8};

struct Bits {
    const uint8_t* p;
    uint32_t pos = 0;
    uint32_t read(int n) {
        uint32_t v = 0;
        for (int k = 0; k < n; k++, pos++) v |= uint32_t((p[pos >> 3] >> (pos & 7)) & 1u) << k;
        return v;
    }
};

inline uint32_t widen(uint32_t v, int bits) { return (v << (8 - bits)) | (v >> (2 * bits - 8)); }
inline uint32_t interp(uint32_t a, uint32_t b, int w) { return ((64 - w) * a + w * b + 32) >> 6; }
inline int weight(int bits, uint32_t i) { return bits == 2 ? kW2[i] : bits == 3 ? kW3[i] : kW4[i]; }

// All 16 texels of one block, row order.
inline void decode(const uint8_t* blk, RGBA8 out[16]) {
    int mode = 0;
    while (mode < 8 && !((blk[0] >> mode) & 1)) mode++;
    if (mode == 8) {
        for (int t = 0; t < 16; t++) out[t] = RGBA8{{0, 0, 0, 0}};
        return;
    }
    const Mode& m = kModes[mode];
    Bits b{blk};
    b.read(mode + 1);
    const uint32_t part = b.read(m.pb), rot = b.read(m.rb), isb = b.read(m.isb);
    const int ne = 2 * m.ns;
    uint32_t ep[6][4] = {};
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < ne; e++) ep[e][c] = b.read(m.cb);
    for (int e = 0; e < ne; e++) ep[e][3] = m.ab ? b.read(m.ab) : 255u;
    uint32_t pbit[6] = {};
    if (m.epb)
        for (int e = 0; e < ne; e++) pbit[e] = b.read(1);
    if (m.spb)
        for (int s = 0; s < m.ns; s++) pbit[2 * s] = pbit[2 * s + 1] = b.read(1);
    for (int e = 0; e < ne; e++)
        for (int c = 0; c < 4; c++) {
            if (c == 3 && !m.ab) continue;
            int bits = c < 3 ? m.cb : m.ab;
            uint32_t v = ep[e][c];
            if (m.epb || m.spb) {
                v = (v << 1) | pbit[e];
                bits++;
            }
            ep[e][c] = widen(v, bits);
        }
    auto subset = [&](int t) { return m.ns == 1 ? 0 : m.ns == 2 ? int(kPart2[part][t]) : int(kPart3[part][t]); };
    auto anchor = [&](int t) {
        if (t == 0) return true;
        if (m.ns == 2) return t == kAnchor2[part];
        if (m.ns == 3) return t == kAnchor3a[part] || t == kAnchor3b[part];
        return false;
    };
    uint32_t idx[16], idx2[16] = {};
    for (int t = 0; t < 16; t++) idx[t] = b.read(m.ib - (anchor(t) ? 1 : 0));
    if (m.ib2)
        for (int t = 0; t < 16; t++) idx2[t] = b.read(m.ib2 - (t == 0 ? 1 : 0));
    for (int t = 0; t < 16; t++) {
        const int s = subset(t);
        const uint32_t* e0 = ep[2 * s];
        const uint32_t* e1 = ep[2 * s + 1];
        int cbits = m.ib, abits = m.ib;
        uint32_t ci = idx[t], ai = idx[t];
        if (m.ib2) {
            if (isb) {
                cbits = m.ib2, ci = idx2[t];
            } else {
                abits = m.ib2, ai = idx2[t];
            }
        }
        uint32_t px[4];
        for (int c = 0; c < 3; c++) px[c] = interp(e0[c], e1[c], weight(cbits, ci));
        px[3] = m.ab ? interp(e0[3], e1[3], weight(abits, ai)) : 255u;
        if (rot) std::swap(px[3], px[rot - 1]);
        for (int c = 0; c < 4; c++) out[t].c[c] = uint8_t(px[c]);
    }
}
}  // namespace bc7

// Bytes per block of a block-compressed format (0: uncompressed RGBA8).
inline uint32_t block_bytes(uint32_t format) {
    const uint32_t f = format & LH_TEXTURE_FORMAT_MASK;
    return f == LH_TEXTURE_BC1 ? 8u : f == LH_TEXTURE_BC7 ? 16u : 0u;
}

// Texel (i, j), already wrapped, of a level of width w whose data starts at
// `base` (the texture's level-0 address).
inline RGBA8 fetch(const LhTexture& t, const uint8_t* base, uint32_t level, uint32_t w, uint32_t i, uint32_t j) {
    RGBA8 o;
    const uint32_t f = t.format & LH_TEXTURE_FORMAT_MASK;
    if (f == LH_TEXTURE_RGBA8) {
        std::memcpy(o.c, base + 4 * (size_t(t.level_offset[level]) + size_t(j) * w + i), 4);
        return o;
    }
    if (f == LH_TEXTURE_BC7) {
        RGBA8 all[16];
        bc7::decode(base + 16 * (size_t(t.level_offset[level]) + size_t(j >> 2) * ((w + 3) >> 2) + (i >> 2)), all);
        return all[4 * (j & 3) + (i & 3)];
    }
    const uint8_t* blk = base + 8 * (size_t(t.level_offset[level]) + size_t(j >> 2) * ((w + 3) >> 2) + (i >> 2));
    uint32_t c0 = uint32_t(blk[0]) | (uint32_t(blk[1]) << 8), c1 = uint32_t(blk[2]) | (uint32_t(blk[3]) << 8);
    uint32_t idx = uint32_t(blk[4]) | (uint32_t(blk[5]) << 8) | (uint32_t(blk[6]) << 16) | (uint32_t(blk[7]) << 24);
    RGBA8 c = bc1(c0, c1, (idx >> (2 * (4 * (j & 3) + (i & 3)))) & 3u);
    if (t.format & LH_TEXTURE_RGB) c.c[3] = 255;
    return c;
}

// sRGB EOTF of an 8-bit encoded channel (the decode a sampler applies to
// *_SRGB formats before filtering): c / 12.92 up to 0.04045, else
// ((c + 0.055) / 1.055)^2.4.
template <class T>
inline T srgb_to_linear(uint8_t v) {
    const T c = T(v) / T(255);
    return c <= T(0.04045) ? c / T(12.92) : std::pow((c + T(0.055)) / T(1.055), T(2.4));
}

}  // namespace orc_texel
