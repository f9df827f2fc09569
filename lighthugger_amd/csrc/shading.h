// Per-pixel shading of the visibility-buffer resolve (device code), shared by
// the forward resolve and its adjoint.
//
// Follows src/shaders/render_geometry.comp:70-266 with CalcFullBary and the
// interpolation helpers of src/shaders/common/vbuffer.glsl:11-119 and the
// BRDF of src/shaders/common/pbr.glsl:1-96.
#pragma once

#include "lh_device.h"

namespace shading {

using namespace lhd;

constexpr float kPi = 3.14159265358979323846f;
// Sky radiance without a skybox cubemap (lh_view.skybox NULL): a constant
// colour times the 2^3 exposure of render_geometry.comp:93-97.
constexpr float kSkyR = 0.35f * 8.0f, kSkyG = 0.50f * 8.0f, kSkyB = 0.80f * 8.0f;

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
LH_DEV f3 load_normal(const MeshInfo& mi, uint32_t v) {
    if (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) {
        const float* q = ptr<float>(mi.normals) + 3 * size_t(v);
        return {q[0], q[1], q[2]};
    }
    const int8_t* q = ptr<int8_t>(mi.normals) + 3 * size_t(v);
    return {float(q[0]), float(q[1]), float(q[2])};
}
LH_DEV f2 load_uv(const MeshInfo& mi, uint32_t v) {
    if (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) {
        const float* q = ptr<float>(mi.uvs) + 2 * size_t(v);
        return {q[0], q[1]};
    }
    const uint16_t* q = ptr<uint16_t>(mi.uvs) + 2 * size_t(v);
    return {float(q[0]), float(q[1])};
}

struct BaryDeriv {
    f3 lambda, ddx, ddy;
};

// CalcFullBary (src/shaders/common/vbuffer.glsl:11-63).
LH_DEV BaryDeriv calc_full_bary(f4 p0, f4 p1, f4 p2, f2 ndc, f2 win) {
    BaryDeriv r;
    const f3 invW{1.0f / p0.w, 1.0f / p1.w, 1.0f / p2.w};
    const f2 n0{p0.x * invW.x, p0.y * invW.x}, n1{p1.x * invW.y, p1.y * invW.y}, n2{p2.x * invW.z, p2.y * invW.z};
    const f2 c0{n2.x - n1.x, n2.y - n1.y}, c1{n0.x - n1.x, n0.y - n1.y};
    const float invDet = 1.0f / (c0.x * c1.y - c1.x * c0.y);
    r.ddx = mul(mul(f3{n1.y - n2.y, n2.y - n0.y, n0.y - n1.y}, invDet), invW);
    r.ddy = mul(mul(f3{n2.x - n1.x, n0.x - n2.x, n1.x - n0.x}, invDet), invW);
    float ddxSum = (r.ddx.x + r.ddx.y) + r.ddx.z;
    float ddySum = (r.ddy.x + r.ddy.y) + r.ddy.z;
    const f2 dv{ndc.x - n0.x, ndc.y - n0.y};
    const float interpInvW = (invW.x + dv.x * ddxSum) + dv.y * ddySum;
    const float interpW = 1.0f / interpInvW;
    r.lambda.x = interpW * ((invW.x + dv.x * r.ddx.x) + dv.y * r.ddy.x);
    r.lambda.y = interpW * ((0.0f + dv.x * r.ddx.y) + dv.y * r.ddy.y);
    r.lambda.z = interpW * ((0.0f + dv.x * r.ddx.z) + dv.y * r.ddy.z);
    const float sx = 2.0f / win.x, sy = 2.0f / win.y;
    r.ddx = mul(r.ddx, sx);
    r.ddy = mul(r.ddy, sy);
    ddxSum *= sx;
    ddySum *= sy;
    r.ddy = mul(r.ddy, -1.0f);
    ddySum *= -1.0f;
    const float wdx = 1.0f / (interpInvW + ddxSum), wdy = 1.0f / (interpInvW + ddySum);
    r.ddx = f3{wdx * (r.lambda.x * interpInvW + r.ddx.x) - r.lambda.x,
               wdx * (r.lambda.y * interpInvW + r.ddx.y) - r.lambda.y,
               wdx * (r.lambda.z * interpInvW + r.ddx.z) - r.lambda.z};
    r.ddy = f3{wdy * (r.lambda.x * interpInvW + r.ddy.x) - r.lambda.x,
               wdy * (r.lambda.y * interpInvW + r.ddy.y) - r.lambda.y,
               wdy * (r.lambda.z * interpInvW + r.ddy.z) - r.lambda.z};
    return r;
}

LH_DEV f3 interp1(const BaryDeriv& d, float a, float b, float c) {
    const f3 m{a, b, c};
    return {dot3(m, d.lambda), dot3(m, d.ddx), dot3(m, d.ddy)};
}
struct Interp3 {
    f3 v, dx, dy;
};
struct Interp2 {
    f2 v, dx, dy;
};
LH_DEV Interp3 interp3(const BaryDeriv& d, f3 a, f3 b, f3 c) {
    const f3 x = interp1(d, a.x, b.x, c.x), y = interp1(d, a.y, b.y, c.y), z = interp1(d, a.z, b.z, c.z);
    return {{x.x, y.x, z.x}, {x.y, y.y, z.y}, {x.z, y.z, z.z}};
}
LH_DEV Interp2 interp2(const BaryDeriv& d, f2 a, f2 b, f2 c) {
    const f3 x = interp1(d, a.x, b.x, c.x), y = interp1(d, a.y, b.y, c.y);
    return {{x.x, y.x}, {x.y, y.y}, {x.z, y.z}};
}

LH_DEV f4 mix4(f4 a, f4 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t)}; }

// BC1 (S3TC) colour of index `sel` of a block with RGB565 endpoints c0, c1:
// endpoints widened by bit replication; c0 > c1: c0, c1, (2 c0 + c1 + 1) / 3,
// (c0 + 2 c1 + 1) / 3, else c0, c1, (c0 + c1 + 1) / 2 and transparent black.
// Packed RGBA8, red in the low byte.
LH_DEV uint32_t widen565(uint32_t c) {
    const uint32_t r = (c >> 11) & 31u, g = (c >> 5) & 63u, b = c & 31u;
    return ((r << 3) | (r >> 2)) | (((g << 2) | (g >> 4)) << 8) | (((b << 3) | (b >> 2)) << 16);
}
LH_DEV uint32_t bc1_colour(uint32_t c0, uint32_t c1, uint32_t sel) {
    const uint32_t a = widen565(c0), b = widen565(c1);
    if (sel == 0) return a | 0xFF000000u;
    if (sel == 1) return b | 0xFF000000u;
    if (c0 <= c1 && sel == 3) return 0u;
    uint32_t out = 0xFF000000u;
    for (uint32_t k = 0; k < 24; k += 8) {
        const uint32_t x = (a >> k) & 255u, y = (b >> k) & 255u;
        const uint32_t m = c0 > c1 ? (sel == 2 ? (2 * x + y + 1) / 3 : (x + 2 * y + 1) / 3) : (x + y + 1) / 2;
        out |= m << k;
    }
    return out;
}
// BC7 (BPTC) texel `t` (row order, x fastest) of one 16-byte block, packed
// RGBA8.  Random access into the block: the mode (unary prefix of the first
// byte) fixes the field widths, so the texel's subset, its two endpoints,
// their p-bits and its index(es) are read at computed bit offsets without
// decoding the rest of the block.  Endpoints with a p-bit append it and are
// widened to 8 bits by replicating their top bits; the texel is
// ((64 - w) e0 + w e1 + 32) >> 6 with the 2-, 3- or 4-bit weights; rotation
// swaps alpha with R, G or B.  A block without a mode bit in its first byte
// is transparent black.
namespace bc7 {
// per mode: subsets, partition bits, rotation bits, index-selection bits,
// colour bits, alpha bits, per-endpoint p-bits, shared p-bits, index bits,
// second index bits - packed 4 bits each, lowest first
static __constant__ uint64_t kModes[8] = {
    0x0000000301040043ull, 0x0000000310060062ull, 0x0000000200050063ull, 0x0000000201070062ull,
    0x0000003200651201ull, 0x0000002200870201ull, 0x0000000401770001ull, 0x0000000201550062ull,
};
// two-subset partitions: bit t of entry p = subset of texel t
static __constant__ uint16_t kPart2[64] = {
    0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8,
    0xFF00, 0xFFF0, 0xF000, 0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110,
    0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C, 0xAAAA, 0xF0F0, 0x5A5A, 0x33CC, 0x3C3C, 0x55AA, 0x9696,
    0xA55A, 0x73CE, 0x13C8, 0x324C, 0x3BDC, 0x6996, 0xC33C, 0x9966, 0x0660, 0x0272, 0x04E4, 0x4E40, 0x2720,
    0xC936, 0x936C, 0x39C6, 0x639C, 0x9336, 0x9CC6, 0x817E, 0xE718, 0xCCF0, 0x0FCC, 0x7744, 0xEE22,
};
// three-subset partitions: bits 2t..2t+1 of entry p = subset of texel t
static __constant__ uint32_t kPart3[64] = {
    0xAA685050, 0x6A5A5040, 0x5A5A4200, 0x5450A0A8, 0xA5A50000, 0xA0A05050, 0x5555A0A0, 0x5A5A5050,
    0xAA550000, 0xAA555500, 0xAAAA5500, 0x90909090, 0x94949494, 0xA4A4A4A4, 0xA9A59450, 0x2A0A4250,
    0xA5945040, 0x0A425054, 0xA5A5A500, 0x55A0A0A0, 0xA8A85454, 0x6A6A4040, 0xA4A45000, 0x1A1A0500,
    0x0050A4A4, 0xAAA59090, 0x14696914, 0x69691400, 0xA08585A0, 0xAA821414, 0x50A4A450, 0x6A5A0200,
    0xA9A58000, 0x5090A0A8, 0xA8A09050, 0x24242424, 0x00AA5500, 0x24924924, 0x24499224, 0x50A50A50,
    0x500AA550, 0xAAAA4444, 0x66660000, 0xA5A0A5A0, 0x50A050A0, 0x69286928, 0x44AAAA44, 0x66666600,
    0xAA444444, 0x54A854A8, 0x95809580, 0x96969600, 0xA85454A8, 0x80959580, 0xAA141414, 0x96960000,
    0xAAAA1414, 0xA05050A0, 0xA0A5A5A0, 0x96000000, 0x40804080, 0xA9A8A9A8, 0xAAAAAA44, 0x2A4A5254,
};
// anchor texels: subset 1 of two; subsets 1 and 2 of three
static __constant__ uint8_t kAnchor2[64] = {15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15,
                                            15, 2,  8,  2,  2,  8,  8,  15, 2,  8,  2,  2,  8,  8,  2,  2,
                                            15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2,  2,  2,  15, 15, 6,
                                            6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15};
static __constant__ uint8_t kAnchor3a[64] = {3,  3,  15, 15, 8,  3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,
                                             3,  3,  8,  15, 3,  3,  6,  10, 5,  8,  8,  6,  8,  5,  15, 15,
                                             8,  15, 3,  5,  6,  10, 8,  15, 15, 3,  15, 5,  15, 15, 15, 15,
                                             3,  15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3};
static __constant__ uint8_t kAnchor3b[64] = {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,
                                             15, 8,  15, 3,  15, 8,  15, 8,  3,  15, 6,  10, 15, 15, 10, 8,
                                             15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15, 3,  6,  6,  8,
                                             15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8};
// interpolation weights of 2-, 3- and 4-bit indices: kWeights[offset + index], offsets 0, 4, 12
static __constant__ uint8_t kWeights[28] = {0,  21, 43, 64, 0,  9,  18, 27, 37, 46, 55, 64, 0,  4,
                                            9,  13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};

struct Block {
    uint64_t lo, hi;
    // n <= 32 bits at bit position pos (< 128)
    LH_DEV uint32_t bits(uint32_t pos, uint32_t n) const {
        const uint64_t v = pos >= 64 ? hi >> (pos - 64) : pos == 0 ? lo : (lo >> pos) | (hi << (64 - pos));
        return uint32_t(v) & uint32_t((1ull << n) - 1);
    }
};
LH_DEV uint32_t widen(uint32_t v, uint32_t bits) { return (v << (8 - bits)) | (v >> (2 * bits - 8)); }
LH_DEV uint32_t interp(uint32_t a, uint32_t b, uint32_t w) { return ((64 - w) * a + w * b + 32) >> 6; }
LH_DEV uint32_t weight(uint32_t bits, uint32_t i) { return kWeights[(bits == 2 ? 0u : bits == 3 ? 4u : 12u) + i]; }

LH_DEV uint32_t texel(uint4 q, uint32_t t) {
    const uint32_t first = q.x & 255u;
    if (first == 0) return 0u;
    const uint32_t mode = uint32_t(__builtin_ctz(first));
    const uint64_t mi = kModes[mode];
    auto f = [&](int k) { return uint32_t(mi >> (4 * k)) & 15u; };
    const uint32_t ns = f(0), pb = f(1), rb = f(2), isb = f(3), cb = f(4), ab = f(5), epb = f(6), spb = f(7),
                   ib = f(8), ib2 = f(9);
    const Block b{uint64_t(q.x) | (uint64_t(q.y) << 32), uint64_t(q.z) | (uint64_t(q.w) << 32)};
    uint32_t pos = mode + 1;
    const uint32_t part = b.bits(pos, pb);
    pos += pb;
    const uint32_t rot = b.bits(pos, rb);
    pos += rb;
    const uint32_t sel = b.bits(pos, isb);
    pos += isb;
    const uint32_t s = ns == 1 ? 0u : ns == 2 ? (kPart2[part] >> t) & 1u : (kPart3[part] >> (2 * t)) & 3u;
    const uint32_t ne = 2 * ns, k0 = 2 * s, k1 = 2 * s + 1;
    uint32_t e0[4], e1[4];
    for (uint32_t c = 0; c < 3; c++) {
        e0[c] = b.bits(pos + (c * ne + k0) * cb, cb);
        e1[c] = b.bits(pos + (c * ne + k1) * cb, cb);
    }
    pos += 3 * ne * cb;
    e0[3] = e1[3] = 255u;
    if (ab) {
        e0[3] = b.bits(pos + k0 * ab, ab);
        e1[3] = b.bits(pos + k1 * ab, ab);
        pos += ne * ab;
    }
    uint32_t p0 = 0, p1 = 0, extra = 0;
    if (epb) {
        p0 = b.bits(pos + k0, 1);
        p1 = b.bits(pos + k1, 1);
        pos += ne;
        extra = 1;
    } else if (spb) {
        p0 = p1 = b.bits(pos + s, 1);
        pos += ns;
        extra = 1;
    }
    for (uint32_t c = 0; c < 4; c++) {
        if (c == 3 && !ab) continue;
        const uint32_t n = (c < 3 ? cb : ab) + extra;
        e0[c] = widen(extra ? (e0[c] << 1) | p0 : e0[c], n);
        e1[c] = widen(extra ? (e1[c] << 1) | p1 : e1[c], n);
    }
    // primary indices: the anchor texel of each subset stores ib - 1 bits
    const uint32_t a1 = ns == 2 ? kAnchor2[part] : ns == 3 ? kAnchor3a[part] : 16u;
    const uint32_t a2 = ns == 3 ? kAnchor3b[part] : 16u;
    const uint32_t off = t * ib - (t > 0 ? 1u : 0u) - (t > a1 ? 1u : 0u) - (t > a2 ? 1u : 0u);
    const bool anchor = t == 0 || t == a1 || t == a2;
    const uint32_t i1 = b.bits(pos + off, ib - (anchor ? 1u : 0u));
    uint32_t ci = i1, cbits = ib, ai = i1, abits = ib;
    if (ib2) {  // modes 4 and 5: a second index plane, anchor texel 0 only
        const uint32_t i2 = b.bits(pos + 16 * ib - ns + t * ib2 - (t > 0 ? 1u : 0u), ib2 - (t == 0 ? 1u : 0u));
        if (sel) {
            ci = i2, cbits = ib2;
        } else {
            ai = i2, abits = ib2;
        }
    }
    uint32_t px[4];
    for (uint32_t c = 0; c < 3; c++) px[c] = interp(e0[c], e1[c], weight(cbits, ci));
    px[3] = ab ? interp(e0[3], e1[3], weight(abits, ai)) : 255u;
    if (rot) {
        const uint32_t tmp = px[3];
        px[3] = px[rot - 1];
        px[rot - 1] = tmp;
    }
    return px[0] | (px[1] << 8) | (px[2] << 16) | (px[3] << 24);
}
}  // namespace bc7

// Texel (i, j) of a level of width w (already wrapped), packed RGBA8.
LH_DEV uint32_t texel_rgba8(const LhTexture& t, uint32_t level, uint32_t w, uint32_t i, uint32_t j) {
    const uint32_t fmt = t.format & LH_TEXTURE_FORMAT_MASK;
    const size_t block = size_t(t.level_offset[level]) + size_t(j >> 2) * ((w + 3) >> 2) + (i >> 2);
    if (fmt == LH_TEXTURE_BC1) {
        const uint2 blk = ptr<uint2>(t.texels)[block];
        return bc1_colour(blk.x & 0xFFFFu, blk.x >> 16, (blk.y >> (2 * (4 * (j & 3) + (i & 3)))) & 3u) |
               ((t.format & LH_TEXTURE_RGB) ? 0xFF000000u : 0u);
    }
    if (fmt == LH_TEXTURE_BC7) return bc7::texel(ptr<uint4>(t.texels)[block], 4 * (j & 3) + (i & 3));
    return ptr<uint32_t>(t.texels)[size_t(t.level_offset[level]) + size_t(j) * w + i];
}

// v / 255 for an 8-bit v, correctly rounded: the product with the rounded
// reciprocal and one FMA correction equal the IEEE quotient for each of the
// 256 values (checked exhaustively in exact arithmetic), at three
// instructions instead of a full division.
LH_DEV float unorm8(uint32_t v) {
    const float x = float(v), r = 1.0f / 255.0f;
    const float q = x * r;
    return __builtin_fmaf(__builtin_fmaf(-q, 255.0f, x), r, q);
}

// sRGB EOTF of an 8-bit channel: c / 12.92 up to 0.04045, else ((c + 0.055) / 1.055)^2.4.
LH_DEV float srgb_to_linear(uint32_t v) {
    const float c = unorm8(v);
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}
// A packed RGBA8 texel as floats in [0, 1]; sRGB formats decode RGB to linear.
LH_DEV f4 unpack_texel(const LhTexture& t, uint32_t p) {
    if (t.format & LH_TEXTURE_SRGB)
        return {srgb_to_linear(p & 255u), srgb_to_linear((p >> 8) & 255u), srgb_to_linear((p >> 16) & 255u),
                unorm8(p >> 24)};
    return {unorm8(p & 255u), unorm8((p >> 8) & 255u), unorm8((p >> 16) & 255u), unorm8(p >> 24)};
}

// i modulo n in [0, n) (repeat wrapping); indices already inside skip the
// division.
LH_DEV uint32_t wrap_index(int32_t i, int32_t n) {
    if (uint32_t(i) < uint32_t(n)) return uint32_t(i);
    i %= n;
    return uint32_t(i < 0 ? i + n : i);
}
// The 2x2 texels of a bilinear footprint with corner (i, j) at one level,
// wrapped once: (i + 1) mod w is the next index after i mod w.
struct TexelQuad {
    f4 t00, t10, t01, t11;
};
LH_DEV TexelQuad texel_quad(const LhTexture& t, uint32_t level, int32_t w, int32_t h, int32_t i, int32_t j) {
    const uint32_t i0 = wrap_index(i, w), j0 = wrap_index(j, h);
    const uint32_t i1 = i0 + 1 == uint32_t(w) ? 0u : i0 + 1, j1 = j0 + 1 == uint32_t(h) ? 0u : j0 + 1;
    const uint32_t W = uint32_t(w);
    return {unpack_texel(t, texel_rgba8(t, level, W, i0, j0)), unpack_texel(t, texel_rgba8(t, level, W, i1, j0)),
            unpack_texel(t, texel_rgba8(t, level, W, i0, j1)), unpack_texel(t, texel_rgba8(t, level, W, i1, j1))};
}
LH_DEV f4 bilinear(const LhTexture& t, uint32_t level, f2 uv) {
    const int32_t wi = int32_t(max(1u, t.width >> level)), hi = int32_t(max(1u, t.height >> level));
    const float w = float(wi), h = float(hi);
    const float u = uv.x * w - 0.5f, v = uv.y * h - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const TexelQuad q = texel_quad(t, level, wi, hi, int32_t(fu), int32_t(fv));
    return mix4(mix4(q.t00, q.t10, a), mix4(q.t01, q.t11, a), b);
}

// Mip selection of the isotropic trilinear filter: levels l0, l1 and the
// blend f between them, from the analytic uv derivatives.
struct Lod {
    uint32_t l0, l1;
    float f;
};
LH_DEV Lod texture_lod(const LhTexture& t, const Interp2& uv) {
    const float W = float(t.width), H = float(t.height);
    const float ax = uv.dx.x * W, ay = uv.dx.y * H, bx = uv.dy.x * W, by = uv.dy.y * H;
    const float rho = fmaxf(sqrtf(ax * ax + ay * ay), sqrtf(bx * bx + by * by));
    float lambda = rho > 0.0f ? log2f(rho) : 0.0f;
    lambda = clampf(lambda, 0.0f, float(t.levels - 1));
    const float d0 = floorf(lambda);
    const uint32_t l0 = uint32_t(d0);
    return {l0, min(l0 + 1, t.levels - 1), lambda - d0};
}

LH_DEV bool has_texture(const lh_scene& s, uint32_t index) {
    return index != UNUSED_TEXTURE_INDEX && index < s.num_textures;
}

// sample_texture (render_geometry.comp:59-66) as isotropic trilinear
// filtering with the LOD from the analytic uv derivatives.  An unused index
// reads as opaque white.
LH_DEV f4 sample_texture_lod(const lh_scene& s, uint32_t index, const Lod& lod, const Interp2& uv) {
    if (!has_texture(s, index)) return {1, 1, 1, 1};
    const LhTexture& t = s.textures[index];
    const f4 c0 = bilinear(t, lod.l0, uv.v);
    if (lod.f == 0.0f) return c0;
    return mix4(c0, bilinear(t, lod.l1, uv.v), lod.f);
}
LH_DEV Lod lod_of(const lh_scene& s, uint32_t index, const Interp2& uv) {
    return has_texture(s, index) ? texture_lod(s.textures[index], uv) : Lod{0, 0, 0.0f};
}
LH_DEV f4 sample_texture(const lh_scene& s, uint32_t index, const Interp2& uv) {
    return sample_texture_lod(s, index, lod_of(s, index, uv), uv);
}

// Per-pixel record of the quantities the adjoint holds fixed
// (lh_resolve_params.frozen): shadow factor, LOD blends of the base-colour
// and metallic-roughness textures, and a word with the flip (bit 0), their
// base levels (bits 8-15, 16-23) and a valid bit (31).
LH_DEV uint4 pack_frozen(float shadow_sum, bool flip, const Lod& b, const Lod& m) {
    return make_uint4(__float_as_uint(shadow_sum), __float_as_uint(b.f), __float_as_uint(m.f),
                      (flip ? 1u : 0u) | (b.l0 << 8) | (m.l0 << 16) | 0x80000000u);
}

// sampler2DArrayShadow tap: linear filter, clamp to edge, compare LESS.
LH_DEV float shadow_tap(const lh_view& v, uint32_t layer, f2 uv, float ref) {
    const uint32_t S = v.shadow_size;
    const uint32_t* map = v.shadow_maps + size_t(layer) * S * S;
    const float u = uv.x * float(S) - 0.5f, w = uv.y * float(S) - 0.5f;
    const float fu = floorf(u), fw = floorf(w);
    const float a = u - fu, b = w - fw;
    const int32_t i = int32_t(fu), j = int32_t(fw);
    auto cmp = [&](int32_t x, int32_t y) {
        x = min(max(x, 0), int32_t(S) - 1);
        y = min(max(y, 0), int32_t(S) - 1);
        return ref < __uint_as_float(map[size_t(y) * S + x]) ? 1.0f : 0.0f;
    };
    return mixf(mixf(cmp(i, j), cmp(i + 1, j), a), mixf(cmp(i, j + 1), cmp(i + 1, j + 1), a), b);
}

LH_DEV float pow5(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}
LH_DEV float F_Schlick1(float u, float f0, float f90) { return f0 + (f90 - f0) * pow5(1.0f - u); }
LH_DEV float Fd_Burley(float NoV, float NoL, float LoH, float roughness) {
    const float f90 = 0.5f + ((2.0f * roughness) * LoH) * LoH;
    return (F_Schlick1(NoL, 1.0f, f90) * F_Schlick1(NoV, 1.0f, f90)) * (1.0f / kPi);
}
LH_DEV float D_GGX(float NoH, float a) {
    const float a2 = a * a;
    const float f = (NoH * a2 - NoH) * NoH + 1.0f;
    return a2 / ((kPi * f) * f);
}
LH_DEV float V_SmithGGXCorrelated(float NoV, float NoL, float a) {
    const float a2 = a * a;
    const float GGXL = NoV * sqrtf((-NoL * a2 + NoL) * NoL + a2);
    const float GGXV = NoL * sqrtf((-NoV * a2 + NoV) * NoV + a2);
    return 0.5f / (GGXV + GGXL);
}

// BRDF (src/shaders/common/pbr.glsl:44-96).
LH_DEV f3 brdf(f3 v, f3 l, f3 n, float pr, float metallic, f3 base, bool thin) {
    const f3 h = normalize3(add(v, l));
    const f3 diffuse = mul(base, 1.0f - metallic);
    const float refl = ((1.5f - 1.0f) / (1.5f + 1.0f)) * ((1.5f - 1.0f) / (1.5f + 1.0f));
    const float d0 = refl * (1.0f - metallic);
    const f3 f0 = add(f3{d0, d0, d0}, mul(base, metallic));
    const float NoV = clampf(dot3(n, v), 1e-5f, 1.0f);
    const float NoL = clampf(dot3(n, l), 0.0f, 1.0f);
    const float NoH = clampf(dot3(n, h), 0.0f, 1.0f);
    const float LoH = clampf(dot3(l, h), 0.0f, 1.0f);
    const float a = pr * pr;
    const float D = D_GGX(NoH, a);
    const float p = pow5(1.0f - LoH);
    const f3 F{f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
    const float V = V_SmithGGXCorrelated(NoV, NoL, a);
    const f3 Fr = mul(F, D * V);
    const f3 Fd = mul(diffuse, Fd_Burley(NoV, NoL, LoH, a));
    f3 irr = mul(add(Fd, Fr), NoL);
    irr = f3{fmaxf(irr.x, 0.0f), fmaxf(irr.y, 0.0f), fmaxf(irr.z, 0.0f)};
    if (thin) {
        const f3 nn = neg(n);
        const float rNoV = clampf(dot3(nn, v), 1e-5f, 1.0f), rNoL = clampf(dot3(nn, l), 0.0f, 1.0f);
        irr = add(irr, mul(diffuse, Fd_Burley(rNoV, rNoL, LoH, a) * rNoL));
    }
    return {isnan(irr.x) ? 0.0f : irr.x, isnan(irr.y) ? 0.0f : irr.y, isnan(irr.z) ? 0.0f : irr.z};
}

// Decoded geometry of one visibility-buffer sample.
struct Surface {
    uint32_t instance, mesh_info, vi[3];
    f3 local[3];  // mesh-space positions
    f3 pos[3];    // world positions
    f4 clip[3];
};

LH_DEV bool decode(const lh_scene& s, const lh_view& v, uint64_t key, Surface& sf) {
    if (~uint32_t(key >> 32) == 0u) return false;  // depth 0.0: sky
    const uint32_t mid = uint32_t(key >> LH_KEY_TRI_BITS) & (LH_MAX_SCENE_MESHLETS - 1);
    const uint32_t tri = uint32_t(key) & ((1u << LH_KEY_TRI_BITS) - 1);
    const MeshletReference ref = s.meshlet_table[mid];
    const Instance& in = s.instances[ref.instance_index];
    const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
    const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
    const uint8_t* micro = ptr<uint8_t>(mi.micro_indices) + ml.triangle_offset + 3 * tri;
    sf.instance = ref.instance_index;
    sf.mesh_info = uint32_t((in.mesh_info_address - uint64_t(s.mesh_infos)) / sizeof(MeshInfo));
    for (int j = 0; j < 3; j++) {
        sf.vi[j] = load_index(mi, ml.index_offset + micro[j]);
        sf.local[j] = load_position(mi, sf.vi[j]);
        sf.pos[j] = mat4_point(in.transform, sf.local[j]);
        sf.clip[j] = mat4_mul_vec4(v.uniforms->combined_perspective_view, f4{sf.pos[j].x, sf.pos[j].y, sf.pos[j].z, 1.0f});
    }
    return true;
}

// Alpha test of the alpha-clip pipelines (visbuffer_alpha_clip_pixel and
// shadowmap_alpha_clipped_pixel, src/shaders/rasterization.glsl:84-97 and
// :130-145): the base-colour texture's alpha at the pixel centre, discarded
// below 0.5.  The uv is interpolated perspective-correctly and its LOD taken
// from the analytic derivatives (CalcFullBary, as the resolve does) in place
// of the hardware's 2x2-quad differences.  `m` is the pass's view-projection
// matrix, W x H its target; the triangle is the one the key id names.  A
// triangle with a vertex at or behind w = 0, or a non-finite uv, passes.
LH_DEV bool alpha_visible(const lh_scene& s, const LhMat4& m, uint32_t W, uint32_t H, uint32_t id, int32_t px,
                          int32_t py) {
    const MeshletReference ref = s.meshlet_table[id >> LH_KEY_TRI_BITS];
    const Instance& in = s.instances[ref.instance_index];
    const MeshInfo& mi = mesh_info_at(in.mesh_info_address);
    if (!has_texture(s, mi.base_color_texture_index)) return true;  // white, alpha 1
    const Meshlet& ml = ptr<Meshlet>(mi.meshlets)[ref.meshlet_index];
    const uint8_t* micro = ptr<uint8_t>(mi.micro_indices) + ml.triangle_offset + 3 * (id & ((1u << LH_KEY_TRI_BITS) - 1));
    const f2 ts{mi.texture_scale.x, mi.texture_scale.y};
    f4 clip[3];
    f2 uvs[3];
    for (int j = 0; j < 3; j++) {
        const uint32_t vi = load_index(mi, ml.index_offset + micro[j]);
        const f3 w = mat4_point(in.transform, load_position(mi, vi));
        clip[j] = mat4_mul_vec4(m, f4{w.x, w.y, w.z, 1.0f});
        if (!(clip[j].w > 0.0f)) return true;
        const f2 t = load_uv(mi, vi);
        uvs[j] = {t.x * ts.x, t.y * ts.y};
    }
    const f2 win{float(W), float(H)};
    const f2 ndc{(float(px) + 0.5f) / win.x * 2.0f - 1.0f, (float(py) + 0.5f) / win.y * 2.0f - 1.0f};
    const BaryDeriv bary = calc_full_bary(clip[0], clip[1], clip[2], ndc, win);
    Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);
    uv.v.x += mi.texture_offset.x;
    uv.v.y += mi.texture_offset.y;
    if (!(isfinite(uv.v.x) && isfinite(uv.v.y) && isfinite(uv.dx.x) && isfinite(uv.dx.y) && isfinite(uv.dy.x) &&
          isfinite(uv.dy.y)))
        return true;
    return !(sample_texture(s, mi.base_color_texture_index, uv).w < 0.5f);
}

// Differentiable parameters: each pointer, when set, replaces the value the
// scene or the uniforms hold (lh_resolve_params).
LH_DEV f3 param_base_color(const lh_resolve_params& p, const MeshInfo& mi, uint32_t k) {
    return p.base_color_factor ? f3{p.base_color_factor[3 * k], p.base_color_factor[3 * k + 1],
                                    p.base_color_factor[3 * k + 2]}
                               : mk3(mi.base_color_factor);
}
LH_DEV float param_roughness(const lh_resolve_params& p, const lh_scene& s, uint32_t k) {
    return p.roughness_factor ? p.roughness_factor[k] : s.factors[k].roughness_factor;
}
LH_DEV float param_metallic(const lh_resolve_params& p, const lh_scene& s, uint32_t k) {
    return p.metallic_factor ? p.metallic_factor[k] : s.factors[k].metallic_factor;
}
LH_DEV f3 param_sun_dir(const lh_resolve_params& p, const Uniforms& u) {
    return p.sun_dir ? f3{p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]} : mk3(u.sun_dir);
}
LH_DEV f3 param_sun_intensity(const lh_resolve_params& p, const Uniforms& u) {
    return p.sun_intensity ? f3{p.sun_intensity[0], p.sun_intensity[1], p.sun_intensity[2]} : mk3(u.sun_intensity);
}

// Cascade selection and 3x3 PCF (render_geometry.comp:183-226).
LH_DEV float shadow_factor(const lh_view& v, const MeshInfo& mi, f3 world, uint32_t* cascade_out = nullptr) {
    const MiscStorage& misc = *v.misc;
    uint32_t cascade;
    f4 sc{0, 0, 0, 0};
    for (cascade = 0; cascade < 4; cascade++) {
        sc = mat4_mul_vec4(misc.uv_space_shadow_matrices[cascade], f4{world.x, world.y, world.z, 1.0f});
        if (fabsf(sc.x - 0.5f) < 0.5f && fabsf(sc.y - 0.5f) < 0.5f) break;
    }
    if (cascade_out) *cascade_out = cascade;  // 4: outside every cascade
    const float bias = (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) ? 0.000002f : 0.0f;
    sc = f4{sc.x / sc.w, sc.y / sc.w, sc.z / sc.w, sc.w / sc.w};
    const uint32_t layer = min(cascade, 3u);
    float shadow_sum = 0.0f;
    const float S = float(v.shadow_size);
    const float ref = sc.z - bias;
    // The nine taps' 2x2 cells normally tile the 4x4 texels around the centre
    // tap's cell: those 16 comparisons are made once (a bit each), and a tap
    // uses them when its own cell, computed exactly as shadow_tap computes
    // it, is the expected one; a tap that rounding puts elsewhere reads its
    // own texels.  The sum is the same, bit for bit.
    const int32_t N = int32_t(v.shadow_size);
    const uint32_t* map = v.shadow_maps + size_t(layer) * v.shadow_size * v.shadow_size;
    const int32_t i0 = int32_t(floorf(sc.x * S - 0.5f)), j0 = int32_t(floorf(sc.y * S - 0.5f));
    uint32_t lit = 0;  // bit 4 dy + dx: texel (i0 - 1 + dx, j0 - 1 + dy) passes the compare
#pragma unroll
    for (int dy = 0; dy < 4; dy++) {
        const int32_t y = min(max(j0 - 1 + dy, 0), N - 1);
#pragma unroll
        for (int dx = 0; dx < 4; dx++) {
            const int32_t x = min(max(i0 - 1 + dx, 0), N - 1);
            lit |= (ref < __uint_as_float(map[size_t(y) * uint32_t(N) + uint32_t(x)]) ? 1u : 0u) << (4 * dy + dx);
        }
    }
#pragma unroll
    for (int ox = -1; ox <= 1; ox++)
#pragma unroll
        for (int oy = -1; oy <= 1; oy++) {
            const f2 uv{sc.x + float(ox) / S, sc.y + float(oy) / S};
            const float u = uv.x * S - 0.5f, w = uv.y * S - 0.5f;
            const float fu = floorf(u), fw = floorf(w);
            if (int32_t(fu) == i0 + ox && int32_t(fw) == j0 + oy) {
                const float a = u - fu, b = w - fw;
                const int k = 4 * (oy + 1) + (ox + 1);
                auto bit = [&](int q) { return float((lit >> q) & 1u); };
                shadow_sum += mixf(mixf(bit(k), bit(k + 1), a), mixf(bit(k + 4), bit(k + 5), a), b);
            } else {
                shadow_sum += shadow_tap(v, layer, uv, ref);
            }
        }
    shadow_sum /= 9.0f;
    if (sc.z > 1.0f) shadow_sum = 1.0f;
    return shadow_sum;
}

// Spherical-harmonics ambient term of render_geometry.comp:40-57 with its
// coefficients (only the constant and y bands are non-zero).
LH_DEV float ambient_sh(f3 nrm) {
    const float L = 0.025f;
    return (((((((L + 0.0f * nrm.x) + L * nrm.y) + 0.0f * nrm.z) + 0.0f * (nrm.x * nrm.z)) + 0.0f * (nrm.z * nrm.y)) +
             0.0f * (nrm.y * nrm.x)) +
            0.0f * ((3.0f * nrm.z) * nrm.z - 1.0f)) +
           0.0f * (nrm.x * nrm.x - nrm.y * nrm.y);
}

// Debug colours (src/shaders/common/debug.glsl:1-11).
LH_DEV f3 debug_colour(uint32_t i) {
    switch (i) {
        case 0: return {0.0f / 255.0f, 2.0f / 255.0f, 91.0f / 255.0f};
        case 1: return {0.0f / 255.0f, 108.0f / 255.0f, 251.0f / 255.0f};
        case 2: return {0.0f / 255.0f, 221.0f / 255.0f, 221.0f / 255.0f};
        case 3: return {51.0f / 255.0f, 221.0f / 255.0f, 0.0f / 255.0f};
        case 4: return {255.0f / 255.0f, 252.0f / 255.0f, 0.0f / 255.0f};
        case 5: return {255.0f / 255.0f, 180.0f / 255.0f, 0.0f / 255.0f};
        case 6: return {255.0f / 255.0f, 104.0f / 255.0f, 0.0f / 255.0f};
        case 7: return {226.0f / 255.0f, 22.0f / 255.0f, 0.0f / 255.0f};
        case 8: return {191.0f / 255.0f, 0.0f / 255.0f, 83.0f / 255.0f};
        default: return {145.0f / 255.0f, 0.0f / 255.0f, 65.0f / 255.0f};
    }
}

// Texel (i, j) of skybox face f, clamped to the face's edge.
LH_DEV f3 sky_texel(const lh_view& v, uint32_t f, int32_t i, int32_t j) {
    const int32_t n = int32_t(v.skybox_size);
    i = min(max(i, 0), n - 1);
    j = min(max(j, 0), n - 1);
    const float4 t = reinterpret_cast<const float4*>(v.skybox)[(size_t(f) * uint32_t(n) + uint32_t(j)) * uint32_t(n) +
                                                               uint32_t(i)];
    return {t.x, t.y, t.z};
}
LH_DEV f3 mix3(f3 a, f3 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }

// Skybox radiance of a pixel (render_geometry.comp:86-101): the ray through
// the pixel corner unprojected by perspective_inverse at z = 0 (the infinitely
// far plane of the reverse-Z projection), normalized, turned to world space by
// view_inverse, and looked up at LOD 0 times 2^3.  Face by the major axis
// (ties go to x, then y) with Vulkan's face table (sc, tc, rc):
//   +X (-z, -y, x)  -X (z, -y, x)  +Y (x, z, y)  -Y (x, -z, y)  +Z (x, -y, z)  -Z (-x, -y, z),
// s = (sc / |rc| + 1) / 2, t likewise; bilinear within the face with texel
// centres at (k + 0.5) / size, clamped to the face's edge (Vulkan filters
// seamlessly across faces; the two agree except within half a texel of an edge).
LH_DEV f3 sky_radiance(const lh_view& v, f2 ndc) {
    if (!v.skybox || v.skybox_size == 0) return {kSkyR, kSkyG, kSkyB};
    const Uniforms& u = *v.uniforms;
    const f4 un = mat4_mul_vec4(u.perspective_inverse, f4{ndc.x, ndc.y, 0.0f, 1.0f});
    const f3 local = normalize3(f3{un.x, un.y, un.z});
    const f4 r4 = mat4_mul_vec4(u.view_inverse, f4{local.x, local.y, local.z, 0.0f});
    const float ax = fabsf(r4.x), ay = fabsf(r4.y), az = fabsf(r4.z);
    uint32_t face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        face = r4.x >= 0.0f ? 0u : 1u;
        sc = r4.x >= 0.0f ? -r4.z : r4.z;
        tc = -r4.y;
        ma = ax;
    } else if (ay >= az) {
        face = r4.y >= 0.0f ? 2u : 3u;
        sc = r4.x;
        tc = r4.y >= 0.0f ? r4.z : -r4.z;
        ma = ay;
    } else {
        face = r4.z >= 0.0f ? 4u : 5u;
        sc = r4.z >= 0.0f ? r4.x : -r4.x;
        tc = -r4.y;
        ma = az;
    }
    const float n = float(v.skybox_size);
    const float fs = (sc / ma + 1.0f) * 0.5f, ft = (tc / ma + 1.0f) * 0.5f;
    const float uu = fs * n - 0.5f, vv = ft * n - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv);
    const float a = uu - fu, b = vv - fv;
    const int32_t i = int32_t(fu), j = int32_t(fv);
    const f3 c = mix3(mix3(sky_texel(v, face, i, j), sky_texel(v, face, i + 1, j), a),
                      mix3(sky_texel(v, face, i, j + 1), sky_texel(v, face, i + 1, j + 1), a), b);
    return mul(c, 8.0f);
}

LH_DEV float smoothstepf(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
// Heat-map colour of t in [0, 1] (temperature(), src/shaders/common/debug.glsl:15-43):
// the debug colours blended around index 10 t.
LH_DEV f3 temperature(float t) {
    const float s = t * 10.0f;
    const int32_t cur = int32_t(s) <= 9 ? int32_t(s) : 9;
    const int32_t prv = cur >= 1 ? cur - 1 : 0, nxt = cur < 9 ? cur + 1 : 9;
    const float blur = 0.8f;
    const float wc = smoothstepf(float(cur) - blur, float(cur) + blur, s) *
                     (1.0f - smoothstepf(float(cur + 1) - blur, float(cur + 1) + blur, s));
    const float wp = 1.0f - smoothstepf(float(cur) - blur, float(cur) + blur, s);
    const float wn = smoothstepf(float(cur + 1) - blur, float(cur + 1) + blur, s);
    const f3 a = debug_colour(uint32_t(cur)), b = debug_colour(uint32_t(prv)), c = debug_colour(uint32_t(nxt));
    return {clampf((wc * a.x + wp * b.x) + wn * c.x, 0.0f, 1.0f), clampf((wc * a.y + wp * b.y) + wn * c.y, 0.0f, 1.0f),
            clampf((wc * a.z + wp * b.z) + wn * c.z, 0.0f, 1.0f)};
}
// UNIFORMS_DEBUG_SHADER_CLOCK: device clock ticks spent on the pixel over
// 65000, as a heat-map colour (render_geometry.comp:287-301).
constexpr float kClockHeatScale = 65000.0f;

LH_DEV f3 resolve_pixel(const lh_scene& s, const lh_view& v, uint32_t x, uint32_t y, const lh_resolve_params& prm) {
    const uint64_t start_time = clock64();
    const Uniforms& u = *v.uniforms;
    const f2 win{float(u.window_size.x), float(u.window_size.y)};
    const f2 ndc{float(x) / win.x * 2.0f - 1.0f, float(y) / win.y * 2.0f - 1.0f};  // pixel corner
    Surface sf;
    const uint64_t key = v.visbuffer[size_t(y) * v.width + x];
    if (!decode(s, v, key, sf)) return sky_radiance(v, ndc);
    const Instance& in = s.instances[sf.instance];
    const MeshInfo& mi = s.mesh_infos[sf.mesh_info];
    const BaryDeriv bary = calc_full_bary(sf.clip[0], sf.clip[1], sf.clip[2], ndc, win);
    const Interp3 world = interp3(bary, sf.pos[0], sf.pos[1], sf.pos[2]);
    const f2 ts{mi.texture_scale.x, mi.texture_scale.y};
    f2 uvs[3];
    for (int j = 0; j < 3; j++) {
        const f2 t = load_uv(mi, sf.vi[j]);
        uvs[j] = {t.x * ts.x, t.y * ts.y};
    }
    Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);
    uv.v.x += mi.texture_offset.x;
    uv.v.y += mi.texture_offset.y;
    const f3 view_vector = normalize3(sub(mk3(u.camera_pos), world.v));
    f3 nrm = interp3(bary, normalize3(load_normal(mi, sf.vi[0])), normalize3(load_normal(mi, sf.vi[1])),
                     normalize3(load_normal(mi, sf.vi[2])))
                 .v;
    nrm = normalize3(mat3_mul_vec3(in.normal_transform, nrm));
    const f3 gn = cross3(sub(sf.pos[1], sf.pos[0]), sub(sf.pos[2], sf.pos[0]));
    const bool flip = dot3(gn, view_vector) < 0.0f;
    if (flip) nrm = neg(nrm);
    uint32_t cascade = 0;
    const float shadow_sum = shadow_factor(v, mi, world.v, &cascade);
    const Lod lod_b = lod_of(s, mi.base_color_texture_index, uv), lod_m = lod_of(s, mi.metallic_roughness_texture_index, uv);
    if (prm.frozen)
        reinterpret_cast<uint4*>(prm.frozen)[size_t(y) * v.width + x] = pack_frozen(shadow_sum, flip, lod_b, lod_m);
    const f4 bc = sample_texture_lod(s, mi.base_color_texture_index, lod_b, uv);
    const f3 base = mul(f3{bc.x, bc.y, bc.z}, param_base_color(prm, mi, sf.mesh_info));
    const f4 mr = sample_texture_lod(s, mi.metallic_roughness_texture_index, lod_m, uv);
    const float roughness = mr.y * param_roughness(prm, s, sf.mesh_info);
    const float metallic = mr.z * param_metallic(prm, s, sf.mesh_info);
    if (mi.normal_texture_index != UNUSED_TEXTURE_INDEX) {
        const f4 mn = sample_texture(s, mi.normal_texture_index, uv);
        const f3 map{mn.x * 255.0f / 127.0f - 128.0f / 127.0f, mn.y * 255.0f / 127.0f - 128.0f / 127.0f,
                     mn.z * 255.0f / 127.0f - 128.0f / 127.0f};
        const f3 dpy_perp = cross3(world.dy, nrm), dpx_perp = cross3(nrm, world.dx);
        f3 t = add(mul(dpy_perp, uv.dx.x), mul(dpx_perp, uv.dy.x));
        f3 b = add(mul(dpy_perp, uv.dx.y), mul(dpx_perp, uv.dy.y));
        const float invmax = 1.0f / sqrtf(fmaxf(dot3(t, t), dot3(b, b)));
        t = mul(t, invmax);
        b = mul(b, invmax);
        nrm = normalize3(f3{(t.x * map.x + b.x * map.y) + nrm.x * map.z, (t.y * map.x + b.y * map.y) + nrm.y * map.z,
                            (t.z * map.x + b.z * map.y) + nrm.z * map.z});
    }
    // Debug views (render_geometry.comp:268-308).
    switch (u.debug) {
        case UNIFORMS_DEBUG_CASCADES: return mul(base, debug_colour(cascade));
        case UNIFORMS_DEBUG_TRIANGLE_INDEX: return debug_colour(uint32_t(key & ((1u << LH_KEY_TRI_BITS) - 1)) % 10u);
        case UNIFORMS_DEBUG_INSTANCE_INDEX:
            return debug_colour(
                uint32_t(s.meshlet_table[uint32_t(key >> LH_KEY_TRI_BITS) & (LH_MAX_SCENE_MESHLETS - 1)].meshlet_index) %
                10u);
        case UNIFORMS_DEBUG_NORMALS: return {nrm.x * 0.5f + 0.5f, nrm.y * 0.5f + 0.5f, nrm.z * 0.5f + 0.5f};
        default: break;
    }
    const f3 ambient = mul(mul(base, 1.0f - metallic), ambient_sh(nrm));
    const f3 sun = mul(mul(brdf(view_vector, param_sun_dir(prm, u), nrm, roughness, metallic, base,
                                (mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) != 0),
                           kPi * shadow_sum),
                       param_sun_intensity(prm, u));
    const f3 lit = add(sun, ambient);
    if (u.debug == UNIFORMS_DEBUG_SHADER_CLOCK) {
        const uint64_t end_time = clock64();
        // timediff(): at most one wrap of the counter (debug.glsl)
        const uint64_t dt = end_time >= start_time ? end_time - start_time : uint64_t(~0u) - (start_time - end_time);
        return temperature(clampf(float(dt) / kClockHeatScale, 0.0f, 1.0f));
    }
    return lit;
}

LH_DEV float srgb_oetf(float x) { return x <= 0.003130f ? x * 12.92f : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f; }
// The LUT-free display curve: x / (x + 1), the sRGB OETF, 8 bits.  Only the
// 8-bit code leaves this function, so it runs on the hardware reciprocal,
// log2 and exp2 (about 1 ulp each) instead of the IEEE division and powf
// (about 50 instructions a channel, 13% of the resolve): the code moves by
// at most one step, and only where the exact value lies within a few ulp of
// a rounding boundary.
LH_DEV uint32_t encode8(float x) {
    x = fmaxf(x, 0.0f);
    const float y = x * __builtin_amdgcn_rcpf(x + 1.0f);
    const float c = y <= 0.003130f ? y * 12.92f
                                   : 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(y) * (1.0f / 2.4f)) - 0.055f;
    return uint32_t(clampf(c, 0.0f, 1.0f) * 255.0f + 0.5f);
}
LH_DEV uint32_t tonemap_pixel(f3 c) { return encode8(c.z) | (encode8(c.y) << 8) | (encode8(c.x) << 16) | (255u << 24); }

// Shadow-map depth with linear filtering and clamp to edge (textureLod of
// the depth array at level 0).
LH_DEV float shadow_depth(const lh_view& v, uint32_t layer, f2 uv) {
    const int32_t S = int32_t(v.shadow_size);
    const uint32_t* map = v.shadow_maps + size_t(layer) * v.shadow_size * v.shadow_size;
    const float fx = uv.x * float(S) - 0.5f, fy = uv.y * float(S) - 0.5f;
    const float x0 = floorf(fx), y0 = floorf(fy);
    const float a = fx - x0, b = fy - y0;
    const int32_t i = int32_t(x0), j = int32_t(y0);
    auto at = [&](int32_t ii, int32_t jj) {
        ii = min(max(ii, 0), S - 1);
        jj = min(max(jj, 0), S - 1);
        return __uint_as_float(map[size_t(jj) * v.shadow_size + uint32_t(ii)]);
    };
    return mixf(mixf(at(i, j), at(i + 1, j), a), mixf(at(i, j + 1), at(i + 1, j + 1), a), b);
}

// tony_mc_mapface (display_transform.comp:5-17): the stimulus encoded as
// x / (x + 1), aligned to texel centres, looked up with trilinear filtering
// and clamp to edge in the view's LH_TONEMAP_LUT_DIMS^3 RGBA float LUT.
LH_DEV f3 lut_lookup(const float* lut, f3 c) {
    constexpr int32_t N = int32_t(LH_TONEMAP_LUT_DIMS);
    constexpr float dims = float(LH_TONEMAP_LUT_DIMS);
    const float x[3] = {fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f)};
    int32_t i0[3];
    float fr[3];
    for (int k = 0; k < 3; k++) {
        const float uvw = (x[k] / (x[k] + 1.0f)) * ((dims - 1.0f) / dims) + 0.5f / dims;
        const float t = uvw * dims - 0.5f, fl = floorf(t);
        fr[k] = t - fl;
        i0[k] = int32_t(fl);
    }
    auto at = [&](int32_t r, int32_t g, int32_t b) {
        r = min(max(r, 0), N - 1);
        g = min(max(g, 0), N - 1);
        b = min(max(b, 0), N - 1);
        const float4 q = reinterpret_cast<const float4*>(lut)[(size_t(b) * N + g) * N + r];
        return f3{q.x, q.y, q.z};
    };
    auto mix3 = [](f3 a, f3 b, float t) { return f3{mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; };
    f3 plane[2];
    for (int db = 0; db < 2; db++) {
        const f3 g0 = mix3(at(i0[0], i0[1], i0[2] + db), at(i0[0] + 1, i0[1], i0[2] + db), fr[0]);
        const f3 g1 = mix3(at(i0[0], i0[1] + 1, i0[2] + db), at(i0[0] + 1, i0[1] + 1, i0[2] + db), fr[0]);
        plane[db] = mix3(g0, g1, fr[1]);
    }
    return mix3(plane[0], plane[1], fr[2]);
}
LH_DEV uint32_t encode_display(float x) { return uint32_t(clampf(srgb_oetf(x), 0.0f, 1.0f) * 255.0f + 0.5f); }

// Display transform of one pixel (display_transform.comp:37-81): the tone
// curve (the view's LUT, or x / (x + 1) without one) and sRGB encoding, or,
// with Uniforms.debug_shadowmaps, the four cascades in the top-left quarter
// of the window (2 x 2, depth d shown as (d - 0.5) * 100 + 0.5 without the
// tone curve).
LH_DEV uint32_t display_pixel(const lh_view& v, uint32_t x, uint32_t y, f3 c) {
    const Uniforms& u = *v.uniforms;
    if (u.debug_shadowmaps) {
        const float ux = (float(x) + 0.5f) / float(u.window_size.x), uy = (float(y) + 0.5f) / float(u.window_size.y);
        if (ux < 0.5f && uy < 0.5f) {
            const float sx = ux * 2.0f / 0.5f, sy = uy * 2.0f / 0.5f;
            const uint32_t layer = (sy > 1.0f ? 2u : 0u) + (sx > 1.0f ? 1u : 0u);
            const float g = (shadow_depth(v, layer, f2{sx - floorf(sx), sy - floorf(sy)}) - 0.5f) * 100.0f + 0.5f;
            const uint32_t e = uint32_t(clampf(srgb_oetf(fmaxf(g, 0.0f)), 0.0f, 1.0f) * 255.0f + 0.5f);
            return e | (e << 8) | (e << 16) | (255u << 24);
        }
    }
    if (v.tonemap_lut) {
        const f3 d = lut_lookup(v.tonemap_lut, c);
        return encode_display(d.z) | (encode_display(d.y) << 8) | (encode_display(d.x) << 16) | (255u << 24);
    }
    return tonemap_pixel(c);
}

}  // namespace shading
