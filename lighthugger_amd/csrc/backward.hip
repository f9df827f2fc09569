// Adjoint of the visibility-buffer resolve (gfx950).
//
// The forward is the per-pixel function of render_geometry.comp:70-266 as
// implemented by shading::resolve_pixel; the adjoint differentiates it with
// respect to the material factors, the sun direction and intensity, and the
// mesh-space vertex positions and normals of the triangle each pixel shows.
// Held fixed, like the visibility buffer itself: the back-face flip, the
// cascade choice and PCF shadow factor, and the texture LOD (levels and blend
// weight), read from the forward's per-pixel record when the caller kept one
// (lh_resolve_params.frozen), else recomputed here exactly as the forward
// computes them.
//
// Precision: the clip-space chain (positions, clip coordinates, barycentrics
// and their reverse) and the texel-cell choice run in FP64 from the FP32
// inputs; the BRDF, texture filtering, normals and light run in FP32.
//
// Barycentrics are not stored: they are recomputed from the key.  Their
// adjoint uses the homogeneous form of the perspective-correct barycentrics
// (the same function CalcFullBary evaluates, vbuffer.glsl:11-63): with
// q = (ndc.x, ndc.y, 1) and c_j = (x_j, y_j, w_j) of the clip positions,
// lambda_j = D_j / (D_0 + D_1 + D_2), D_j = q . (c_{j+1} x c_{j+2}).
//
// Work split: a persistent grid of one-wave blocks walks the 8x8 pixel tiles.
// Per tile the gradients of the pixels' vertices are merged in an LDS hash
// table and added to the gradient arena once per distinct vec3; material-factor
// gradients are summed over the wave per distinct mesh info and kept per block
// in registers (one lane per mesh info) until the block ends;
// the six light gradients stay in registers and every block writes one
// partial, which the last block to finish sums in a fixed order.
#include "kernels.h"
#include "shading.h"

using namespace lhd;
using namespace shading;

namespace {

// ---- FP64 vector math (the adjoint's working precision) ----
struct d2 {
    double x, y;
};
struct d3 {
    double x, y, z;
};
LH_DEV d3 D3(f3 a) { return {double(a.x), double(a.y), double(a.z)}; }
LH_DEV d3 D3(const LhVec3& a) { return {double(a.x), double(a.y), double(a.z)}; }
LH_DEV d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LH_DEV d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LH_DEV d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
LH_DEV double dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
LH_DEV d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// columns 0..2 of a 4x4 (column-major) applied to a point / the linear part
LH_DEV d3 col3(const LhVec4& c) { return {double(c.x), double(c.y), double(c.z)}; }

// ---- FP32 part: BRDF, textures, normals ----
LH_DEV float pow4f(float x) {
    const float x2 = x * x;
    return x2 * x2;
}
LH_DEV float pow5f(float x) { return pow4f(x) * x; }
LH_DEV f3 madd(f3 a, f3 b, float s) { return add(a, mul(b, s)); }

// Reciprocal and reciprocal square root to about 1 ulp: the FP32 part of the
// adjoint is held to a relative 1e-4, far wider than what the IEEE-rounded
// division and square root it replaces would buy, at a tenth of the
// instructions.
LH_DEV float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
LH_DEV float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
LH_DEV f3 unit(f3 a) { return mul(a, rsqrt(dot3(a, a))); }

// u = a / |a|: the gradient of a from the gradient of u.
LH_DEV f3 normalize_bwd(f3 a, f3 gu) {
    const float il = rsqrt(dot3(a, a));
    const f3 u = mul(a, il);
    return mul(sub(gu, mul(u, dot3(u, gu))), il);
}

struct BrdfGrad {
    f3 v, l, n, base;
    float pr, metallic;
};

// BRDF of shading::brdf (thin = false): value and reverse mode for the
// upstream gradient g of its RGB output.
LH_DEV f3 brdf_bwd(f3 v, f3 l, f3 n, float pr, float m, f3 base, f3 g, BrdfGrad& o) {
    const f3 hv = add(v, l);
    const f3 h = unit(hv);
    const f3 diffuse = mul(base, 1.0f - m);
    const float refl = (0.5f / 2.5f) * (0.5f / 2.5f);
    const float d0 = refl * (1.0f - m);
    const f3 f0 = madd(f3{d0, d0, d0}, base, m);
    const float nv = dot3(n, v), nl = dot3(n, l), nh = dot3(n, h), lh = dot3(l, h);
    const float NoV = clampf(nv, 1e-5f, 1.0f), NoL = clampf(nl, 0.0f, 1.0f);
    const float NoH = clampf(nh, 0.0f, 1.0f), LoH = clampf(lh, 0.0f, 1.0f);
    const float a = pr * pr, a2 = a * a;
    const float fD = (NoH * a2 - NoH) * NoH + 1.0f;
    const float iD = rcp((kPi * fD) * fD);
    const float D = a2 * iD;
    const float p = pow5f(1.0f - LoH);
    const f3 F{f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
    const float sL2 = (-NoL * a2 + NoL) * NoL + a2, sV2 = (-NoV * a2 + NoV) * NoV + a2;
    const float isL = rsqrt(sL2), isV = rsqrt(sV2), sL = sL2 * isL, sV = sV2 * isV;
    const float den = NoL * sV + NoV * sL;
    const float iden = rcp(den);
    const float Vis = 0.5f * iden;
    const float f90 = 0.5f + ((2.0f * a) * LoH) * LoH;
    const float q5L = pow5f(1.0f - NoL), q5V = pow5f(1.0f - NoV);
    const float FSL = 1.0f + (f90 - 1.0f) * q5L, FSV = 1.0f + (f90 - 1.0f) * q5V;
    const float FdB = (FSL * FSV) * (1.0f / kPi);
    const float DV = D * Vis;
    const f3 Fr = mul(F, DV), Fd = mul(diffuse, FdB);
    const f3 irr = mul(add(Fd, Fr), NoL);
    // max(irr, 0) and the NaN guard pass the gradient only where irr > 0
    const f3 gi{irr.x > 0.0f ? g.x : 0.0f, irr.y > 0.0f ? g.y : 0.0f, irr.z > 0.0f ? g.z : 0.0f};
    const f3 gs = mul(gi, NoL);
    float gNoL = dot3(gi, add(Fd, Fr));
    const f3 gF = mul(gs, DV);
    const float gDV = dot3(gs, F);
    const float gD = gDV * Vis, gVis = gDV * D;
    const f3 gdiff = mul(gs, FdB);
    const float gFdB = dot3(gs, diffuse);
    const float gFSL = gFdB * FSV * (1.0f / kPi), gFSV = gFdB * FSL * (1.0f / kPi);
    const float gf90 = gFSL * q5L + gFSV * q5V;
    gNoL += gFSL * (f90 - 1.0f) * (-5.0f * pow4f(1.0f - NoL));
    float gNoV = gFSV * (f90 - 1.0f) * (-5.0f * pow4f(1.0f - NoV));
    float ga = gf90 * (2.0f * LoH * LoH);
    float gLoH = gf90 * (4.0f * a * LoH);
    const float gden = -gVis * Vis * iden;
    gNoV += gden * sL;
    gNoL += gden * sV;
    const float gsL = gden * NoV, gsV = gden * NoL;
    gNoL += gsL * NoL * (1.0f - a2) * isL;
    gNoV += gsV * NoV * (1.0f - a2) * isV;
    float ga2 = gsL * (1.0f - NoL * NoL) * (0.5f * isL) + gsV * (1.0f - NoV * NoV) * (0.5f * isV);
    const f3 gf0 = mul(gF, 1.0f - p);
    const float gp = (gF.x * (1.0f - f0.x) + gF.y * (1.0f - f0.y)) + gF.z * (1.0f - f0.z);
    gLoH += gp * (-5.0f * pow4f(1.0f - LoH));
    const float E = D * rcp(fD);
    ga2 += gD * (iD - 2.0f * E * NoH * NoH);
    const float gNoH = gD * (-2.0f * E * (2.0f * NoH * (a2 - 1.0f)));
    ga += ga2 * 2.0f * a;
    o.pr = ga * 2.0f * pr;
    o.metallic = dot3(gf0, sub(base, f3{refl, refl, refl})) - dot3(gdiff, base);
    o.base = add(mul(gf0, m), mul(gdiff, 1.0f - m));
    const float gnv = (nv > 1e-5f && nv < 1.0f) ? gNoV : 0.0f;
    const float gnl = (nl > 0.0f && nl < 1.0f) ? gNoL : 0.0f;
    const float gnh = (nh > 0.0f && nh < 1.0f) ? gNoH : 0.0f;
    const float glh = (lh > 0.0f && lh < 1.0f) ? gLoH : 0.0f;
    o.n = add(add(mul(v, gnv), mul(l, gnl)), mul(h, gnh));
    const f3 ghv = normalize_bwd(hv, add(mul(n, gnh), mul(l, glh)));
    o.v = add(mul(n, gnv), ghv);
    o.l = add(add(mul(n, gnl), mul(h, glh)), ghv);
    return {irr.x > 0.0f ? irr.x : 0.0f, irr.y > 0.0f ? irr.y : 0.0f, irr.z > 0.0f ? irr.z : 0.0f};
}

// ---- textures with the forward's LOD held fixed ----
// A sample and its derivatives with respect to uv, per channel, so one fetch
// of the texels serves the value and the reverse.
struct TexSample {
    float val[4], du[4], dv[4];
};
// Adds weight times the bilinear sample at one level (and its uv derivatives).
// The texel cell and its fractions come from FP64 coordinates: FP32 would
// pick a neighbouring cell for uvs within rounding of a texel boundary.
LH_DEV void bilinear_acc(const LhTexture& t, uint32_t level, d2 uv, float weight, TexSample& o) {
    const int32_t w = int32_t(max(1u, t.width >> level)), h = int32_t(max(1u, t.height >> level));
    const double u = uv.x * double(w) - 0.5, vv = uv.y * double(h) - 0.5;
    const double fu = floor(u), fv = floor(vv);
    const float a = float(u - fu), b = float(vv - fv);
    const int32_t i = int32_t(fu), j = int32_t(fv);
    const TexelQuad q = texel_quad(t, level, w, h, i, j);
    const float c00[4] = {q.t00.x, q.t00.y, q.t00.z, q.t00.w}, c10[4] = {q.t10.x, q.t10.y, q.t10.z, q.t10.w};
    const float c01[4] = {q.t01.x, q.t01.y, q.t01.z, q.t01.w}, c11[4] = {q.t11.x, q.t11.y, q.t11.z, q.t11.w};
    for (int k = 0; k < 4; k++) {
        const float top = c00[k] * (1.0f - a) + c10[k] * a, bot = c01[k] * (1.0f - a) + c11[k] * a;
        o.val[k] += weight * (top * (1.0f - b) + bot * b);
        o.du[k] += weight * float(w) * ((1.0f - b) * (c10[k] - c00[k]) + b * (c11[k] - c01[k]));
        o.dv[k] += weight * float(h) * (bot - top);
    }
}
// Trilinear sample with the frozen LOD.  Unused index: white, no gradient.
LH_DEV TexSample sample_d(const lh_scene& s, uint32_t index, const Lod& lod, d2 uv) {
    TexSample o;
    for (int k = 0; k < 4; k++) o.val[k] = o.du[k] = o.dv[k] = 0.0f;
    if (!has_texture(s, index)) {
        for (int k = 0; k < 4; k++) o.val[k] = 1.0f;
        return o;
    }
    const LhTexture& t = s.textures[index];
    bilinear_acc(t, lod.l0, uv, 1.0f - lod.f, o);
    if (lod.f != 0.0f) bilinear_acc(t, lod.l1, uv, lod.f, o);
    return o;
}

struct LightGrad {
    f3 sun_dir, sun_intensity;  // a lane's sums over its pixels (a few per block)
};

// Vertex gradients of one tile are merged in an LDS hash table before they
// reach the gradient arena: neighbouring pixels share vertices, and
// same-address atomics from one wave serialize.  One entry per vertex holds
// its position and normal gradients (six floats), keyed by the arena words of
// both vec3s (position word | normal word << 32; ~0u for a vec3 outside the
// arena), so a mesh info whose normals differ from another's over the same
// positions gets entries of its own.  The table is kept across the block's
// tiles and flushed only when it holds more than kFlushAt entries after a tile,
// and after the block's last tile: a tile (about eight triangles of about
// eight pixels each) touches a dozen vertices, so most blocks flush once, and
// the atomics of a flush, which a later load of the wave would wait behind,
// leave the per-tile path.  A tile claims at most 64 pixels x 3 vertices =
// 192 entries, so at most 352 of the 512 slots are taken (69% load) in 16 KB,
// and two one-wave blocks per SIMD fit in LDS.  Claimed slots are listed (one
// LDS atomic per wave and probe round) so the flush visits only those.
constexpr uint32_t kHashSlots = 512;
constexpr uint32_t kTileClaims = 192;
constexpr uint32_t kFlushAt = 160;
constexpr uint32_t kHashKeysMax = kFlushAt + kTileClaims;
static_assert(kHashKeysMax < kHashSlots, "an empty slot ends every probe");
constexpr uint64_t kHashEmpty = ~0ull;
constexpr uint32_t kNoWord = ~0u;
struct VertexHash {
    uint64_t* keys;   // [kHashSlots]
    float* vals;      // [kHashSlots * 6]: position gradient, normal gradient
    uint32_t* used;   // [kHashKeysMax] slots claimed since the last flush
    uint32_t* count;  // number of them
};

// Word index of a vec3 in the gradient arena, or kNoWord outside it.
LH_DEV uint32_t arena_word(const lh_resolve_grads& gr, uint64_t addr) {
    if (!gr.geometry || addr < gr.geometry_base) return kNoWord;
    const uint64_t k = (addr - gr.geometry_base) / 4;
    return (k + 3 > gr.geometry_bytes / 4 || k >= kNoWord) ? kNoWord : uint32_t(k);
}

// Claims the entries of the three vertices of each active lane's pixel, all
// three probes in flight together, and returns their slots (kHashSlots for a
// vertex with neither vec3 in the arena).
LH_DEV void claim_vertices(const VertexHash& h, const uint64_t key[3], uint32_t slot[3]) {
    bool done[3];
    for (int j = 0; j < 3; j++) {
        done[j] = key[j] == kHashEmpty;
        const uint32_t mix = uint32_t(key[j]) * 2654435761u ^ uint32_t(key[j] >> 32) * 0x85EBCA6Bu;
        slot[j] = done[j] ? kHashSlots : mix >> 23;  // 9 bits: kHashSlots == 512
    }
    while (!(done[0] && done[1] && done[2])) {
        uint64_t old[3];
        for (int j = 0; j < 3; j++)
            if (!done[j]) old[j] = atomicCAS(reinterpret_cast<unsigned long long*>(h.keys) + slot[j],
                                             kHashEmpty, key[j]);
        for (int j = 0; j < 3; j++) {
            const bool fresh = !done[j] && old[j] == kHashEmpty;
            const uint64_t m = ballot(fresh);
            if (m) {
                uint32_t base = 0;
                if (lane_id() == first_lane(m)) base = atomicAdd(h.count, uint32_t(__popcll(m)));
                base = __builtin_amdgcn_readlane(base, int(first_lane(m)));
                if (fresh) h.used[base + lanes_below(m)] = slot[j];
            }
            if (!done[j]) {
                if (fresh || old[j] == key[j]) done[j] = true;
                else slot[j] = (slot[j] + 1) & (kHashSlots - 1);
            }
        }
    }
}

LH_DEV void add_vertex(const VertexHash& h, uint32_t slot, uint32_t part, f3 g) {
    if (slot >= kHashSlots) return;
    float* val = h.vals + 6 * slot + part;
    atomicAdd(val, g.x);
    atomicAdd(val + 1, g.y);
    atomicAdd(val + 2, g.z);
}

// Called by the whole wave after a tile: once the table holds more than
// kFlushAt entries, or when `last`, adds every merged vec3 to the arena and
// empties the table.
LH_DEV void flush_geometry(const lh_resolve_grads& gr, const VertexHash& h, bool last) {
    __syncthreads();
    const uint32_t n = *h.count;
    if (n <= kFlushAt && !last) return;
    for (uint32_t i = threadIdx.x; i < n; i += 64) {
        const uint32_t slot = h.used[i];
        const uint64_t key = h.keys[slot];
        float* val = h.vals + 6 * slot;
        const uint32_t kp = uint32_t(key), kn = uint32_t(key >> 32);
        if (kp != kNoWord) {
            atomicAdd(gr.geometry + kp, val[0]);
            atomicAdd(gr.geometry + kp + 1, val[1]);
            atomicAdd(gr.geometry + kp + 2, val[2]);
        }
        if (kn != kNoWord) {
            atomicAdd(gr.geometry + kn, val[3]);
            atomicAdd(gr.geometry + kn + 1, val[4]);
            atomicAdd(gr.geometry + kn + 2, val[5]);
        }
        h.keys[slot] = kHashEmpty;
        for (int c = 0; c < 6; c++) val[c] = 0.0f;
    }
    __syncthreads();
    if (threadIdx.x == 0) *h.count = 0;
    __syncthreads();
}

LH_DEV bool supported(const MeshInfo& mi) {
    return (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) && !(mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) &&
           mi.normal_texture_index == UNUSED_TEXTURE_INDEX;
}

// Quantities held fixed, computed in FP32 exactly as shading::resolve_pixel.
struct Frozen {
    float flip, K;
    Lod lod_b, lod_m;
};
LH_DEV Frozen frozen(const lh_scene& s, const lh_view& v, const MeshInfo& mi, const Surface& sf, f2 ndc, f2 win) {
    const Uniforms& u = *v.uniforms;
    const BaryDeriv bary = calc_full_bary(sf.clip[0], sf.clip[1], sf.clip[2], ndc, win);
    const Interp3 world = interp3(bary, sf.pos[0], sf.pos[1], sf.pos[2]);
    const f2 ts{mi.texture_scale.x, mi.texture_scale.y};
    f2 uvs[3];
    for (int j = 0; j < 3; j++) {
        const f2 t = load_uv(mi, sf.vi[j]);
        uvs[j] = {t.x * ts.x, t.y * ts.y};
    }
    const Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);  // derivatives only: the offset does not matter
    const f3 view_vector = normalize3(sub(mk3(u.camera_pos), world.v));
    const f3 gn = cross3(sub(sf.pos[1], sf.pos[0]), sub(sf.pos[2], sf.pos[0]));
    Frozen fz;
    fz.flip = dot3(gn, view_vector) < 0.0f ? -1.0f : 1.0f;
    fz.K = kPi * shadow_factor(v, mi, world.v);
    fz.lod_b = has_texture(s, mi.base_color_texture_index) ? texture_lod(s.textures[mi.base_color_texture_index], uv)
                                                           : Lod{0, 0, 0.0f};
    fz.lod_m = has_texture(s, mi.metallic_roughness_texture_index)
                   ? texture_lod(s.textures[mi.metallic_roughness_texture_index], uv)
                   : Lod{0, 0, 0.0f};
    return fz;
}

// One pixel: scatters vertex gradients, returns its material-factor gradients
// (mesh info in mesh_info, or ~0u when it contributes nothing) and adds its
// light gradients to lg.  The clip-space chain (barycentrics and their
// reverse) is evaluated in FP64 from the FP32 inputs, because for small,
// distant triangles it cancels ~(depth / size)^2 of the FP32 precision; the
// BRDF, texture and normal parts are well conditioned and run in FP32.
template <bool kRecorded>
LH_DEV void pixel_bwd(const lh_scene& s, const lh_view& v, const lh_resolve_params& prm, const float* grad_radiance,
                      const lh_resolve_grads& gr, const VertexHash& hash, uint32_t x, uint32_t y, uint32_t& mesh_info,
                      f3& g_bcf, float& g_rf, float& g_mf, LightGrad& lg) {
    mesh_info = ~0u;
    const size_t o = size_t(y) * v.width + x;
    Surface sf;
    if (!decode(s, v, v.visbuffer[o], sf)) return;  // sky: constant radiance
    const Uniforms& u = *v.uniforms;
    const Instance& in = s.instances[sf.instance];
    const MeshInfo& mi = s.mesh_infos[sf.mesh_info];
    if (!supported(mi)) {
        atomicOr(&v.counters[LH_CNT_FLAGS], LH_FLAG_BACKWARD_UNSUPPORTED);
        return;
    }
    Frozen fz;
    // The record of the forward (lh_resolve_ex with the same buffer) is used
    // only where its valid bit is set; its levels are clamped to the
    // texture's chain, so a stale or foreign record cannot address texels
    // outside the texture.  Elsewhere the quantities are recomputed.
    const uint4 r = kRecorded ? reinterpret_cast<const uint4*>(prm.frozen)[o] : make_uint4(0, 0, 0, 0);
    if (kRecorded && (r.w >> 31) != 0u) {
        fz.flip = (r.w & 1u) ? -1.0f : 1.0f;
        fz.K = kPi * __uint_as_float(r.x);
        auto lod = [&](uint32_t index, uint32_t l0, uint32_t fbits) {
            if (!has_texture(s, index)) return Lod{0, 0, 0.0f};
            const uint32_t top = s.textures[index].levels - 1;
            const uint32_t a = min(l0, top);
            return Lod{a, a + 1 < top ? a + 1 : top, clampf(__uint_as_float(fbits), 0.0f, 1.0f)};
        };
        fz.lod_b = lod(mi.base_color_texture_index, (r.w >> 8) & 255u, r.y);
        fz.lod_m = lod(mi.metallic_roughness_texture_index, (r.w >> 16) & 255u, r.z);
    } else {
        const f2 win{float(u.window_size.x), float(u.window_size.y)};
        const f2 ndc{float(x) / win.x * 2.0f - 1.0f, float(y) / win.y * 2.0f - 1.0f};
        fz = frozen(s, v, mi, sf, ndc, win);
    }
    const float4 gin = reinterpret_cast<const float4*>(grad_radiance)[o];
    const f3 g{gin.x, gin.y, gin.z};
    // ---- forward: clip-space chain in FP64 ----
    const LhMat4& PV = u.combined_perspective_view;
    const d3 q{double(x) / double(u.window_size.x) * 2.0 - 1.0, double(y) / double(u.window_size.y) * 2.0 - 1.0, 1.0};
    d3 c[3];
    double sumD, invD, lam[3];
    f3 dpos[3], to_cam;  // world positions relative to vertex 0; camera - world
    {
        const LhMat4& T = in.transform;
        d3 pos[3];
        for (int j = 0; j < 3; j++) {
            const d3 L = D3(sf.local[j]);
            pos[j] = ((col3(T.c[0]) * L.x + col3(T.c[1]) * L.y) + col3(T.c[2]) * L.z) + col3(T.c[3]);
            // clip (x, y, w) = rows 0, 1, 3 of PV [pos, 1]
            auto row = [&](int r) {
                auto e = [&](int k) {
                    const LhVec4& cc = PV.c[k];
                    return double(r == 0 ? cc.x : r == 1 ? cc.y : cc.w);
                };
                return ((e(0) * pos[j].x + e(1) * pos[j].y) + e(2) * pos[j].z) + e(3);
            };
            c[j] = {row(0), row(1), row(3)};
        }
        // perspective-correct barycentrics: lambda_j = D_j / sum D, D_j = q . (c_{j+1} x c_{j+2})
        double Dk[3];
        for (int k = 0; k < 3; k++) Dk[k] = dot(q, cross(c[(k + 1) % 3], c[(k + 2) % 3]));
        sumD = (Dk[0] + Dk[1]) + Dk[2];
        invD = 1.0 / sumD;
        for (int k = 0; k < 3; k++) lam[k] = Dk[k] * invD;
        const d3 tc = D3(u.camera_pos) - ((pos[0] * lam[0] + pos[1] * lam[1]) + pos[2] * lam[2]);
        to_cam = {float(tc.x), float(tc.y), float(tc.z)};
        for (int j = 0; j < 3; j++) {
            const d3 dp = pos[j] - pos[0];
            dpos[j] = {float(dp.x), float(dp.y), float(dp.z)};
        }
    }
    const float lamf[3] = {float(lam[0]), float(lam[1]), float(lam[2])};
    d2 uv;  // FP64: it selects the texel cells
    f2 duv[3];
    {
        d2 uvs[3];
        for (int j = 0; j < 3; j++) {
            const f2 t = load_uv(mi, sf.vi[j]);
            uvs[j] = {double(t.x) * double(mi.texture_scale.x), double(t.y) * double(mi.texture_scale.y)};
        }
        uv = {((lam[0] * uvs[0].x + lam[1] * uvs[1].x) + lam[2] * uvs[2].x) + double(mi.texture_offset.x),
              ((lam[0] * uvs[0].y + lam[1] * uvs[1].y) + lam[2] * uvs[2].y) + double(mi.texture_offset.y)};
        for (int j = 0; j < 3; j++) duv[j] = {float(uvs[j].x - uvs[0].x), float(uvs[j].y - uvs[0].y)};
    }
    // ---- forward: shading in FP32 ----
    f3 nraw[3], nn[3];
    for (int j = 0; j < 3; j++) {
        nraw[j] = load_normal(mi, sf.vi[j]);
        nn[j] = unit(nraw[j]);
    }
    const f3 V = unit(to_cam);
    const f3 nI = add(add(mul(nn[0], lamf[0]), mul(nn[1], lamf[1])), mul(nn[2], lamf[2]));
    const LhMat3& N = in.normal_transform;
    const f3 m3 = add(add(mul(mk3(N.c[0]), nI.x), mul(mk3(N.c[1]), nI.y)), mul(mk3(N.c[2]), nI.z));
    const f3 nrm = mul(unit(m3), fz.flip);
    const f3 bcf = param_base_color(prm, mi, sf.mesh_info);
    const float rf = param_roughness(prm, s, sf.mesh_info), mf = param_metallic(prm, s, sf.mesh_info);
    const f3 sun_dir = param_sun_dir(prm, u), sun_int = param_sun_intensity(prm, u);
    const TexSample tb = sample_d(s, mi.base_color_texture_index, fz.lod_b, uv);
    const TexSample tm = sample_d(s, mi.metallic_roughness_texture_index, fz.lod_m, uv);
    const f3 bc{tb.val[0], tb.val[1], tb.val[2]};
    const f3 base = mul(bc, bcf);
    const float roughness = tm.val[1] * rf, metallic = tm.val[2] * mf;
    const float shv = nrm.y * 0.025f + 0.025f;  // SH ambient: constant and y bands
    const float K = fz.K;
    asm volatile("" ::: "memory");
    // ---- reverse ----
    // out = B * K * I + base * (1 - metallic) * shv;  B needs the upstream gB = g * I * K
    BrdfGrad bg;
    const f3 B = brdf_bwd(V, sun_dir, nrm, roughness, metallic, base, mul(mul(g, sun_int), K), bg);
    const f3 gI = mul(mul(g, B), K);
    lg.sun_intensity = add(lg.sun_intensity, gI);
    lg.sun_dir = add(lg.sun_dir, bg.l);
    const f3 gbase = add(mul(g, (1.0f - metallic) * shv), bg.base);
    const float gmet = -dot3(g, base) * shv + bg.metallic;
    f3 gnrm = bg.n;
    gnrm.y += dot3(g, base) * (1.0f - metallic) * 0.025f;
    mesh_info = sf.mesh_info;
    g_bcf = mul(gbase, bc);
    g_rf = bg.pr * tm.val[1];
    g_mf = gmet * tm.val[2];
    // uv gradient: the upstream of the sampled channels times their derivatives
    const f3 gb3 = mul(gbase, bcf);
    const float gr1 = bg.pr * rf, gm2 = gmet * mf;
    const f2 guv{((gb3.x * tb.du[0] + gb3.y * tb.du[1]) + gb3.z * tb.du[2]) + (gr1 * tm.du[1] + gm2 * tm.du[2]),
                 ((gb3.x * tb.dv[0] + gb3.y * tb.dv[1]) + gb3.z * tb.dv[2]) + (gr1 * tm.dv[1] + gm2 * tm.dv[2])};
    // normal chain: nrm = flip * normalize(N nI), nI = sum lambda_j normalize(nraw_j)
    const f3 gm3 = normalize_bwd(m3, mul(gnrm, fz.flip));
    const f3 gnI{dot3(mk3(N.c[0]), gm3), dot3(mk3(N.c[1]), gm3), dot3(mk3(N.c[2]), gm3)};
    // world position through the view vector (shadow and LOD are fixed)
    const f3 gworld = neg(normalize_bwd(to_cam, bg.v));
    // every lambda-weighted quantity relative to vertex 0: lambda sums to 1, so
    // a common shift of the lambda gradients cancels and the relative form
    // keeps their differences exact
    double glam[3];
    f3 gnraw[3];
    for (int j = 0; j < 3; j++) {
        glam[j] = double(dot3(gworld, dpos[j])) + double(dot3(gnI, sub(nn[j], nn[0]))) +
                  double(guv.x * duv[j].x + guv.y * duv[j].y);
        gnraw[j] = normalize_bwd(nraw[j], mul(gnI, lamf[j]));
    }
    const double lg_dot = (glam[0] * lam[0] + glam[1] * lam[1]) + glam[2] * lam[2];
    d3 gc[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 3; k++) {
        const double gD = (glam[k] - lg_dot) * invD;
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        gc[a] = gc[a] + cross(c[b], q) * gD;
        gc[b] = gc[b] + cross(q, c[a]) * gD;
    }
    uint32_t slot[3];
    {
        uint64_t key[3];
        for (int j = 0; j < 3; j++)
            key[j] = uint64_t(arena_word(gr, mi.positions + 12ull * sf.vi[j])) |
                     uint64_t(arena_word(gr, mi.normals + 12ull * sf.vi[j])) << 32;
        claim_vertices(hash, key, slot);
    }
    for (int j = 0; j < 3; j++) add_vertex(hash, slot[j], 3, gnraw[j]);
    const LhMat4& T = in.transform;
    for (int j = 0; j < 3; j++) {
        d3 gpos = d3{gworld.x, gworld.y, gworld.z} * lam[j];
        // clip (x, y, w) <- pos: d clip_r / d pos_k = PV.c[k][r]
        for (int k = 0; k < 3; k++) {
            const LhVec4& cc = PV.c[k];
            const double d = (double(cc.x) * gc[j].x + double(cc.y) * gc[j].y) + double(cc.w) * gc[j].z;
            if (k == 0) gpos.x += d;
            if (k == 1) gpos.y += d;
            if (k == 2) gpos.z += d;
        }
        const d3 gL{dot(col3(T.c[0]), gpos), dot(col3(T.c[1]), gpos), dot(col3(T.c[2]), gpos)};
        add_vertex(hash, slot[j], 0, f3{float(gL.x), float(gL.y), float(gL.z)});
    }
}

// Occupancy target (waves per SIMD) of the adjoint; 0 leaves the register
// budget to the compiler.
#ifndef LH_BACKWARD_WAVES
#define LH_BACKWARD_WAVES 2
#endif
#if LH_BACKWARD_WAVES > 0
#define LH_BACKWARD_OCCUPANCY LH_WAVES_ATTR(LH_BACKWARD_WAVES)
#else
#define LH_BACKWARD_OCCUPANCY
#endif

template <bool kRecorded>
__global__ void __launch_bounds__(64) LH_BACKWARD_OCCUPANCY k_resolve_backward(lh_scene s, OneView vs, lh_resolve_params prm,
                                                         const float* grad_radiance, lh_resolve_grads gr,
                                                         double* partials) {
    __shared__ uint64_t hkeys[kHashSlots];
    __shared__ float hvals[kHashSlots * 6];
    __shared__ uint32_t hused[kHashKeysMax];
    __shared__ uint32_t hcount;
    const VertexHash hash{hkeys, hvals, hused, &hcount};
    for (uint32_t i = threadIdx.x; i < kHashSlots; i += 64) {
        hkeys[i] = kHashEmpty;
        for (int c = 0; c < 6; c++) hvals[6 * i + c] = 0.0f;
    }
    if (threadIdx.x == 0) hcount = 0;
    __syncthreads();
    const lh_view& v = vs.v;
    const uint32_t tiles_x = (v.width + 7) / 8, tiles = tiles_x * ((v.height + 7) / 8);
    LightGrad lg{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    // Material-factor gradients: lane i keeps the sums of the i-th distinct
    // mesh info the block met and adds them to the outputs once, at the end.
    // The few mesh infos that cover most of the view would otherwise take
    // five same-address atomics per tile, which serialize in the L2.
    uint32_t mat_key = ~0u, mat_used = 0;  // mat_used: wave-uniform
    float mat_val[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    auto add_material = [&](uint32_t mi, const float* val) {
        if (gr.base_color_factor) {
            atomicAdd(gr.base_color_factor + 3 * mi, val[0]);
            atomicAdd(gr.base_color_factor + 3 * mi + 1, val[1]);
            atomicAdd(gr.base_color_factor + 3 * mi + 2, val[2]);
        }
        if (gr.roughness_factor) atomicAdd(gr.roughness_factor + mi, val[3]);
        if (gr.metallic_factor) atomicAdd(gr.metallic_factor + mi, val[4]);
    };
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint32_t x = (t % tiles_x) * 8 + (threadIdx.x & 7), y = (t / tiles_x) * 8 + (threadIdx.x >> 3);
        uint32_t mesh_info = ~0u;
        f3 g_bcf{0, 0, 0};
        float g_rf = 0.0f, g_mf = 0.0f;
        if (x < v.width && y < v.height)
            pixel_bwd<kRecorded>(s, v, prm, grad_radiance, gr, hash, x, y, mesh_info, g_bcf, g_rf, g_mf, lg);
        if (gr.geometry) flush_geometry(gr, hash, t + gridDim.x >= tiles);
        // material factors: one add per distinct mesh info of the wave
        uint64_t pending = ballot(mesh_info != ~0u);
        while (pending) {
            const uint32_t lead = __builtin_amdgcn_readlane(mesh_info, int(first_lane(pending)));
            const bool mine = mesh_info == lead;
            const float r[5] = {wave_sum_f32(mine ? g_bcf.x : 0.0f), wave_sum_f32(mine ? g_bcf.y : 0.0f),
                                wave_sum_f32(mine ? g_bcf.z : 0.0f), wave_sum_f32(mine ? g_rf : 0.0f),
                                wave_sum_f32(mine ? g_mf : 0.0f)};
            // every lane holds the wave sums; pick the lane that keeps lead
            const uint64_t hit = ballot(mat_key == lead);
            uint32_t owner = 64;  // none: the table is full, add directly
            if (hit) {
                owner = first_lane(hit);
            } else if (mat_used < 64) {
                owner = mat_used++;
                if (lane_id() == owner) mat_key = lead;
            }
            if (lane_id() == owner) {
                for (int i = 0; i < 5; i++) mat_val[i] += r[i];
            } else if (owner == 64 && lane_id() == first_lane(pending)) {
                add_material(lead, r);
            }
            pending &= ~ballot(mine);
        }
    }
    if (mat_key != ~0u) add_material(mat_key, mat_val);
    double r[6] = {lg.sun_dir.x, lg.sun_dir.y, lg.sun_dir.z, lg.sun_intensity.x, lg.sun_intensity.y,
                   lg.sun_intensity.z};
    for (int i = 0; i < 6; i++) {  // the lanes' sums in FP64, in a fixed order
        double sum = r[i];
        for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
        r[i] = sum;
    }
    if (threadIdx.x < 8) {
        double val = 0.0;
        for (int i = 0; i < 6; i++)
            if (threadIdx.x == uint32_t(i)) val = r[i];
        partials[size_t(blockIdx.x) * 8 + threadIdx.x] = val;
    }
    // The last block to finish sums every block's partial, lane l over blocks
    // l, l + 64, ... then across the lanes in a fixed pattern, so the result
    // does not depend on which block that is or on timing, and adds the sums
    // to the outputs.
    __threadfence();  // this block's partial before its ticket
    uint32_t ticket = 0;
    if (threadIdx.x == 0) ticket = atomicAdd(&v.counters[LH_CNT_BACKWARD_TICKET], 1u);
    ticket = __shfl(ticket, 0, 64);
    if (ticket != gridDim.x - 1) return;
    __threadfence();  // every ticket before the partials are read
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += 64)
        for (int i = 0; i < 6; i++)
            acc[i] += __hip_atomic_load(partials + size_t(b) * 8 + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i = 0; i < 6; i++)
        for (int d = 32; d >= 1; d >>= 1) acc[i] += __shfl_xor(acc[i], d, 64);
    if (threadIdx.x < 6) {
        float* dst = threadIdx.x < 3 ? gr.sun_dir : gr.sun_intensity;
        double val = 0.0;
        for (int i = 0; i < 6; i++)
            if (threadIdx.x == uint32_t(i)) val = acc[i];
        if (dst) dst[threadIdx.x % 3] += float(val);
    }
    if (threadIdx.x == 0) v.counters[LH_CNT_BACKWARD_TICKET] = 0;  // ready for the next call
}

}  // namespace

namespace lhk {

hipError_t resolve_backward(const lh_scene& s, const OneView& vs, const lh_resolve_params& params,
                            const float* grad_radiance, const lh_resolve_grads& grads, hipStream_t st) {
    const lh_view& v = vs.v;
    const uint32_t tiles = ((v.width + 7) / 8) * ((v.height + 7) / 8);
    const uint32_t blocks = std::max(1u, std::min(tiles, kBackwardBlocks));
    // light partials live in the view's raster scratch (free after the raster;
    // the C entry point checks its size against kBackwardScratchBytes)
    double* partials = static_cast<double*>(v.scratch);
    mark("resolve_backward", st);
    if (params.frozen)
        hipLaunchKernelGGL(k_resolve_backward<true>, dim3(blocks), dim3(64), 0, st, s, vs, params, grad_radiance,
                           grads, partials);
    else
        hipLaunchKernelGGL(k_resolve_backward<false>, dim3(blocks), dim3(64), 0, st, s, vs, params, grad_radiance,
                           grads, partials);
    return hipGetLastError();
}

}  // namespace lhk
