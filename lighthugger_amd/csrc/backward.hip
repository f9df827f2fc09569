// Adjoint of the visibility-buffer resolve (gfx950).
//
// The forward is the per-pixel function of render_geometry.comp:70-266 as
// implemented by shading::resolve_pixel; the adjoint differentiates it with
// respect to the material factors, the sun direction and intensity, and the
// mesh-space vertex positions and normals of the triangle each pixel shows.
// Held fixed, like the visibility buffer itself: the back-face flip, the
// cascade choice and PCF shadow factor, and the texture LOD (levels and blend
// weight), all recomputed here exactly as the forward computes them.
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
// gradients are summed over the wave per distinct mesh info and added once;
// the six light gradients stay in registers and every block writes one
// partial, which a second one-block kernel sums in a fixed order.
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
LH_DEV d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
LH_DEV d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
LH_DEV d3 operator*(d3 a, d3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LH_DEV double dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
LH_DEV d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LH_DEV double len(d3 a) { return sqrt(dot(a, a)); }
LH_DEV d3 normalize(d3 a) { return a * (1.0 / len(a)); }
LH_DEV double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
LH_DEV double pow4(double x) {
    const double x2 = x * x;
    return x2 * x2;
}
LH_DEV double pow5(double x) { return pow4(x) * x; }
// columns 0..2 of a 4x4 (column-major) applied to a point / the linear part
LH_DEV d3 col3(const LhVec4& c) { return {double(c.x), double(c.y), double(c.z)}; }

// u = a / |a|: the gradient of a from the gradient of u.
LH_DEV d3 normalize_bwd(d3 a, d3 gu) {
    const double l = len(a);
    const d3 u = a * (1.0 / l);
    return (gu - u * dot(u, gu)) * (1.0 / l);
}

struct BrdfGrad {
    d3 v, l, n, base;
    double pr, metallic;
};

// BRDF of shading::brdf (thin = false) in FP64: value and reverse mode for the
// upstream gradient g of its RGB output.
LH_DEV d3 brdf_bwd(d3 v, d3 l, d3 n, double pr, double m, d3 base, d3 g, BrdfGrad& o) {
    const d3 hv = v + l;
    const d3 h = normalize(hv);
    const d3 diffuse = base * (1.0 - m);
    const double refl = (0.5 / 2.5) * (0.5 / 2.5);
    const double d0 = refl * (1.0 - m);
    const d3 f0 = d3{d0, d0, d0} + base * m;
    const double nv = dot(n, v), nl = dot(n, l), nh = dot(n, h), lh = dot(l, h);
    const double NoV = clampd(nv, 1e-5, 1.0), NoL = clampd(nl, 0.0, 1.0);
    const double NoH = clampd(nh, 0.0, 1.0), LoH = clampd(lh, 0.0, 1.0);
    const double a = pr * pr, a2 = a * a;
    const double fD = (NoH * a2 - NoH) * NoH + 1.0;
    const double D = a2 / ((M_PI * fD) * fD);
    const double p = pow5(1.0 - LoH);
    const d3 F{f0.x + (1.0 - f0.x) * p, f0.y + (1.0 - f0.y) * p, f0.z + (1.0 - f0.z) * p};
    const double sL = sqrt((-NoL * a2 + NoL) * NoL + a2), sV = sqrt((-NoV * a2 + NoV) * NoV + a2);
    const double den = NoL * sV + NoV * sL;
    const double Vis = 0.5 / den;
    const double f90 = 0.5 + ((2.0 * a) * LoH) * LoH;
    const double q5L = pow5(1.0 - NoL), q5V = pow5(1.0 - NoV);
    const double FSL = 1.0 + (f90 - 1.0) * q5L, FSV = 1.0 + (f90 - 1.0) * q5V;
    const double FdB = (FSL * FSV) * (1.0 / M_PI);
    const double DV = D * Vis;
    const d3 Fr = F * DV, Fd = diffuse * FdB;
    const d3 irr = (Fd + Fr) * NoL;
    // max(irr, 0) and the NaN guard pass the gradient only where irr > 0
    const d3 gi{irr.x > 0.0 ? g.x : 0.0, irr.y > 0.0 ? g.y : 0.0, irr.z > 0.0 ? g.z : 0.0};
    const d3 gs = gi * NoL;
    double gNoL = dot(gi, Fd + Fr);
    const d3 gF = gs * DV;
    const double gDV = dot(gs, F);
    const double gD = gDV * Vis, gVis = gDV * D;
    const d3 gdiff = gs * FdB;
    const double gFdB = dot(gs, diffuse);
    const double gFSL = gFdB * FSV * (1.0 / M_PI), gFSV = gFdB * FSL * (1.0 / M_PI);
    const double gf90 = gFSL * q5L + gFSV * q5V;
    gNoL += gFSL * (f90 - 1.0) * (-5.0 * pow4(1.0 - NoL));
    double gNoV = gFSV * (f90 - 1.0) * (-5.0 * pow4(1.0 - NoV));
    double ga = gf90 * (2.0 * LoH * LoH);
    double gLoH = gf90 * (4.0 * a * LoH);
    const double gden = -gVis * Vis / den;
    gNoV += gden * sL;
    gNoL += gden * sV;
    const double gsL = gden * NoV, gsV = gden * NoL;
    gNoL += gsL * NoL * (1.0 - a2) / sL;
    gNoV += gsV * NoV * (1.0 - a2) / sV;
    double ga2 = gsL * (1.0 - NoL * NoL) / (2.0 * sL) + gsV * (1.0 - NoV * NoV) / (2.0 * sV);
    const d3 gf0 = gF * (1.0 - p);
    const double gp = (gF.x * (1.0 - f0.x) + gF.y * (1.0 - f0.y)) + gF.z * (1.0 - f0.z);
    gLoH += gp * (-5.0 * pow4(1.0 - LoH));
    const double E = D / fD;
    ga2 += gD * (1.0 / ((M_PI * fD) * fD) - 2.0 * E * NoH * NoH);
    const double gNoH = gD * (-2.0 * E * (2.0 * NoH * (a2 - 1.0)));
    ga += ga2 * 2.0 * a;
    o.pr = ga * 2.0 * pr;
    o.metallic = dot(gf0, base - d3{refl, refl, refl}) - dot(gdiff, base);
    o.base = gf0 * m + gdiff * (1.0 - m);
    const double gnv = (nv > 1e-5 && nv < 1.0) ? gNoV : 0.0;
    const double gnl = (nl > 0.0 && nl < 1.0) ? gNoL : 0.0;
    const double gnh = (nh > 0.0 && nh < 1.0) ? gNoH : 0.0;
    const double glh = (lh > 0.0 && lh < 1.0) ? gLoH : 0.0;
    o.n = (v * gnv + l * gnl) + h * gnh;
    const d3 ghv = normalize_bwd(hv, n * gnh + l * glh);
    o.v = n * gnv + ghv;
    o.l = (n * gnl + h * glh) + ghv;
    return {irr.x > 0.0 ? irr.x : 0.0, irr.y > 0.0 ? irr.y : 0.0, irr.z > 0.0 ? irr.z : 0.0};
}

// ---- textures in FP64 with the forward's (FP32) LOD held fixed ----
struct Texel4 {
    double c[4];
};
LH_DEV Texel4 texel64(const LhTexture& t, uint32_t level, int32_t i, int32_t j) {
    const int32_t w = int32_t(max(1u, t.width >> level)), h = int32_t(max(1u, t.height >> level));
    i %= w;
    if (i < 0) i += w;
    j %= h;
    if (j < 0) j += h;
    const uint32_t q = ptr<uint32_t>(t.texels)[size_t(t.level_offset[level]) + size_t(j) * w + i];
    return {{double(q & 255u) / 255.0, double((q >> 8) & 255u) / 255.0, double((q >> 16) & 255u) / 255.0,
             double(q >> 24) / 255.0}};
}
// Bilinear value (val) and d/d(uv) contracted with g, at one level.
LH_DEV void bilinear64(const LhTexture& t, uint32_t level, d2 uv, const double* g, double* val, d2& guv) {
    const double w = double(max(1u, t.width >> level)), h = double(max(1u, t.height >> level));
    const double u = uv.x * w - 0.5, vv = uv.y * h - 0.5;
    const double fu = floor(u), fv = floor(vv);
    const double a = u - fu, b = vv - fv;
    const int32_t i = int32_t(fu), j = int32_t(fv);
    const Texel4 t00 = texel64(t, level, i, j), t10 = texel64(t, level, i + 1, j);
    const Texel4 t01 = texel64(t, level, i, j + 1), t11 = texel64(t, level, i + 1, j + 1);
    double ga = 0.0, gb = 0.0;
    for (int k = 0; k < 4; k++) {
        const double top = t00.c[k] * (1.0 - a) + t10.c[k] * a, bot = t01.c[k] * (1.0 - a) + t11.c[k] * a;
        val[k] = top * (1.0 - b) + bot * b;
        ga += g[k] * ((1.0 - b) * (t10.c[k] - t00.c[k]) + b * (t11.c[k] - t01.c[k]));
        gb += g[k] * (bot - top);
    }
    guv = {ga * w, gb * h};
}
// Trilinear sample with the frozen LOD; the uv gradient is for the upstream g.
// Unused index: white, no gradient.
LH_DEV void sample64(const lh_scene& s, uint32_t index, const Lod& lod, d2 uv, const double* g, double* val, d2& guv) {
    guv = {0.0, 0.0};
    if (!has_texture(s, index)) {
        for (int k = 0; k < 4; k++) val[k] = 1.0;
        return;
    }
    const LhTexture& t = s.textures[index];
    bilinear64(t, lod.l0, uv, g, val, guv);
    if (lod.f == 0.0f) return;
    double v1[4];
    d2 g1;
    bilinear64(t, lod.l1, uv, g, v1, g1);
    const double f = lod.f;
    for (int k = 0; k < 4; k++) val[k] = val[k] * (1.0 - f) + v1[k] * f;
    guv = {guv.x * (1.0 - f) + g1.x * f, guv.y * (1.0 - f) + g1.y * f};
}

struct LightGrad {
    d3 sun_dir, sun_intensity;
};

// Vertex gradients of one tile are merged in an LDS hash table keyed by the
// gradient-arena word of the vec3 (positions and normals alike) and flushed
// with one atomic per distinct word: neighbouring pixels share vertices, and
// same-address atomics from one wave serialize.  A tile inserts at most
// 64 pixels x 6 vec3 = 384 keys, so the 512-slot table never fills.
constexpr uint32_t kHashSlots = 512;
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;
struct VertexHash {
    uint32_t* keys;  // [kHashSlots]
    float* vals;     // [kHashSlots * 3]
};

LH_DEV void add_geometry(const lh_resolve_grads& gr, const VertexHash& h, uint64_t addr, d3 g) {
    if (!gr.geometry || addr < gr.geometry_base) return;
    const uint64_t k64 = (addr - gr.geometry_base) / 4;
    if (k64 + 3 > gr.geometry_bytes / 4 || k64 >= kHashEmpty) return;
    const uint32_t k = uint32_t(k64);
    uint32_t slot = (k * 2654435761u) >> 23;  // 9 bits: kHashSlots == 512
    while (true) {
        const uint32_t old = atomicCAS(h.keys + slot, kHashEmpty, k);
        if (old == kHashEmpty || old == k) break;
        slot = (slot + 1) & (kHashSlots - 1);
    }
    atomicAdd(h.vals + 3 * slot, float(g.x));
    atomicAdd(h.vals + 3 * slot + 1, float(g.y));
    atomicAdd(h.vals + 3 * slot + 2, float(g.z));
}

// Called by the whole wave after a tile: adds every merged vec3 to the arena
// and empties the table.
LH_DEV void flush_geometry(const lh_resolve_grads& gr, const VertexHash& h) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHashSlots; i += 64) {
        const uint32_t k = h.keys[i];
        if (k == kHashEmpty) continue;
        atomicAdd(gr.geometry + k, h.vals[3 * i]);
        atomicAdd(gr.geometry + k + 1, h.vals[3 * i + 1]);
        atomicAdd(gr.geometry + k + 2, h.vals[3 * i + 2]);
        h.keys[i] = kHashEmpty;
        h.vals[3 * i] = h.vals[3 * i + 1] = h.vals[3 * i + 2] = 0.0f;
    }
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
// light gradients to lg.  The function and its reverse are evaluated in FP64
// from the FP32 inputs.
LH_DEV void pixel_bwd(const lh_scene& s, const lh_view& v, const lh_resolve_params& prm, const float* grad_radiance,
                      const lh_resolve_grads& gr, const VertexHash& hash, uint32_t x, uint32_t y, uint32_t& mesh_info,
                      d3& g_bcf, double& g_rf, double& g_mf, LightGrad& lg) {
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
    if (prm.frozen) {  // recorded by the forward (lh_resolve_ex with the same buffer)
        const uint4 r = reinterpret_cast<const uint4*>(prm.frozen)[o];
        fz.flip = (r.w & 1u) ? -1.0f : 1.0f;
        fz.K = kPi * __uint_as_float(r.x);
        auto lod = [&](uint32_t index, uint32_t l0, uint32_t fbits) {
            if (!has_texture(s, index)) return Lod{0, 0, 0.0f};
            return Lod{l0, min(l0 + 1, s.textures[index].levels - 1), __uint_as_float(fbits)};
        };
        fz.lod_b = lod(mi.base_color_texture_index, (r.w >> 8) & 255u, r.y);
        fz.lod_m = lod(mi.metallic_roughness_texture_index, (r.w >> 16) & 255u, r.z);
    } else {
        const f2 win{float(u.window_size.x), float(u.window_size.y)};
        const f2 ndc{float(x) / win.x * 2.0f - 1.0f, float(y) / win.y * 2.0f - 1.0f};
        fz = frozen(s, v, mi, sf, ndc, win);
    }
    const float4 gin = reinterpret_cast<const float4*>(grad_radiance)[o];
    const d3 g{gin.x, gin.y, gin.z};
    // ---- forward in FP64 ----
    const LhMat4& T = in.transform;
    const LhMat4& PV = u.combined_perspective_view;
    d3 pos[3], c[3], nraw[3], nn[3];
    d2 uvs[3];
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
        nraw[j] = D3(load_normal(mi, sf.vi[j]));
        nn[j] = normalize(nraw[j]);
        const f2 t = load_uv(mi, sf.vi[j]);
        uvs[j] = {double(t.x) * double(mi.texture_scale.x), double(t.y) * double(mi.texture_scale.y)};
    }
    // perspective-correct barycentrics: lambda_j = D_j / sum D, D_j = q . (c_{j+1} x c_{j+2})
    const d3 q{double(x) / double(u.window_size.x) * 2.0 - 1.0, double(y) / double(u.window_size.y) * 2.0 - 1.0, 1.0};
    double Dk[3];
    for (int k = 0; k < 3; k++) Dk[k] = dot(q, cross(c[(k + 1) % 3], c[(k + 2) % 3]));
    const double sumD = (Dk[0] + Dk[1]) + Dk[2];
    const double lam[3] = {Dk[0] / sumD, Dk[1] / sumD, Dk[2] / sumD};
    const d3 world = (pos[0] * lam[0] + pos[1] * lam[1]) + pos[2] * lam[2];
    const d2 uv{((lam[0] * uvs[0].x + lam[1] * uvs[1].x) + lam[2] * uvs[2].x) + double(mi.texture_offset.x),
                ((lam[0] * uvs[0].y + lam[1] * uvs[1].y) + lam[2] * uvs[2].y) + double(mi.texture_offset.y)};
    const d3 to_cam = D3(u.camera_pos) - world;
    const d3 V = normalize(to_cam);
    const d3 nI = (nn[0] * lam[0] + nn[1] * lam[1]) + nn[2] * lam[2];
    const LhMat3& N = in.normal_transform;
    const d3 m3 = (D3(N.c[0]) * nI.x + D3(N.c[1]) * nI.y) + D3(N.c[2]) * nI.z;
    const d3 nrm = normalize(m3) * double(fz.flip);
    const d3 bcf = D3(param_base_color(prm, mi, sf.mesh_info));
    const double rf = param_roughness(prm, s, sf.mesh_info), mf = param_metallic(prm, s, sf.mesh_info);
    const d3 sun_dir = D3(param_sun_dir(prm, u)), sun_int = D3(param_sun_intensity(prm, u));
    // the texture samples are needed before their upstream gradients are known:
    // take the values first (zero upstream), the uv gradients after the reverse
    const double zero4[4] = {0.0, 0.0, 0.0, 0.0};
    double bc[4], mr[4];
    d2 unused;
    sample64(s, mi.base_color_texture_index, fz.lod_b, uv, zero4, bc, unused);
    sample64(s, mi.metallic_roughness_texture_index, fz.lod_m, uv, zero4, mr, unused);
    const d3 base = d3{bc[0], bc[1], bc[2]} * bcf;
    const double roughness = mr[1] * rf, metallic = mr[2] * mf;
    const double shv = nrm.y * 0.025 + 0.025;  // SH ambient: constant and y bands
    const double K = fz.K;
    // ---- reverse ----
    // out = B * K * I + base * (1 - metallic) * shv;  B needs the upstream gB = g * I * K
    BrdfGrad bg;
    const d3 B = brdf_bwd(V, sun_dir, nrm, roughness, metallic, base, g * sun_int * K, bg);
    lg.sun_intensity = lg.sun_intensity + g * B * K;
    lg.sun_dir = lg.sun_dir + bg.l;
    const d3 gbase = g * ((1.0 - metallic) * shv) + bg.base;
    const double gmet = -dot(g, base) * shv + bg.metallic;
    d3 gnrm = bg.n;
    gnrm.y += dot(g, base) * (1.0 - metallic) * 0.025;
    mesh_info = sf.mesh_info;
    g_bcf = gbase * d3{bc[0], bc[1], bc[2]};
    g_rf = bg.pr * mr[1];
    g_mf = gmet * mr[2];
    const double gb4[4] = {gbase.x * bcf.x, gbase.y * bcf.y, gbase.z * bcf.z, 0.0};
    const double gm4[4] = {0.0, bg.pr * rf, gmet * mf, 0.0};
    double tmp[4];
    d2 guv_b, guv_m;
    sample64(s, mi.base_color_texture_index, fz.lod_b, uv, gb4, tmp, guv_b);
    sample64(s, mi.metallic_roughness_texture_index, fz.lod_m, uv, gm4, tmp, guv_m);
    const d2 guv{guv_b.x + guv_m.x, guv_b.y + guv_m.y};
    // normal chain: nrm = flip * normalize(N nI), nI = sum lambda_j normalize(nraw_j)
    const d3 gm3 = normalize_bwd(m3, gnrm * double(fz.flip));
    const d3 gnI{dot(D3(N.c[0]), gm3), dot(D3(N.c[1]), gm3), dot(D3(N.c[2]), gm3)};
    // world position through the view vector (shadow and LOD are fixed)
    const d3 gworld = -normalize_bwd(to_cam, bg.v);
    // positions relative to vertex 0: lambda sums to 1, so a common shift of
    // the lambda gradients cancels and the relative form is better conditioned
    double glam[3];
    d3 gpos[3];
    for (int j = 0; j < 3; j++) {
        glam[j] = (dot(gworld, pos[j] - pos[0]) + dot(gnI, nn[j])) +
                  (guv.x * (uvs[j].x - uvs[0].x) + guv.y * (uvs[j].y - uvs[0].y));
        gpos[j] = gworld * lam[j];
        add_geometry(gr, hash, mi.normals + 12ull * sf.vi[j], normalize_bwd(nraw[j], gnI * lam[j]));
    }
    const double lg_dot = (glam[0] * lam[0] + glam[1] * lam[1]) + glam[2] * lam[2];
    d3 gc[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 3; k++) {
        const double gD = (glam[k] - lg_dot) / sumD;
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        gc[a] = gc[a] + cross(c[b], q) * gD;
        gc[b] = gc[b] + cross(q, c[a]) * gD;
    }
    for (int j = 0; j < 3; j++) {
        // clip (x, y, w) <- pos: d clip_r / d pos_k = PV.c[k][r]
        for (int k = 0; k < 3; k++) {
            const LhVec4& cc = PV.c[k];
            const double d = (double(cc.x) * gc[j].x + double(cc.y) * gc[j].y) + double(cc.w) * gc[j].z;
            if (k == 0) gpos[j].x += d;
            if (k == 1) gpos[j].y += d;
            if (k == 2) gpos[j].z += d;
        }
        const d3 gL{dot(col3(T.c[0]), gpos[j]), dot(col3(T.c[1]), gpos[j]), dot(col3(T.c[2]), gpos[j])};
        add_geometry(gr, hash, mi.positions + 12ull * sf.vi[j], gL);
    }
}

// Occupancy target (waves per SIMD) of the adjoint; 0 leaves the register
// budget to the compiler.
#ifndef LH_BACKWARD_WAVES
#define LH_BACKWARD_WAVES 0
#endif
#if LH_BACKWARD_WAVES > 0
#define LH_BACKWARD_OCCUPANCY LH_WAVES_ATTR(LH_BACKWARD_WAVES)
#else
#define LH_BACKWARD_OCCUPANCY
#endif

__global__ void __launch_bounds__(64) LH_BACKWARD_OCCUPANCY k_resolve_backward(lh_scene s, OneView vs, lh_resolve_params prm,
                                                         const float* grad_radiance, lh_resolve_grads gr,
                                                         double* partials) {
    __shared__ uint32_t hkeys[kHashSlots];
    __shared__ float hvals[kHashSlots * 3];
    const VertexHash hash{hkeys, hvals};
    for (uint32_t i = threadIdx.x; i < kHashSlots; i += 64) {
        hkeys[i] = kHashEmpty;
        hvals[3 * i] = hvals[3 * i + 1] = hvals[3 * i + 2] = 0.0f;
    }
    __syncthreads();
    const lh_view& v = vs.v;
    const uint32_t tiles_x = (v.width + 7) / 8, tiles = tiles_x * ((v.height + 7) / 8);
    LightGrad lg{{0, 0, 0}, {0, 0, 0}};
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint32_t x = (t % tiles_x) * 8 + (threadIdx.x & 7), y = (t / tiles_x) * 8 + (threadIdx.x >> 3);
        uint32_t mesh_info = ~0u;
        d3 g_bcf{0, 0, 0};
        double g_rf = 0.0, g_mf = 0.0;
        if (x < v.width && y < v.height)
            pixel_bwd(s, v, prm, grad_radiance, gr, hash, x, y, mesh_info, g_bcf, g_rf, g_mf, lg);
        if (gr.geometry) flush_geometry(gr, hash);
        // material factors: one add per distinct mesh info of the wave
        uint64_t pending = ballot(mesh_info != ~0u);
        while (pending) {
            const uint32_t lead = __builtin_amdgcn_readlane(mesh_info, int(first_lane(pending)));
            const bool mine = mesh_info == lead;
            double r[5] = {mine ? g_bcf.x : 0.0, mine ? g_bcf.y : 0.0, mine ? g_bcf.z : 0.0, mine ? g_rf : 0.0,
                           mine ? g_mf : 0.0};
            for (int i = 0; i < 5; i++) {
                double sum = r[i];
                for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
                r[i] = sum;
            }
            if (lane_id() == first_lane(pending)) {
                if (gr.base_color_factor) {
                    atomicAdd(gr.base_color_factor + 3 * lead, float(r[0]));
                    atomicAdd(gr.base_color_factor + 3 * lead + 1, float(r[1]));
                    atomicAdd(gr.base_color_factor + 3 * lead + 2, float(r[2]));
                }
                if (gr.roughness_factor) atomicAdd(gr.roughness_factor + lead, float(r[3]));
                if (gr.metallic_factor) atomicAdd(gr.metallic_factor + lead, float(r[4]));
            }
            pending &= ~ballot(mine);
        }
    }
    double r[6] = {lg.sun_dir.x, lg.sun_dir.y, lg.sun_dir.z, lg.sun_intensity.x, lg.sun_intensity.y,
                   lg.sun_intensity.z};
    for (int i = 0; i < 6; i++) {
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
}

// Sums the per-block light partials with a fixed reduction tree (the result
// does not depend on timing) and adds them to the outputs: 256 threads each
// sum a strided subset per component, then a tree in LDS.
__global__ void __launch_bounds__(256) k_reduce_light(const double* partials, uint32_t blocks, lh_resolve_grads gr) {
    __shared__ double acc[6][256];
    const uint32_t t = threadIdx.x;
    for (int i = 0; i < 6; i++) {
        double sum = 0.0;
        for (uint32_t b = t; b < blocks; b += 256) sum += partials[size_t(b) * 8 + i];
        acc[i][t] = sum;
    }
    __syncthreads();
    for (uint32_t w = 128; w >= 1; w >>= 1) {
        if (t < w)
            for (int i = 0; i < 6; i++) acc[i][t] += acc[i][t + w];
        __syncthreads();
    }
    if (t < 6) {
        float* dst = t < 3 ? gr.sun_dir : gr.sun_intensity;
        if (dst) dst[t % 3] += float(acc[t][0]);
    }
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
    hipLaunchKernelGGL(k_resolve_backward, dim3(blocks), dim3(64), 0, st, s, vs, params, grad_radiance, grads,
                       partials);
    mark("reduce_light", st);
    hipLaunchKernelGGL(k_reduce_light, dim3(1), dim3(256), 0, st, partials, blocks, grads);
    return hipGetLastError();
}

}  // namespace lhk
