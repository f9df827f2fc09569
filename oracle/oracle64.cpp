// FP64 restatement of the visibility-buffer resolve as a function of its
// differentiable parameters, and its gradient by forward-mode dual numbers.
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker of the hand-written
// adjoint kernel, never part of the product path.
//
// The function is render_geometry main() (src/shaders/render_geometry.comp:70-266)
// with CalcFullBary (src/shaders/common/vbuffer.glsl:11-63) and the BRDF of
// src/shaders/common/pbr.glsl:44-96, evaluated in double precision, with the
// visibility buffer and the per-pixel quantities of orc_frozen (back-face flip,
// PCF shadow factor, texture LOD) taken as constants.  Derivatives are exact
// up to FP64 rounding: every operation propagates the 29 partials of one
// pixel's parameters (3 vertex positions, 3 vertex normals, base-colour,
// roughness and metallic factors, sun direction and intensity).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "texel.h"

#include "oracle.h"

namespace {

constexpr int NP = 29;
enum { P_POS = 0, P_NRM = 9, P_BCF = 18, P_RF = 21, P_MF = 22, P_SUN = 23, P_INT = 26 };
constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- duals

struct Dn {
    double v = 0.0;
    double d[NP] = {};
    Dn() = default;
    Dn(double x) : v(x) {}  // NOLINT: constants promote implicitly
};
Dn var(double x, int i) {
    Dn r(x);
    r.d[i] = 1.0;
    return r;
}
Dn operator+(const Dn& a, const Dn& b) {
    Dn r(a.v + b.v);
    for (int k = 0; k < NP; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
Dn operator-(const Dn& a, const Dn& b) {
    Dn r(a.v - b.v);
    for (int k = 0; k < NP; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
Dn operator-(const Dn& a) {
    Dn r(-a.v);
    for (int k = 0; k < NP; k++) r.d[k] = -a.d[k];
    return r;
}
Dn operator*(const Dn& a, const Dn& b) {
    Dn r(a.v * b.v);
    for (int k = 0; k < NP; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
Dn operator/(const Dn& a, const Dn& b) {
    Dn r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int k = 0; k < NP; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
Dn sqrt(const Dn& a) {
    Dn r(std::sqrt(a.v));
    for (int k = 0; k < NP; k++) r.d[k] = a.d[k] / (2.0 * r.v);
    return r;
}
double value(double x) { return x; }
double value(const Dn& x) { return x.v; }
using std::sqrt;

// clamp / max pass the argument through where it is strictly inside
template <class T>
T clampv(const T& x, double lo, double hi) {
    if (value(x) <= lo) return T(lo);
    if (value(x) >= hi) return T(hi);
    return x;
}

template <class T>
struct V2 {
    T x, y;
};
template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
struct V4 {
    T x, y, z, w;
};
template <class T>
V3<T> operator+(const V3<T>& a, const V3<T>& b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
V3<T> operator-(const V3<T>& a, const V3<T>& b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T, class S>
V3<T> scale(const V3<T>& a, const S& s) {
    return {a.x * s, a.y * s, a.z * s};
}
template <class T>
V3<T> hadamard(const V3<T>& a, const V3<T>& b) {
    return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class T>
T dot(const V3<T>& a, const V3<T>& b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <class T>
V3<T> normalize(const V3<T>& a) {
    T l = sqrt(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}
template <class T>
V3<T> promote(const V3<double>& a) {
    return {T(a.x), T(a.y), T(a.z)};
}

// column-major 4x4 / 3x3 in double
struct M4 {
    double c[4][4];
};
struct M3 {
    double c[3][3];
};
M4 m4(const LhMat4& m) {
    M4 r;
    const LhVec4* c = m.c;
    for (int k = 0; k < 4; k++) {
        r.c[k][0] = c[k].x;
        r.c[k][1] = c[k].y;
        r.c[k][2] = c[k].z;
        r.c[k][3] = c[k].w;
    }
    return r;
}
M3 m3(const LhMat3& m) {
    M3 r;
    for (int k = 0; k < 3; k++) {
        r.c[k][0] = m.c[k].x;
        r.c[k][1] = m.c[k].y;
        r.c[k][2] = m.c[k].z;
    }
    return r;
}
template <class T>
V4<T> mul4(const M4& m, const V4<T>& v) {
    T o[4];
    for (int r = 0; r < 4; r++) o[r] = ((v.x * m.c[0][r] + v.y * m.c[1][r]) + v.z * m.c[2][r]) + v.w * m.c[3][r];
    return {o[0], o[1], o[2], o[3]};
}
template <class T>
V3<T> mul3(const M3& m, const V3<T>& v) {
    T o[3];
    for (int r = 0; r < 3; r++) o[r] = (v.x * m.c[0][r] + v.y * m.c[1][r]) + v.z * m.c[2][r];
    return {o[0], o[1], o[2]};
}

template <class T>
const T* at(uint64_t a) {
    return reinterpret_cast<const T*>(static_cast<uintptr_t>(a));
}

// ---------------------------------------------------------------- inputs

struct Ctx {
    const lh_scene* sc;
    const lh_view* v;
    const orc_params64* p;
};

// FP32 component at host address `addr` of the geometry arena, or its FP64
// replacement.
double geo(const Ctx& c, uint64_t addr) {
    if (c.p->geometry && addr >= c.p->geometry_base && addr + 4 <= c.p->geometry_base + c.p->geometry_bytes)
        return c.p->geometry[(addr - c.p->geometry_base) / 4];
    return double(*at<float>(addr));
}
uint32_t load_index(const MeshInfo& mi, uint32_t i) {
    if (mi.flags & MESH_INFO_FLAGS_32_BIT_INDICES) return at<uint32_t>(mi.indices)[i];
    return at<uint16_t>(mi.indices)[i];
}

// Texel of a level with repeat addressing, as doubles in [0, 1] (sRGB
// formats decoded to linear).
V4<double> texel(const LhTexture& t, uint32_t level, int64_t i, int64_t j) {
    const int64_t w = std::max<int64_t>(1, t.width >> level), h = std::max<int64_t>(1, t.height >> level);
    i %= w;
    if (i < 0) i += w;
    j %= h;
    if (j < 0) j += h;
    const orc_texel::RGBA8 c =
        orc_texel::fetch(t, at<uint8_t>(t.texels), level, uint32_t(w), uint32_t(i), uint32_t(j));
    const uint8_t* q = c.c;
    if (t.format & LH_TEXTURE_SRGB)
        return {orc_texel::srgb_to_linear<double>(q[0]), orc_texel::srgb_to_linear<double>(q[1]),
                orc_texel::srgb_to_linear<double>(q[2]), q[3] / 255.0};
    return {q[0] / 255.0, q[1] / 255.0, q[2] / 255.0, q[3] / 255.0};
}
template <class T>
V4<T> bilinear(const LhTexture& t, uint32_t level, const V2<T>& uv) {
    const double w = double(std::max(1u, t.width >> level)), h = double(std::max(1u, t.height >> level));
    const T u = uv.x * w - 0.5, vv = uv.y * h - 0.5;
    const double fu = std::floor(value(u)), fv = std::floor(value(vv));
    const T a = u - fu, b = vv - fv;
    const int64_t i = int64_t(fu), j = int64_t(fv);
    const V4<double> t00 = texel(t, level, i, j), t10 = texel(t, level, i + 1, j), t01 = texel(t, level, i, j + 1),
                     t11 = texel(t, level, i + 1, j + 1);
    auto lerp2 = [&](double x00, double x10, double x01, double x11) {
        const T top = T(x00) * (1.0 - a) + T(x10) * a;
        const T bot = T(x01) * (1.0 - a) + T(x11) * a;
        return top * (1.0 - b) + bot * b;
    };
    return {lerp2(t00.x, t10.x, t01.x, t11.x), lerp2(t00.y, t10.y, t01.y, t11.y), lerp2(t00.z, t10.z, t01.z, t11.z),
            lerp2(t00.w, t10.w, t01.w, t11.w)};
}
// Trilinear sample with the frozen LOD; an unused texture reads as white.
template <class T>
V4<T> sample(const Ctx& c, uint32_t index, const V2<T>& uv, const orc_frozen& fz, int k) {
    if (index == UNUSED_TEXTURE_INDEX || index >= c.sc->num_textures) return {T(1.0), T(1.0), T(1.0), T(1.0)};
    const LhTexture& t = c.sc->textures[index];
    const V4<T> c0 = bilinear(t, fz.lod_l0[k], uv);
    if (fz.lod_f[k] == 0.0f) return c0;
    const V4<T> c1 = bilinear(t, fz.lod_l1[k], uv);
    const double f = fz.lod_f[k];
    return {c0.x * (1.0 - f) + c1.x * f, c0.y * (1.0 - f) + c1.y * f, c0.z * (1.0 - f) + c1.z * f,
            c0.w * (1.0 - f) + c1.w * f};
}

template <class T>
T pow5(const T& x) {
    const T x2 = x * x;
    return (x2 * x2) * x;
}

// BRDF (pbr.glsl:44-96), thin-walled term absent (alpha clip is unsupported).
template <class T>
V3<T> brdf(const V3<T>& v, const V3<T>& l, const V3<T>& n, const T& pr, const T& metallic, const V3<T>& base) {
    const V3<T> h = normalize(v + l);
    const V3<T> diffuse = scale(base, 1.0 - metallic);
    const double refl = (0.5 / 2.5) * (0.5 / 2.5);
    const T d0 = (1.0 - metallic) * refl;
    const V3<T> f0 = V3<T>{d0, d0, d0} + scale(base, metallic);
    const T NoV = clampv(dot(n, v), 1e-5, 1.0), NoL = clampv(dot(n, l), 0.0, 1.0);
    const T NoH = clampv(dot(n, h), 0.0, 1.0), LoH = clampv(dot(l, h), 0.0, 1.0);
    const T a = pr * pr, a2 = a * a;
    const T fD = (NoH * a2 - NoH) * NoH + 1.0;
    const T D = a2 / ((fD * kPi) * fD);
    const T p = pow5(1.0 - LoH);
    const V3<T> F{f0.x + (1.0 - f0.x) * p, f0.y + (1.0 - f0.y) * p, f0.z + (1.0 - f0.z) * p};
    const T V = 0.5 / (NoL * sqrt((-NoV * a2 + NoV) * NoV + a2) + NoV * sqrt((-NoL * a2 + NoL) * NoL + a2));
    const T f90 = ((a * 2.0) * LoH) * LoH + 0.5;
    const T FdB = ((1.0 + (f90 - 1.0) * pow5(1.0 - NoL)) * (1.0 + (f90 - 1.0) * pow5(1.0 - NoV))) / kPi;
    const V3<T> Fr = scale(F, D * V), Fd = scale(diffuse, FdB);
    V3<T> irr = scale(Fr + Fd, NoL);
    auto fix = [](const T& x) { return (std::isnan(value(x)) || value(x) <= 0.0) ? T(0.0) : x; };
    return {fix(irr.x), fix(irr.y), fix(irr.z)};
}

struct Pixel {
    const MeshInfo* mi;
    uint32_t mesh_info;
    uint64_t pos_addr[3], nrm_addr[3];
    M4 T4, PV;
    M3 N3;
    V3<double> cam;
    double ndc_x, ndc_y;
    V2<double> uvs[3];
    double off_x, off_y;
};

bool decode(const Ctx& c, uint32_t x, uint32_t y, Pixel& px) {
    const lh_view& v = *c.v;
    const uint64_t key = v.visbuffer[size_t(y) * v.width + x];
    if (~uint32_t(key >> 32) == 0u) return false;
    const uint32_t mid = uint32_t(key >> LH_KEY_TRI_BITS) & (LH_MAX_SCENE_MESHLETS - 1);
    const uint32_t tri = uint32_t(key) & ((1u << LH_KEY_TRI_BITS) - 1);
    const MeshletReference& ref = c.sc->meshlet_table[mid];
    const Instance& in = c.sc->instances[ref.instance_index];
    const MeshInfo& mi = *at<MeshInfo>(in.mesh_info_address);
    const Meshlet& ml = at<Meshlet>(mi.meshlets)[ref.meshlet_index];
    px.mi = &mi;
    px.mesh_info = uint32_t((in.mesh_info_address - uint64_t(uintptr_t(c.sc->mesh_infos))) / sizeof(MeshInfo));
    for (int j = 0; j < 3; j++) {
        const uint8_t m = at<uint8_t>(mi.micro_indices)[size_t(ml.triangle_offset) + 3 * tri + j];
        const uint32_t vi = load_index(mi, ml.index_offset + m);
        px.pos_addr[j] = mi.positions + 12ull * vi;
        px.nrm_addr[j] = mi.normals + 12ull * vi;
        const float* uvp = at<float>(mi.uvs) + 2 * size_t(vi);
        px.uvs[j] = {double(uvp[0]) * mi.texture_scale.x, double(uvp[1]) * mi.texture_scale.y};
    }
    const Uniforms& u = *v.uniforms;
    px.T4 = m4(in.transform);
    px.PV = m4(u.combined_perspective_view);
    px.N3 = m3(in.normal_transform);
    px.cam = {u.camera_pos.x, u.camera_pos.y, u.camera_pos.z};
    px.ndc_x = double(x) / double(u.window_size.x) * 2.0 - 1.0;
    px.ndc_y = double(y) / double(u.window_size.y) * 2.0 - 1.0;
    px.off_x = mi.texture_offset.x;
    px.off_y = mi.texture_offset.y;
    return true;
}

// The resolve of one pixel as a function of its parameters (all of type T).
template <class T>
V3<T> shade(const Ctx& c, const Pixel& px, const orc_frozen& fz, const V3<T> L[3], const V3<T> Nr[3],
            const V3<T>& bcf, const T& rf, const T& mf, const V3<T>& sun_dir, const V3<T>& sun_int) {
    V3<T> pos[3];
    V4<T> clip[3];
    for (int j = 0; j < 3; j++) {
        const V4<T> w = mul4(px.T4, V4<T>{L[j].x, L[j].y, L[j].z, T(1.0)});
        pos[j] = {w.x, w.y, w.z};
        clip[j] = mul4(px.PV, V4<T>{w.x, w.y, w.z, T(1.0)});
    }
    // CalcFullBary (vbuffer.glsl:11-63), barycentric part
    const T iw0 = 1.0 / clip[0].w, iw1 = 1.0 / clip[1].w, iw2 = 1.0 / clip[2].w;
    const V2<T> n0{clip[0].x * iw0, clip[0].y * iw0}, n1{clip[1].x * iw1, clip[1].y * iw1},
        n2{clip[2].x * iw2, clip[2].y * iw2};
    const T invDet = 1.0 / ((n2.x - n1.x) * (n0.y - n1.y) - (n0.x - n1.x) * (n2.y - n1.y));
    const T ddx0 = (n1.y - n2.y) * invDet * iw0, ddx1 = (n2.y - n0.y) * invDet * iw1, ddx2 = (n0.y - n1.y) * invDet * iw2;
    const T ddy0 = (n2.x - n1.x) * invDet * iw0, ddy1 = (n0.x - n2.x) * invDet * iw1, ddy2 = (n1.x - n0.x) * invDet * iw2;
    const T dvx = px.ndc_x - n0.x, dvy = px.ndc_y - n0.y;
    const T interpInvW = iw0 + dvx * ((ddx0 + ddx1) + ddx2) + dvy * ((ddy0 + ddy1) + ddy2);
    const T lam[3] = {(iw0 + dvx * ddx0 + dvy * ddy0) / interpInvW, (dvx * ddx1 + dvy * ddy1) / interpInvW,
                      (dvx * ddx2 + dvy * ddy2) / interpInvW};
    const V3<T> world = (scale(pos[0], lam[0]) + scale(pos[1], lam[1])) + scale(pos[2], lam[2]);
    const V2<T> uv{(lam[0] * px.uvs[0].x + lam[1] * px.uvs[1].x) + lam[2] * px.uvs[2].x + px.off_x,
                   (lam[0] * px.uvs[0].y + lam[1] * px.uvs[1].y) + lam[2] * px.uvs[2].y + px.off_y};
    const V3<T> view_vector = normalize(promote<T>(px.cam) - world);
    const V3<T> nI = (scale(normalize(Nr[0]), lam[0]) + scale(normalize(Nr[1]), lam[1])) + scale(normalize(Nr[2]), lam[2]);
    const V3<T> nrm = scale(normalize(mul3(px.N3, nI)), double(fz.flip));
    const V4<T> bc = sample(c, px.mi->base_color_texture_index, uv, fz, 0);
    const V3<T> base = hadamard(V3<T>{bc.x, bc.y, bc.z}, bcf);
    const V4<T> mr = sample(c, px.mi->metallic_roughness_texture_index, uv, fz, 1);
    const T roughness = mr.y * rf, metallic = mr.z * mf;
    const T shv = nrm.y * 0.025 + 0.025;  // SH ambient (render_geometry.comp:40-57): constant + y band
    const V3<T> ambient = scale(base, (1.0 - metallic) * shv);
    const V3<T> B = brdf(view_vector, sun_dir, nrm, roughness, metallic, base);
    const double K = kPi * double(fz.shadow);
    return hadamard(scale(B, K), sun_int) + ambient;
}

struct PixelParams {
    double L[3][3], Nr[3][3], bcf[3], rf, mf, sun[3], inten[3];
};

void gather(const Ctx& c, const Pixel& px, PixelParams& q) {
    const orc_params64& p = *c.p;
    const MeshInfo& mi = *px.mi;
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
            q.L[j][k] = geo(c, px.pos_addr[j] + 4 * k);
            q.Nr[j][k] = geo(c, px.nrm_addr[j] + 4 * k);
        }
    const uint32_t m = px.mesh_info;
    const Uniforms& u = *c.v->uniforms;
    const double mi_bcf[3] = {mi.base_color_factor.x, mi.base_color_factor.y, mi.base_color_factor.z};
    for (int k = 0; k < 3; k++) {
        q.bcf[k] = p.base_color_factor ? p.base_color_factor[3 * m + k] : mi_bcf[k];
        const double sd[3] = {u.sun_dir.x, u.sun_dir.y, u.sun_dir.z}, si[3] = {u.sun_intensity.x, u.sun_intensity.y,
                                                                              u.sun_intensity.z};
        q.sun[k] = p.sun_dir ? p.sun_dir[k] : sd[k];
        q.inten[k] = p.sun_intensity ? p.sun_intensity[k] : si[k];
    }
    q.rf = p.roughness_factor ? p.roughness_factor[m] : c.sc->factors[m].roughness_factor;
    q.mf = p.metallic_factor ? p.metallic_factor[m] : c.sc->factors[m].metallic_factor;
}

bool valid(const lh_scene* sc, const lh_view* v, const orc_params64* p, const orc_frozen* f) {
    return sc && v && p && f && v->visbuffer && v->uniforms && v->misc;
}

}  // namespace

extern "C" {

lh_status orc64_resolve(const lh_scene* sc, const lh_view* v, const orc_params64* p, const orc_frozen* frozen,
                        double* rgb) {
    if (!valid(sc, v, p, frozen) || !rgb) return LH_ERR_INVALID;
    const Ctx c{sc, v, p};
    for (uint32_t y = 0; y < v->height; y++)
        for (uint32_t x = 0; x < v->width; x++) {
            const size_t o = size_t(y) * v->width + x;
            double* out = rgb + 3 * o;
            out[0] = out[1] = out[2] = 0.0;
            Pixel px;
            if (frozen[o].kind != 1 || !decode(c, x, y, px)) continue;
            PixelParams q;
            gather(c, px, q);
            V3<double> L[3], Nr[3];
            for (int j = 0; j < 3; j++) {
                L[j] = {q.L[j][0], q.L[j][1], q.L[j][2]};
                Nr[j] = {q.Nr[j][0], q.Nr[j][1], q.Nr[j][2]};
            }
            const V3<double> r = shade<double>(c, px, frozen[o], L, Nr, {q.bcf[0], q.bcf[1], q.bcf[2]}, q.rf, q.mf,
                                               {q.sun[0], q.sun[1], q.sun[2]}, {q.inten[0], q.inten[1], q.inten[2]});
            out[0] = r.x;
            out[1] = r.y;
            out[2] = r.z;
        }
    return LH_OK;
}

lh_status orc64_resolve_grad(const lh_scene* sc, const lh_view* v, const orc_params64* p, const orc_frozen* frozen,
                             const float* grad_radiance, double* g_bcf, double* g_rf, double* g_mf, double* g_sun,
                             double* g_int, double* g_geo) {
    if (!valid(sc, v, p, frozen) || !grad_radiance) return LH_ERR_INVALID;
    const Ctx c{sc, v, p};
    for (uint32_t y = 0; y < v->height; y++)
        for (uint32_t x = 0; x < v->width; x++) {
            const size_t o = size_t(y) * v->width + x;
            Pixel px;
            if (frozen[o].kind != 1 || !decode(c, x, y, px)) continue;
            PixelParams q;
            gather(c, px, q);
            V3<Dn> L[3], Nr[3];
            for (int j = 0; j < 3; j++) {
                L[j] = {var(q.L[j][0], P_POS + 3 * j), var(q.L[j][1], P_POS + 3 * j + 1),
                        var(q.L[j][2], P_POS + 3 * j + 2)};
                Nr[j] = {var(q.Nr[j][0], P_NRM + 3 * j), var(q.Nr[j][1], P_NRM + 3 * j + 1),
                         var(q.Nr[j][2], P_NRM + 3 * j + 2)};
            }
            const V3<Dn> r = shade<Dn>(c, px, frozen[o], L, Nr,
                                       {var(q.bcf[0], P_BCF), var(q.bcf[1], P_BCF + 1), var(q.bcf[2], P_BCF + 2)},
                                       var(q.rf, P_RF), var(q.mf, P_MF),
                                       {var(q.sun[0], P_SUN), var(q.sun[1], P_SUN + 1), var(q.sun[2], P_SUN + 2)},
                                       {var(q.inten[0], P_INT), var(q.inten[1], P_INT + 1), var(q.inten[2], P_INT + 2)});
            const double* gr = nullptr;
            double gbuf[3] = {grad_radiance[4 * o], grad_radiance[4 * o + 1], grad_radiance[4 * o + 2]};
            gr = gbuf;
            double dl[NP];
            for (int k = 0; k < NP; k++) dl[k] = (gr[0] * r.x.d[k] + gr[1] * r.y.d[k]) + gr[2] * r.z.d[k];
            const uint32_t m = px.mesh_info;
            if (g_geo && p->geometry_bytes) {
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 3; k++) {
                        const uint64_t pa = px.pos_addr[j] + 4 * k, na = px.nrm_addr[j] + 4 * k;
                        if (pa >= p->geometry_base && pa + 4 <= p->geometry_base + p->geometry_bytes)
                            g_geo[(pa - p->geometry_base) / 4] += dl[P_POS + 3 * j + k];
                        if (na >= p->geometry_base && na + 4 <= p->geometry_base + p->geometry_bytes)
                            g_geo[(na - p->geometry_base) / 4] += dl[P_NRM + 3 * j + k];
                    }
            }
            for (int k = 0; k < 3; k++) {
                if (g_bcf) g_bcf[3 * m + k] += dl[P_BCF + k];
                if (g_sun) g_sun[k] += dl[P_SUN + k];
                if (g_int) g_int[k] += dl[P_INT + k];
            }
            if (g_rf) g_rf[m] += dl[P_RF];
            if (g_mf) g_mf[m] += dl[P_MF];
        }
    return LH_OK;
}

}  // extern "C"
