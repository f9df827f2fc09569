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
// Per pixel the gradients of its three vertices go to the gradient arena with
// float atomics; material-factor gradients are summed over the wave when the
// wave sees one mesh info (the common case) and added once; the six light
// gradients stay in registers and every block writes one partial, which a
// second one-block kernel sums in a fixed order.
#include "kernels.h"
#include "shading.h"

using namespace lhd;
using namespace shading;

namespace {


LH_DEV f3 madd(f3 acc, f3 a, float s) { return add(acc, mul(a, s)); }
LH_DEV float pow4(float x) {
    const float x2 = x * x;
    return x2 * x2;
}

// u = a / |a|: the gradient of a from the gradient of u.
LH_DEV f3 normalize_bwd(f3 a, f3 gu) {
    const float l = length3(a);
    const f3 u = mul(a, 1.0f / l);
    return mul(sub(gu, mul(u, dot3(u, gu))), 1.0f / l);
}

struct BrdfGrad {
    f3 v, l, n, base;
    float pr, metallic;
};

// Reverse mode of shading::brdf (thin = false) for the upstream gradient g of
// its RGB output.
LH_DEV void brdf_bwd(f3 v, f3 l, f3 n, float pr, float m, f3 base, f3 g, BrdfGrad& o) {
    const f3 hv = add(v, l);
    const f3 h = normalize3(hv);
    const f3 diffuse = mul(base, 1.0f - m);
    const float refl = ((1.5f - 1.0f) / (1.5f + 1.0f)) * ((1.5f - 1.0f) / (1.5f + 1.0f));
    const float d0 = refl * (1.0f - m);
    const f3 f0 = add(f3{d0, d0, d0}, mul(base, m));
    const float nv = dot3(n, v), nl = dot3(n, l), nh = dot3(n, h), lh = dot3(l, h);
    const float NoV = clampf(nv, 1e-5f, 1.0f), NoL = clampf(nl, 0.0f, 1.0f);
    const float NoH = clampf(nh, 0.0f, 1.0f), LoH = clampf(lh, 0.0f, 1.0f);
    const float a = pr * pr, a2 = a * a;
    const float fD = (NoH * a2 - NoH) * NoH + 1.0f;
    const float D = a2 / ((kPi * fD) * fD);
    const float p = pow5(1.0f - LoH);
    const f3 F{f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
    const float sL = sqrtf((-NoL * a2 + NoL) * NoL + a2), sV = sqrtf((-NoV * a2 + NoV) * NoV + a2);
    const float den = NoL * sV + NoV * sL;
    const float Vis = 0.5f / den;
    const float f90 = 0.5f + ((2.0f * a) * LoH) * LoH;
    const float q5L = pow5(1.0f - NoL), q5V = pow5(1.0f - NoV);
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
    gNoL += gFSL * (f90 - 1.0f) * (-5.0f * pow4(1.0f - NoL));
    float gNoV = gFSV * (f90 - 1.0f) * (-5.0f * pow4(1.0f - NoV));
    float ga = gf90 * (2.0f * LoH * LoH);
    float gLoH = gf90 * (4.0f * a * LoH);
    const float gden = -gVis * Vis / den;
    gNoV += gden * sL;
    gNoL += gden * sV;
    const float gsL = gden * NoV, gsV = gden * NoL;
    gNoL += gsL * NoL * (1.0f - a2) / sL;
    gNoV += gsV * NoV * (1.0f - a2) / sV;
    float ga2 = gsL * (1.0f - NoL * NoL) / (2.0f * sL) + gsV * (1.0f - NoV * NoV) / (2.0f * sV);
    const f3 gf0{gF.x * (1.0f - p), gF.y * (1.0f - p), gF.z * (1.0f - p)};
    const float gp = (gF.x * (1.0f - f0.x) + gF.y * (1.0f - f0.y)) + gF.z * (1.0f - f0.z);
    gLoH += gp * (-5.0f * pow4(1.0f - LoH));
    const float E = D / fD;
    ga2 += gD * (1.0f / ((kPi * fD) * fD) - 2.0f * E * NoH * NoH);
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
    o.v = mul(n, gnv);
    o.l = add(mul(n, gnl), mul(h, glh));
    const f3 ghv = normalize_bwd(hv, add(mul(n, gnh), mul(l, glh)));
    o.v = add(o.v, ghv);
    o.l = add(o.l, ghv);
}

// d(bilinear)/d(uv) at one level, contracted with the upstream gradient g.
LH_DEV f2 bilinear_bwd(const LhTexture& t, uint32_t level, f2 uv, f4 g) {
    const float w = float(max(1u, t.width >> level)), h = float(max(1u, t.height >> level));
    const float u = uv.x * w - 0.5f, v = uv.y * h - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int32_t i = int32_t(fu), j = int32_t(fv);
    const f4 t00 = texel(t, level, i, j), t10 = texel(t, level, i + 1, j);
    const f4 t01 = texel(t, level, i, j + 1), t11 = texel(t, level, i + 1, j + 1);
    auto da = [&](float x00, float x10, float x01, float x11) { return (1.0f - b) * (x10 - x00) + b * (x11 - x01); };
    auto db = [&](float x00, float x10, float x01, float x11) {
        return ((1.0f - a) * x01 + a * x11) - ((1.0f - a) * x00 + a * x10);
    };
    const float ga = ((g.x * da(t00.x, t10.x, t01.x, t11.x) + g.y * da(t00.y, t10.y, t01.y, t11.y)) +
                      g.z * da(t00.z, t10.z, t01.z, t11.z)) + g.w * da(t00.w, t10.w, t01.w, t11.w);
    const float gb = ((g.x * db(t00.x, t10.x, t01.x, t11.x) + g.y * db(t00.y, t10.y, t01.y, t11.y)) +
                      g.z * db(t00.z, t10.z, t01.z, t11.z)) + g.w * db(t00.w, t10.w, t01.w, t11.w);
    return {ga * w, gb * h};
}

// Gradient wrt uv of the trilinear sample (LOD held fixed).
LH_DEV f2 sample_texture_bwd(const lh_scene& s, uint32_t index, const Interp2& uv, f4 g) {
    if (!has_texture(s, index)) return {0.0f, 0.0f};
    const LhTexture& t = s.textures[index];
    const Lod lod = texture_lod(t, uv);
    f2 r = bilinear_bwd(t, lod.l0, uv.v, g);
    if (lod.f == 0.0f) return r;
    const f2 r1 = bilinear_bwd(t, lod.l1, uv.v, g);
    return {r.x * (1.0f - lod.f) + r1.x * lod.f, r.y * (1.0f - lod.f) + r1.y * lod.f};
}

struct LightGrad {
    f3 sun_dir, sun_intensity;
};

LH_DEV void add_geometry(const lh_resolve_grads& gr, uint64_t addr, f3 g) {
    if (!gr.geometry || addr < gr.geometry_base) return;
    const uint64_t k = (addr - gr.geometry_base) / 4;
    if (k + 3 > gr.geometry_bytes / 4) return;
    atomicAdd(gr.geometry + k, g.x);
    atomicAdd(gr.geometry + k + 1, g.y);
    atomicAdd(gr.geometry + k + 2, g.z);
}

LH_DEV bool supported(const MeshInfo& mi) {
    return (mi.flags & MESH_INFO_FLAGS_FP32_VERTICES) && !(mi.flags & MESH_INFO_FLAGS_ALPHA_CLIP) &&
           mi.normal_texture_index == UNUSED_TEXTURE_INDEX;
}

// One pixel: scatters vertex gradients, returns its material-factor gradients
// (mesh info in *mesh_info, or ~0u when it contributes nothing) and adds its
// light gradients to lg.
LH_DEV void pixel_bwd(const lh_scene& s, const lh_view& v, const lh_resolve_params& prm, const float* grad_radiance,
                      const lh_resolve_grads& gr, uint32_t x, uint32_t y, uint32_t& mesh_info, f3& g_bcf, float& g_rf,
                      float& g_mf, LightGrad& lg) {
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
    const float4 gin = reinterpret_cast<const float4*>(grad_radiance)[o];
    const f3 g{gin.x, gin.y, gin.z};
    // ---- forward recomputation (as shading::resolve_pixel) ----
    const f2 win{float(u.window_size.x), float(u.window_size.y)};
    const f2 ndc{float(x) / win.x * 2.0f - 1.0f, float(y) / win.y * 2.0f - 1.0f};
    const BaryDeriv bary = calc_full_bary(sf.clip[0], sf.clip[1], sf.clip[2], ndc, win);
    const Interp3 world = interp3(bary, sf.pos[0], sf.pos[1], sf.pos[2]);
    const f2 ts{mi.texture_scale.x, mi.texture_scale.y};
    f2 uvs[3];
    f3 nraw[3], nn[3];
    for (int j = 0; j < 3; j++) {
        const f2 t = load_uv(mi, sf.vi[j]);
        uvs[j] = {t.x * ts.x, t.y * ts.y};
        nraw[j] = load_normal(mi, sf.vi[j]);
        nn[j] = normalize3(nraw[j]);
    }
    Interp2 uv = interp2(bary, uvs[0], uvs[1], uvs[2]);
    uv.v.x += mi.texture_offset.x;
    uv.v.y += mi.texture_offset.y;
    const f3 cam = mk3(u.camera_pos);
    const f3 to_cam = sub(cam, world.v);
    const f3 view_vector = normalize3(to_cam);
    const f3 nI = interp3(bary, nn[0], nn[1], nn[2]).v;
    const f3 m3 = mat3_mul_vec3(in.normal_transform, nI);
    f3 nrm = normalize3(m3);
    const f3 gn = cross3(sub(sf.pos[1], sf.pos[0]), sub(sf.pos[2], sf.pos[0]));
    const float flip = dot3(gn, view_vector) < 0.0f ? -1.0f : 1.0f;
    nrm = mul(nrm, flip);
    const float shadow_sum = shadow_factor(v, mi, world.v);
    const f4 bc = sample_texture(s, mi.base_color_texture_index, uv);
    const f3 bcf = param_base_color(prm, mi, sf.mesh_info);
    const f3 base = mul(f3{bc.x, bc.y, bc.z}, bcf);
    const f4 mr = sample_texture(s, mi.metallic_roughness_texture_index, uv);
    const float rf = param_roughness(prm, s, sf.mesh_info), mf = param_metallic(prm, s, sf.mesh_info);
    const float roughness = mr.y * rf, metallic = mr.z * mf;
    const float shv = ambient_sh(nrm);
    const f3 sun_dir = param_sun_dir(prm, u), sun_int = param_sun_intensity(prm, u);
    const f3 B = brdf(view_vector, sun_dir, nrm, roughness, metallic, base, false);
    const float K = kPi * shadow_sum;
    // ---- reverse ----
    // out = B * K * I + base * (1 - metallic) * shv
    lg.sun_intensity = add(lg.sun_intensity, mul(mul(g, B), K));
    const f3 gB = mul(mul(g, sun_int), K);
    f3 gbase = mul(g, (1.0f - metallic) * shv);
    float gmet = -dot3(g, base) * shv;
    const float gshv = dot3(g, base) * (1.0f - metallic);
    BrdfGrad bg;
    brdf_bwd(view_vector, sun_dir, nrm, roughness, metallic, base, gB, bg);
    lg.sun_dir = add(lg.sun_dir, bg.l);
    gbase = add(gbase, bg.base);
    gmet += bg.metallic;
    f3 gnrm = bg.n;
    gnrm.y += gshv * 0.025f;  // d(ambient_sh)/d(nrm.y)
    // factors and texture samples
    mesh_info = sf.mesh_info;
    g_bcf = mul(gbase, f3{bc.x, bc.y, bc.z});
    g_rf = bg.pr * mr.y;
    g_mf = gmet * mr.z;
    const f2 guv_b = sample_texture_bwd(s, mi.base_color_texture_index, uv, f4{gbase.x * bcf.x, gbase.y * bcf.y,
                                                                               gbase.z * bcf.z, 0.0f});
    const f2 guv_m = sample_texture_bwd(s, mi.metallic_roughness_texture_index, uv, f4{0.0f, bg.pr * rf, gmet * mf, 0.0f});
    const f2 guv{guv_b.x + guv_m.x, guv_b.y + guv_m.y};
    // normal chain: nrm = flip * normalize(N nI), nI = sum lambda_j normalize(nraw_j)
    const f3 gm3 = normalize_bwd(m3, mul(gnrm, flip));
    const LhMat3& N = in.normal_transform;
    const f3 gnI{dot3(mk3(N.c[0]), gm3), dot3(mk3(N.c[1]), gm3), dot3(mk3(N.c[2]), gm3)};
    // world position: view vector (and nothing else: shadow and LOD are fixed)
    const f3 gworld = neg(normalize_bwd(to_cam, bg.v));
    const float lam[3] = {bary.lambda.x, bary.lambda.y, bary.lambda.z};
    float glam[3];
    f3 gpos[3];
    for (int j = 0; j < 3; j++) {
        glam[j] = (dot3(gworld, sf.pos[j]) + dot3(gnI, nn[j])) + (guv.x * uvs[j].x + guv.y * uvs[j].y);
        gpos[j] = mul(gworld, lam[j]);
        add_geometry(gr, mi.normals + 12ull * sf.vi[j], normalize_bwd(nraw[j], mul(gnI, lam[j])));
    }
    // barycentrics: lambda_j = D_j / sum D, D_j = q . (c_{j+1} x c_{j+2})
    const f3 q{ndc.x, ndc.y, 1.0f};
    f3 c[3];
    for (int j = 0; j < 3; j++) c[j] = f3{sf.clip[j].x, sf.clip[j].y, sf.clip[j].w};
    float Dk[3];
    for (int k = 0; k < 3; k++) Dk[k] = dot3(q, cross3(c[(k + 1) % 3], c[(k + 2) % 3]));
    const float sumD = (Dk[0] + Dk[1]) + Dk[2];
    const float lg_dot = (glam[0] * Dk[0] + glam[1] * Dk[1] + glam[2] * Dk[2]) / sumD;
    f3 gc[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 3; k++) {
        const float gD = (glam[k] - lg_dot) / sumD;
        const int a = (k + 1) % 3, b = (k + 2) % 3;
        gc[a] = madd(gc[a], cross3(c[b], q), gD);
        gc[b] = madd(gc[b], cross3(q, c[a]), gD);
    }
    const LhMat4& PV = u.combined_perspective_view;
    const LhMat4& T = in.transform;
    for (int j = 0; j < 3; j++) {
        const f4 gclip{gc[j].x, gc[j].y, 0.0f, gc[j].z};
        for (int k = 0; k < 3; k++) {
            const LhVec4 col = PV.c[k];
            const float d = ((col.x * gclip.x + col.y * gclip.y) + col.z * gclip.z) + col.w * gclip.w;
            if (k == 0) gpos[j].x += d;
            if (k == 1) gpos[j].y += d;
            if (k == 2) gpos[j].z += d;
        }
        const f3 gl{dot3(xyz(T.c[0]), gpos[j]), dot3(xyz(T.c[1]), gpos[j]), dot3(xyz(T.c[2]), gpos[j])};
        add_geometry(gr, mi.positions + 12ull * sf.vi[j], gl);
    }
}

__global__ void __launch_bounds__(64) k_resolve_backward(lh_scene s, OneView vs, lh_resolve_params prm,
                                                         const float* grad_radiance, lh_resolve_grads gr,
                                                         float* partials) {
    const lh_view& v = vs.v;
    const uint32_t tiles_x = (v.width + 7) / 8, tiles = tiles_x * ((v.height + 7) / 8);
    LightGrad lg{{0, 0, 0}, {0, 0, 0}};
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint32_t x = (t % tiles_x) * 8 + (threadIdx.x & 7), y = (t / tiles_x) * 8 + (threadIdx.x >> 3);
        uint32_t mesh_info = ~0u;
        f3 g_bcf{0, 0, 0};
        float g_rf = 0.0f, g_mf = 0.0f;
        if (x < v.width && y < v.height) pixel_bwd(s, v, prm, grad_radiance, gr, x, y, mesh_info, g_bcf, g_rf, g_mf, lg);
        // material factors: one add per wave when every contributing lane shares the mesh info
        const uint64_t contrib = ballot(mesh_info != ~0u);
        if (contrib == 0) continue;
        const uint32_t lead = __builtin_amdgcn_readlane(mesh_info, int(first_lane(contrib)));
        const bool uniform = ballot(mesh_info != ~0u && mesh_info != lead) == 0;
        if (uniform) {
            float r[5] = {g_bcf.x, g_bcf.y, g_bcf.z, g_rf, g_mf};
            for (int i = 0; i < 5; i++) {
                float sum = r[i];
                for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
                r[i] = sum;
            }
            if (threadIdx.x == 0) {
                if (gr.base_color_factor) {
                    atomicAdd(gr.base_color_factor + 3 * lead, r[0]);
                    atomicAdd(gr.base_color_factor + 3 * lead + 1, r[1]);
                    atomicAdd(gr.base_color_factor + 3 * lead + 2, r[2]);
                }
                if (gr.roughness_factor) atomicAdd(gr.roughness_factor + lead, r[3]);
                if (gr.metallic_factor) atomicAdd(gr.metallic_factor + lead, r[4]);
            }
        } else if (mesh_info != ~0u) {
            if (gr.base_color_factor) {
                atomicAdd(gr.base_color_factor + 3 * mesh_info, g_bcf.x);
                atomicAdd(gr.base_color_factor + 3 * mesh_info + 1, g_bcf.y);
                atomicAdd(gr.base_color_factor + 3 * mesh_info + 2, g_bcf.z);
            }
            if (gr.roughness_factor) atomicAdd(gr.roughness_factor + mesh_info, g_rf);
            if (gr.metallic_factor) atomicAdd(gr.metallic_factor + mesh_info, g_mf);
        }
    }
    float r[6] = {lg.sun_dir.x, lg.sun_dir.y, lg.sun_dir.z, lg.sun_intensity.x, lg.sun_intensity.y,
                  lg.sun_intensity.z};
    for (int i = 0; i < 6; i++) {
        float sum = r[i];
        for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
        r[i] = sum;
    }
    if (threadIdx.x < 8) {
        float val = 0.0f;
        for (int i = 0; i < 6; i++)
            if (threadIdx.x == uint32_t(i)) val = r[i];
        partials[size_t(blockIdx.x) * 8 + threadIdx.x] = val;
    }
}

// Sums the per-block light partials in block order (deterministic) and adds
// them to the outputs.
__global__ void __launch_bounds__(64) k_reduce_light(const float* partials, uint32_t blocks, lh_resolve_grads gr) {
    const uint32_t i = threadIdx.x;
    if (i >= 6) return;
    float sum = 0.0f;
    for (uint32_t b = 0; b < blocks; b++) sum += partials[size_t(b) * 8 + i];
    float* dst = i < 3 ? gr.sun_dir : gr.sun_intensity;
    if (dst) dst[i % 3] += sum;
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
    float* partials = static_cast<float*>(v.scratch);
    mark("resolve_backward", st);
    hipLaunchKernelGGL(k_resolve_backward, dim3(blocks), dim3(64), 0, st, s, vs, params, grad_radiance, grads,
                       partials);
    mark("reduce_light", st);
    hipLaunchKernelGGL(k_reduce_light, dim3(1), dim3(64), 0, st, partials, blocks, grads);
    return hipGetLastError();
}

}  // namespace lhk
