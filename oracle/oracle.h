/*
 * Scalar C++ oracle of the GPU-driven frame.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; never part of the product path.
 *
 * Every function restates the reference's algorithm for one pass, citing the
 * file:line it follows, in plain scalar C++ compiled without FMA contraction
 * and with IEEE division / square root, so that culling decisions, prefix
 * sums and visibility-buffer keys are reproducible bit for bit by the HIP
 * kernels.  The entry points take the same lh_scene / lh_view structs as
 * include/lh/lh.h, but every address in them (and inside Instance / MeshInfo
 * / Uniforms) is a host address.
 *
 * The software raster rules have no reference code (the reference uses the
 * Vulkan fixed-function rasterizer); they are this project's restatement of
 * the Vulkan rasterization rules: 8-bit subpixel snapping, pixel-centre
 * sampling, top-left fill rule, CCW front faces in framebuffer coordinates,
 * clipping to 0 <= z <= w and to a guard band |x|,|y| <= 4w, screen-linear
 * depth clamped to [0, 1].
 */
#ifndef LH_ORACLE_H
#define LH_ORACLE_H

#include "lh/lh.h"

#ifdef __cplusplus
extern "C" {
#endif

lh_status orc_view_reset(const lh_scene* scene, const lh_view* view);
lh_status orc_cull_instances(const lh_scene* scene, const lh_view* view);
lh_status orc_cull_meshlets(const lh_scene* scene, const lh_view* view, uint32_t pass);
lh_status orc_raster_view(const lh_scene* scene, const lh_view* view);
lh_status orc_depth_bounds(const lh_scene* scene, const lh_view* view);
lh_status orc_fit_cascades(const lh_scene* scene, const lh_view* view);
lh_status orc_cull_instances_shadow(const lh_scene* scene, const lh_view* view);
lh_status orc_raster_shadow(const lh_scene* scene, const lh_view* view, uint32_t cascade);
lh_status orc_resolve(const lh_scene* scene, const lh_view* view);
lh_status orc_tonemap(const lh_view* view);
lh_status orc_frame(const lh_scene* scene, const lh_view* view);
/* orc_frame with the rasters and the resolve on `threads` threads (same outputs). */
lh_status orc_frame_threads(const lh_scene* scene, const lh_view* view, uint32_t threads);

/* ---- building blocks exposed for known-answer tests ---- */

/* Rasterizes one clip-space triangle (clip[v][x,y,z,w]).  kind 0: main view
 * (back faces culled, 64-bit keys into vis64 with id = meshlet << 7 | tri);
 * kind 1: shadow (front faces culled, float bits into depth32).
 * Returns the number of fragments that passed coverage. */
uint64_t orc_raster_clip_triangle(const float* clip, uint32_t width, uint32_t height, uint32_t kind,
                                  uint32_t id, uint64_t* vis64, uint32_t* depth32);

/* Sphere tests of src/shaders/common/culling.glsl:1-46 and :82-135 (1 = culled). */
int orc_cull_sphere(const Instance* instance, const float* sphere, const Uniforms* u);
int orc_cull_sphere_shadow(const Instance* instance, const float* sphere, const MiscStorage* misc, uint32_t cascade);
/* Cone tests of src/shaders/common/culling.glsl:137-151 (1 = culled). */
int orc_cull_cone(const Instance* instance, const Meshlet* meshlet, const Uniforms* u, int orthographic);

/* CalcFullBary (src/shaders/common/vbuffer.glsl:11-63): out = lambda[3], ddx[3], ddy[3]. */
void orc_calc_full_bary(const float* pt0, const float* pt1, const float* pt2, const float* pixel_ndc,
                        const float* win_size, float* out9);
/* BRDF (src/shaders/common/pbr.glsl:44-96): v, l, n, roughness, metallic, base colour -> rgb. */
void orc_brdf(const float* v, const float* l, const float* n, float roughness, float metallic,
              const float* base_color, int thin_walled, float* out3);
/* sRGB OETF (src/shaders/display_transform.comp:19-25). */
float orc_srgb_oetf(float x);

/* Two-phase occlusion bound check: for every meshlet of draws[0], the pixel
 * bounding box of each of its triangles (within the image) must lie in the
 * meshlet's conservative screen rectangle and its vertex depths must not
 * exceed the bound's depth (1e-4 relative slack).  Returns the number of
 * violations; *checked counts the triangles examined. */
int64_t orc_hiz_bound_violations(const lh_scene* scene, const lh_view* view, uint64_t* checked);
/* One texel (x, y) of a BC1 block (8 bytes), packed RGBA8 with red in the low byte. */
uint32_t orc_bc1_texel(const uint8_t* block, uint32_t x, uint32_t y);
/* The 16 texels of a BC7 block (16 bytes) as RGBA8, row order (64 bytes). */
void orc_bc7_block(const uint8_t* block, uint8_t* rgba64);
/* sRGB EOTF of an 8-bit channel value (the *_SRGB sampler decode). */
float orc_srgb_to_linear(uint32_t v);

/* Shading building blocks of the resolve (render_geometry.comp), exposed so
 * that closed forms can pin them:
 *   orc_shadow_tap     one sampler2DArrayShadow tap (linear filter, clamp to
 *                      edge, compare LESS) of layer `layer` at uv[2] against ref;
 *   orc_shadow_factor  cascade selection + 3x3 PCF (:183-226) of a world
 *                      point; *cascade gets the loop's end index (0..4);
 *   orc_texture_lod    mip levels and blend of the trilinear filter from
 *                      uv6 = {u, v, du/dx, dv/dx, du/dy, dv/dy};
 *   orc_bilinear       bilinear fetch of one level with repeat addressing;
 *   orc_sample_texture trilinear sample_texture (:59-66) of scene texture `index`;
 *   orc_normal_map     the normal-map step (:238-246) with the cotangent frame
 *                      (:25-38): normal[3], world9 = {p, dp/dx, dp/dy},
 *                      uv6 as above, texel[4] the sampled RGBA in [0, 1]. */
float orc_shadow_tap(const lh_view* view, uint32_t layer, const float* uv, float ref);
/* Sky radiance of the pixel corner ndc[2] (render_geometry.comp:86-101):
 * lh_view.skybox looked up along the unprojected view ray, times 2^3; the
 * constant sky without a skybox. */
void orc_sky_radiance(const lh_view* view, const float* ndc, float* rgb3);
/* Heat-map colour temperature(t) of the shader-clock debug view
 * (src/shaders/common/debug.glsl:15-43). */
void orc_temperature(float t, float* rgb3);
float orc_shadow_factor(const lh_view* view, const MeshInfo* mesh_info, const float* world, uint32_t* cascade);
void orc_texture_lod(const LhTexture* texture, const float* uv6, uint32_t* l0, uint32_t* l1, float* f);
void orc_bilinear(const LhTexture* texture, uint32_t level, const float* uv, float* rgba);
void orc_sample_texture(const lh_scene* scene, uint32_t index, const float* uv6, float* rgba);
void orc_normal_map(const float* normal, const float* world9, const float* uv6, const float* texel, float* out3);

/* ---- differentiable resolve ----
 * Per-pixel quantities the adjoint holds fixed (FP32, as the forward resolve
 * computes them): kind 0 sky, 1 shaded, 2 unsupported mesh info (16-bit
 * vertices, alpha clip or a normal texture); back-face flip (+1/-1); PCF
 * shadow factor; texture LOD of [0] base colour, [1] metallic-roughness. */
typedef struct orc_frozen {
    uint32_t kind;
    float flip;
    float shadow;
    uint32_t lod_l0[2];
    uint32_t lod_l1[2];
    float lod_f[2];
} orc_frozen;
lh_status orc_resolve_frozen(const lh_scene* scene, const lh_view* view, orc_frozen* out /* width*height */);

/* FP64 restatement of the resolve as a function of its differentiable
 * parameters (each NULL pointer = the scene / uniform value).  geometry, when
 * set, holds the geometry arena as doubles (one per 4-byte word from
 * geometry_base) and replaces the FP32 positions and normals. */
typedef struct orc_params64 {
    const double* base_color_factor; /* [num_mesh_infos * 3] */
    const double* roughness_factor;  /* [num_mesh_infos] */
    const double* metallic_factor;   /* [num_mesh_infos] */
    const double* sun_dir;           /* [3] */
    const double* sun_intensity;     /* [3] */
    const double* geometry;          /* [geometry_bytes / 4] */
    uint64_t geometry_base;
    uint64_t geometry_bytes;
} orc_params64;
/* FP64 radiance RGB (width*height*3) with the frozen quantities of `frozen`. */
lh_status orc64_resolve(const lh_scene* scene, const lh_view* view, const orc_params64* params,
                        const orc_frozen* frozen, double* rgb);
/* Gradients of sum(grad_radiance . radiance) (grad_radiance: width*height*4
 * floats, alpha ignored) by forward-mode dual numbers over orc64_resolve,
 * ADDED into the non-NULL outputs (g_geometry mirrors the geometry arena). */
lh_status orc64_resolve_grad(const lh_scene* scene, const lh_view* view, const orc_params64* params,
                             const orc_frozen* frozen, const float* grad_radiance, double* g_base_color,
                             double* g_roughness, double* g_metallic, double* g_sun_dir, double* g_sun_intensity,
                             double* g_geometry);

#ifdef __cplusplus
}
#endif
#endif /* LH_ORACLE_H */
