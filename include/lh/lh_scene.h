/*
 * Host-side scene producer: synthetic meshes, meshlet builder, instances,
 * materials, value-noise textures and camera/uniform setup.
 *
 * Replaces the reference's asset ingestion for this path
 * (src/resources/mesh_loading.cpp:241-738 glTF load,
 *  src/resources/meshlets.cpp:3-140 meshopt meshlet build,
 *  src/resources/bounding_sphere.h:3-74 bounding sphere,
 *  src/main.cpp:433-440 Instance construction,
 *  src/main.cpp:646-670,734-760 uniform setup,
 *  src/projection.h:3-20 infinite reverse-Z perspective).
 *
 * Everything here runs on the CPU.  The arrays it produces hold offsets in
 * place of device addresses; lh_scene_relocate() turns them into absolute
 * addresses for a given set of base pointers (device memory for the HIP
 * path, host memory for CPU consumers).
 */
#ifndef LH_SCENE_H
#define LH_SCENE_H

#include "shared_cpu_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lh_scene_host lh_scene_host; /* opaque */

/* Vertex formats the producer writes.
 * FP32: float3 positions and normals, float2 uvs, u32 remap indices
 *       (MESH_INFO_FLAGS_FP32_VERTICES | MESH_INFO_FLAGS_32_BIT_INDICES).
 * QUANTIZED: the reference's formats (src/shaders/common/buffer_references.glsl:24-50):
 *       u16x3 positions on a per-mesh grid (its origin and step are folded
 *       into every instance transform), i8x3 normals (n * 127), u16x2 uvs (the
 *       grid folded into the mesh info's texture_scale / texture_offset), u16
 *       remap indices when the mesh has at most 65,536 vertices, else u32. */
#define LH_VERTEX_FORMAT_FP32 0
#define LH_VERTEX_FORMAT_QUANTIZED 1

/* Synthetic workloads. */
#define LH_PRESET_SMALL 1  /* 64 instances of one 8,192-triangle sphere, 4 constant materials */
#define LH_PRESET_LARGE 2  /* 2,000 instances of 50 meshes (5k-500k tris), 64 textured materials */
#define LH_PRESET_STRESS 3 /* 20,000 instances of the same 50 meshes */

typedef struct lh_preset_config {
    uint32_t preset;
    uint32_t num_instances;   /* 0 = preset default */
    uint64_t seed;            /* 0 = preset default (preset number) */
    uint32_t texture_size;    /* 0 = preset default (1024 for LARGE/STRESS) */
    uint32_t mesh_tri_scale_pct; /* 0 = 100; scales per-mesh triangle targets (CPU-sized tests) */
    uint32_t vertex_format;   /* LH_VERTEX_FORMAT_* (0 = FP32) */
    uint32_t texture_format;  /* LH_TEXTURE_RGBA8 / BC1 / BC7, optionally | LH_TEXTURE_SRGB (shared_cpu_gpu.h) */
    const char* cache_dir;    /* NULL, or an existing directory for the built-mesh cache */
} lh_preset_config;

/* Camera of one view.  `target` is a point on the view axis; up is +Y. */
typedef struct lh_camera {
    LhVec3 position;
    LhVec3 target;
    float fov_y_rad;
    float near_plane;
    uint32_t width;
    uint32_t height;
    LhVec3 sun_dir;           /* unit vector towards the sun */
    LhVec3 sun_intensity;
    float shadow_cam_distance;
    float cascade_split_pow;
} lh_camera;

typedef struct lh_scene_arrays {
    const uint8_t* geometry;          /* positions, normals, uvs, remap, micro indices, meshlets */
    uint64_t geometry_bytes;
    const MeshInfo* mesh_infos;       /* address fields are byte offsets into `geometry` */
    uint32_t num_mesh_infos;
    const LhMaterialFactors* factors; /* [num_mesh_infos] */
    const Instance* instances;        /* mesh_info_address holds the MeshInfo index */
    uint32_t num_instances;
    const uint32_t* instance_meshlet_base; /* [num_instances] exclusive scan of num_meshlets */
    const MeshletReference* meshlet_table; /* [total_meshlets] stable id -> (instance, meshlet) */
    uint32_t total_meshlets;
    const uint8_t* texels;            /* RGBA8 texel arena */
    uint64_t texel_bytes;
    const LhTexture* textures;        /* texels field is a byte offset into `texels` */
    uint32_t num_textures;
    uint32_t num_meshes;
    uint64_t total_triangles;         /* sum over instances */
    uint64_t unique_triangles;        /* sum over meshes */
} lh_scene_arrays;

/* ---- incremental builder (hand-built fixtures and presets) ---- */
lh_scene_host* lh_scene_create(void);
/* On-disk cache of built meshes (the reference's FsCache): meshes added
 * afterwards are looked up in `dir` by a hash of their source arrays and the
 * vertex format, and stored there when built.  `dir` must exist; NULL or ""
 * switches the cache off.  Only before the first mesh.  0 on success. */
int32_t lh_scene_set_cache_dir(lh_scene_host* s, const char* dir);
/* Texture storage of the finished scene for textures added as RGBA8:
 * LH_TEXTURE_RGBA8 (default), LH_TEXTURE_BC1 (each mip level encoded to BC1
 * blocks, alpha dropped) or LH_TEXTURE_BC7 (mode-6 blocks, alpha kept),
 * decoded on fetch by the kernels; | LH_TEXTURE_SRGB marks the data
 * sRGB-encoded (decoded to linear per texel before filtering).  Encoded
 * textures keep their own format.  Before lh_scene_finish.  0 on success. */
int32_t lh_scene_set_texture_format(lh_scene_host* s, uint32_t format);
/* Selects the vertex format (LH_VERTEX_FORMAT_*) of the meshes added after
 * it; only before the first mesh.  0 on success. */
int32_t lh_scene_set_vertex_format(lh_scene_host* s, uint32_t format);
/* Adds a mesh with FP32 positions/normals (3 floats each), uvs (2 floats) and
 * uint32 triangle indices; builds its meshlets.  Returns the mesh id or -1. */
int32_t lh_scene_add_mesh(lh_scene_host* s, const float* positions, const float* normals,
                          const float* uvs, uint32_t num_vertices, const uint32_t* indices,
                          uint32_t num_triangles);
/* Adds an RGBA8 texture (level 0 only; the mip chain is built by 2x2 box
 * filtering).  Returns the texture index or -1. */
int32_t lh_scene_add_texture(lh_scene_host* s, const uint8_t* rgba, uint32_t width, uint32_t height);
/* Adds a texture as stored in a file (DDS, KTX2): `levels` mip levels of
 * `format` (LH_TEXTURE_RGBA8 / BC1 / BC7, optionally | LH_TEXTURE_SRGB) packed
 * level after level, level l of size max(1, width >> l) x max(1, height >> l)
 * (block formats: ceil(w/4) x ceil(h/4) blocks of 8 / 16 bytes).  `bytes` must
 * equal the sum of the level sizes.  Returns the texture index or -1. */
int32_t lh_scene_add_texture_encoded(lh_scene_host* s, uint32_t format, uint32_t width, uint32_t height,
                                     uint32_t levels, const uint8_t* data, uint64_t bytes);
/* Material = (base colour factor, roughness/metallic factors, texture indices
 * or UNUSED_TEXTURE_INDEX).  Returns the material id or -1. */
int32_t lh_scene_add_material(lh_scene_host* s, LhVec3 base_color_factor, float roughness_factor,
                              float metallic_factor, uint32_t base_color_texture,
                              uint32_t metallic_roughness_texture, uint32_t normal_texture,
                              LhVec2 texture_scale, LhVec2 texture_offset);
/* Marks a material alpha-clipped (MESH_INFO_FLAGS_ALPHA_CLIP on its mesh
 * infos: double-sided raster, samples whose base-colour alpha is below 0.5
 * discarded, thin-walled BRDF, shadow bias 2e-6).  0 on success. */
int32_t lh_scene_set_material_alpha_clip(lh_scene_host* s, uint32_t material, uint32_t enabled);
/* Adds an instance of (mesh, material) with a column-major 4x4 transform. */
int32_t lh_scene_add_instance(lh_scene_host* s, uint32_t mesh, uint32_t material,
                              const LhMat4* transform);
/* Freezes the scene; builds MeshInfos, meshlet tables and scans. 0 on success. */
int32_t lh_scene_finish(lh_scene_host* s);

/* ---- presets ---- */
lh_scene_host* lh_scene_build_preset(const lh_preset_config* cfg);
/* Default camera of a preset (view 0). */
int32_t lh_preset_camera(uint32_t preset, lh_camera* out);
/* Seeded cameras for the batched-views workload: square views with side
 * log-uniform in [min_side, max_side] (multiples of 8), placed in the scene. */
int32_t lh_batch_cameras(uint64_t seed, uint32_t count, uint32_t min_side, uint32_t max_side,
                         lh_camera* out);

void lh_scene_destroy(lh_scene_host* s);
int32_t lh_scene_get(const lh_scene_host* s, lh_scene_arrays* out);

/* Writes copies of the MeshInfo / Instance / LhTexture arrays with absolute
 * addresses: geometry offsets + geometry_base, MeshInfo index * 104 +
 * mesh_info_base, texel offsets + texel_base. */
int32_t lh_scene_relocate(const lh_scene_host* s, uint64_t geometry_base, uint64_t mesh_info_base,
                          uint64_t texel_base, MeshInfo* mesh_infos_out, Instance* instances_out,
                          LhTexture* textures_out);

/* Camera -> Uniforms (matrices, sun, window size, cascade parameters); the
 * six address fields are set to zero.  Follows src/main.cpp:734-760 and
 * src/projection.h:3-20. */
int32_t lh_make_uniforms(const lh_camera* cam, uint32_t num_instances, Uniforms* out);

/* Column-major transform = translate(t) * rotate_y(yaw) * scale(s). */
void lh_make_transform(LhVec3 translation, float yaw, float scale, LhMat4* out);

#ifdef __cplusplus
}
#endif
#endif /* LH_SCENE_H */
