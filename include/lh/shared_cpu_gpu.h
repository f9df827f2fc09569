/*
 * Host/device data contract of the GPU-driven frame.
 *
 * Byte layout follows the reference's shared header
 * (src/shared_cpu_gpu.h:8-166), which is compiled there both as C++ (glm
 * types, 4-byte aligned) and as GLSL with the `scalar` block layout.  The
 * structs below are a plain-C re-expression of that layout with every offset
 * pinned by a static assertion, so the same bytes can be produced on the host,
 * read by the HIP kernels and read by the scalar oracle.
 *
 * Changes from the reference, all additive (no existing offset moves):
 *   - MESH_INFO_FLAGS_FP32_VERTICES: positions float3, normals float3,
 *     uvs float2 (the reference only has quantized u16x3 / i8x3 / u16x2).
 *   - Capacities (instances per prefix-sum sector, meshlets per draw list)
 *     are runtime parameters instead of the compile-time MAX_INSTANCES = 1024
 *     and MAX_OPAQUE_DRAWS = 200000 (src/shared_cpu_gpu.h:119-125,159-162).
 *   - The near plane is read from the perspective matrix (column 3, row 2,
 *     src/projection.h:18) instead of the compile-time NEAR_PLANE
 *     (src/shared_cpu_gpu.h:127), so it can differ per view.
 *   - Roughness / metallic factors live in a side table (LhMaterialFactors)
 *     indexed like the MeshInfo array, so MeshInfo keeps its 104 bytes.
 *   - Instance.normal_transform is written by the host as
 *     mat3(transpose(inverse(transform))); the reference writes
 *     mat3(inverse(transform)) (src/shared_cpu_gpu.h:92-97), which only equals
 *     the correct normal matrix for scale+translate transforms.
 *
 * This header is valid C11 and C++20 and is included by HIP device code.
 */
#ifndef LH_SHARED_CPU_GPU_H
#define LH_SHARED_CPU_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define LH_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define LH_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

/* ---- 4-byte aligned vector / matrix types (glm layout, column major) ---- */
typedef struct LhVec2 { float x, y; } LhVec2;
typedef struct LhVec3 { float x, y, z; } LhVec3;
typedef struct LhVec4 { float x, y, z, w; } LhVec4;
typedef struct LhUVec2 { uint32_t x, y; } LhUVec2;
typedef struct LhMat3 { LhVec3 c[3]; } LhMat3; /* c[column] */
typedef struct LhMat4 { LhVec4 c[4]; } LhMat4; /* c[column] */

/* src/shared_cpu_gpu.h:8-9 plus one new bit. */
#define MESH_INFO_FLAGS_32_BIT_INDICES 0x01u
#define MESH_INFO_FLAGS_ALPHA_CLIP 0x02u
#define MESH_INFO_FLAGS_FP32_VERTICES 0x04u

/* src/shared_cpu_gpu.h:12-16 */
typedef struct DispatchIndirectCommand {
    uint32_t x, y, z;
} DispatchIndirectCommand;

/* src/shared_cpu_gpu.h:18-34 */
typedef struct MeshInfo {
    uint64_t positions;     /* device address of vertex positions */
    uint64_t indices;       /* meshlet vertex remap table (u16 or u32) */
    uint64_t normals;
    uint64_t uvs;
    uint64_t micro_indices; /* u8 triangle corners, 3 per triangle */
    uint64_t meshlets;      /* Meshlet[num_meshlets] */
    uint16_t num_meshlets;
    uint8_t flags;
    LhVec4 bounding_sphere; /* xyz centre, w radius, mesh space */
    LhVec2 texture_scale;
    LhVec2 texture_offset;
    uint16_t base_color_texture_index;
    uint16_t metallic_roughness_texture_index;
    uint16_t normal_texture_index;
    LhVec3 base_color_factor;
} MeshInfo;

/* src/shared_cpu_gpu.h:36-43 */
typedef struct MiscStorage {
    LhMat4 shadow_matrices[4];
    LhMat4 uv_space_shadow_matrices[4];
    LhMat4 shadow_view_matrices[4];
    float shadow_sphere_radii[4];
    uint32_t min_depth; /* float bits, min over non-zero depths */
    uint32_t max_depth; /* float bits */
} MiscStorage;

/* src/shared_cpu_gpu.h:46-51 */
#define UNIFORMS_DEBUG_OFF 0
#define UNIFORMS_DEBUG_CASCADES 1
#define UNIFORMS_DEBUG_TRIANGLE_INDEX 2
#define UNIFORMS_DEBUG_INSTANCE_INDEX 3
#define UNIFORMS_DEBUG_SHADER_CLOCK 4
#define UNIFORMS_DEBUG_NORMALS 5

/* src/shared_cpu_gpu.h:53-76 */
typedef struct Uniforms {
    LhMat4 combined_perspective_view;
    LhMat4 inv_perspective_view;
    LhMat4 view;
    LhMat4 initial_view;
    LhMat4 perspective;
    LhMat4 perspective_inverse;
    LhMat4 view_inverse;
    uint64_t meshlet_references;
    uint64_t instances;
    uint64_t draw_calls;
    uint64_t misc_storage;
    uint64_t num_meshlets_prefix_sum;
    uint64_t dispatches;
    LhVec3 camera_pos;
    LhVec3 sun_dir;
    LhVec3 sun_intensity;
    uint32_t num_instances;
    LhUVec2 window_size;
    float shadow_cam_distance;
    float cascade_split_pow;
    int32_t debug;
    uint8_t debug_shadowmaps; /* C++ bool in the reference: 1 byte + 3 pad */
} Uniforms;

/* src/shared_cpu_gpu.h:79-84 */
typedef struct DrawIndirectCommand {
    uint32_t vertexCount, instanceCount, firstVertex, firstInstance;
} DrawIndirectCommand;

/* src/shared_cpu_gpu.h:86-99 */
typedef struct Instance {
    LhMat4 transform;
    uint64_t mesh_info_address;
    LhMat3 normal_transform;
} Instance;

/* src/shared_cpu_gpu.h:129-139 */
typedef struct Meshlet {
    LhVec3 cone_apex;
    LhVec3 cone_axis;
    float cone_cutoff;
    LhVec4 bounding_sphere;
    uint32_t triangle_offset; /* byte offset into micro_indices */
    uint32_t index_offset;    /* element offset into the remap table */
    uint8_t triangle_count;
    uint8_t index_count;
} Meshlet;

/* src/shared_cpu_gpu.h:142-145 */
typedef struct MeshletReference {
    uint32_t instance_index;
    uint16_t meshlet_index;
} MeshletReference;

/* src/shared_cpu_gpu.h:148-152 */
#define MAX_MESHLET_UNIQUE_VERTICES 64u
#define MAX_MESHLET_TRIANGLES 124u
#define MAX_MESHLET_VERTICES (MAX_MESHLET_TRIANGLES * 3u)
#define UNUSED_TEXTURE_INDEX 0xFFFFu

/* src/shared_cpu_gpu.h:154-157; a sector is { u64 counter; PrefixSumValue values[capacity]; } */
typedef struct PrefixSumValue {
    uint32_t index;
    uint32_t sum;
} PrefixSumValue;

/* src/shared_cpu_gpu.h:164-166 */
#define PER_INSTANCE_DISPATCH 0u
#define PER_SHADOW_INSTANCE_DISPATCH 1u
#define PER_MESHLET_DISPATCH 2u

/* ---- additions (no counterpart in the reference) ---- */

/* Per-MeshInfo material factors; identity (1, 1) reproduces the reference,
 * which takes roughness = G and metallic = B of the metallic-roughness texture
 * (src/shaders/render_geometry.comp:233-236). */
typedef struct LhMaterialFactors {
    float roughness_factor;
    float metallic_factor;
} LhMaterialFactors;

/* One texture with a full mip chain packed level after level, RGBA8 texels
 * or BC1 / BC7 blocks (decoded on fetch).  format = storage | flags:
 *   LH_TEXTURE_RGBA8: 4 bytes per texel, row-major; level_offset in texels.
 *   LH_TEXTURE_BC1:   8-byte blocks of 4x4 texels, ceil(w/4) blocks per row;
 *                     level_offset in blocks.
 *   LH_TEXTURE_BC7:   16-byte blocks of 4x4 texels, laid out as BC1.
 *   LH_TEXTURE_SRGB:  the RGB channels are sRGB-encoded and decoded to linear
 *                     per texel before filtering, as Vulkan samples the
 *                     *_SRGB formats the reference uploads
 *                     (src/resources/image_loading.cpp:14-39); alpha is linear.
 *   LH_TEXTURE_RGB:   alpha reads as 1 (Vulkan's BC1_RGB formats, where a
 *                     three-colour block's fourth entry is opaque black). */
#define LH_MAX_TEXTURE_LEVELS 16u
#define LH_TEXTURE_RGBA8 0u
#define LH_TEXTURE_BC1 1u
#define LH_TEXTURE_BC7 2u
#define LH_TEXTURE_FORMAT_MASK 0xFFu
#define LH_TEXTURE_SRGB 0x100u
#define LH_TEXTURE_RGB 0x200u
typedef struct LhTexture {
    uint64_t texels; /* device address of level 0 */
    uint32_t width;
    uint32_t height;
    uint32_t levels;
    uint32_t format; /* LH_TEXTURE_* */
    uint32_t level_offset[LH_MAX_TEXTURE_LEVELS];
} LhTexture;

/* Visibility-buffer key: depth_key << 32 | meshlet_id << 7 | triangle.
 * depth_key = ~float_bits(z_ndc) so that atomicMin keeps the nearest sample
 * under reverse-Z (depth compare GREATER, clear 0: src/pipelines.cpp:67-71,
 * src/rendering.cpp:206-211).  An empty pixel holds LH_VISBUFFER_CLEAR
 * (depth 0.0).  meshlet_id is the stable per-scene id
 * instance_meshlet_base[instance] + meshlet_index. */
#define LH_VISBUFFER_CLEAR 0xFFFFFFFFFFFFFFFFull
#define LH_KEY_TRI_BITS 7u
#define LH_KEY_MESHLET_BITS 25u
#define LH_MAX_SCENE_MESHLETS (1u << LH_KEY_MESHLET_BITS)
/* Shadow maps hold float bits of z in [0, 1]; cleared to 1.0 (compare LESS,
 * src/pipelines.cpp:73-77, src/rendering.cpp:384). */
#define LH_SHADOW_CLEAR 0x3F800000u

/* ---- layout pins (offsets measured from the reference's C++ structs) ---- */
LH_STATIC_ASSERT(sizeof(LhVec3) == 12 && sizeof(LhVec4) == 16, "vec");
LH_STATIC_ASSERT(sizeof(LhMat3) == 36 && sizeof(LhMat4) == 64, "mat");
LH_STATIC_ASSERT(sizeof(MeshInfo) == 104, "MeshInfo size");
LH_STATIC_ASSERT(offsetof(MeshInfo, num_meshlets) == 48, "MeshInfo.num_meshlets");
LH_STATIC_ASSERT(offsetof(MeshInfo, flags) == 50, "MeshInfo.flags");
LH_STATIC_ASSERT(offsetof(MeshInfo, bounding_sphere) == 52, "MeshInfo.bounding_sphere");
LH_STATIC_ASSERT(offsetof(MeshInfo, texture_scale) == 68, "MeshInfo.texture_scale");
LH_STATIC_ASSERT(offsetof(MeshInfo, texture_offset) == 76, "MeshInfo.texture_offset");
LH_STATIC_ASSERT(offsetof(MeshInfo, base_color_texture_index) == 84, "MeshInfo.tex0");
LH_STATIC_ASSERT(offsetof(MeshInfo, metallic_roughness_texture_index) == 86, "MeshInfo.tex1");
LH_STATIC_ASSERT(offsetof(MeshInfo, normal_texture_index) == 88, "MeshInfo.tex2");
LH_STATIC_ASSERT(offsetof(MeshInfo, base_color_factor) == 92, "MeshInfo.base_color_factor");
LH_STATIC_ASSERT(sizeof(Instance) == 112, "Instance size");
LH_STATIC_ASSERT(offsetof(Instance, mesh_info_address) == 64, "Instance.mesh_info_address");
LH_STATIC_ASSERT(offsetof(Instance, normal_transform) == 72, "Instance.normal_transform");
LH_STATIC_ASSERT(sizeof(Meshlet) == 56, "Meshlet size");
LH_STATIC_ASSERT(offsetof(Meshlet, cone_axis) == 12, "Meshlet.cone_axis");
LH_STATIC_ASSERT(offsetof(Meshlet, cone_cutoff) == 24, "Meshlet.cone_cutoff");
LH_STATIC_ASSERT(offsetof(Meshlet, bounding_sphere) == 28, "Meshlet.bounding_sphere");
LH_STATIC_ASSERT(offsetof(Meshlet, triangle_offset) == 44, "Meshlet.triangle_offset");
LH_STATIC_ASSERT(offsetof(Meshlet, index_offset) == 48, "Meshlet.index_offset");
LH_STATIC_ASSERT(offsetof(Meshlet, triangle_count) == 52, "Meshlet.triangle_count");
LH_STATIC_ASSERT(offsetof(Meshlet, index_count) == 53, "Meshlet.index_count");
LH_STATIC_ASSERT(sizeof(MeshletReference) == 8, "MeshletReference size");
LH_STATIC_ASSERT(sizeof(Uniforms) == 560, "Uniforms size");
LH_STATIC_ASSERT(offsetof(Uniforms, meshlet_references) == 448, "Uniforms.meshlet_references");
LH_STATIC_ASSERT(offsetof(Uniforms, dispatches) == 488, "Uniforms.dispatches");
LH_STATIC_ASSERT(offsetof(Uniforms, camera_pos) == 496, "Uniforms.camera_pos");
LH_STATIC_ASSERT(offsetof(Uniforms, sun_dir) == 508, "Uniforms.sun_dir");
LH_STATIC_ASSERT(offsetof(Uniforms, sun_intensity) == 520, "Uniforms.sun_intensity");
LH_STATIC_ASSERT(offsetof(Uniforms, num_instances) == 532, "Uniforms.num_instances");
LH_STATIC_ASSERT(offsetof(Uniforms, window_size) == 536, "Uniforms.window_size");
LH_STATIC_ASSERT(offsetof(Uniforms, shadow_cam_distance) == 544, "Uniforms.shadow_cam_distance");
LH_STATIC_ASSERT(offsetof(Uniforms, cascade_split_pow) == 548, "Uniforms.cascade_split_pow");
LH_STATIC_ASSERT(offsetof(Uniforms, debug) == 552, "Uniforms.debug");
LH_STATIC_ASSERT(offsetof(Uniforms, debug_shadowmaps) == 556, "Uniforms.debug_shadowmaps");
LH_STATIC_ASSERT(sizeof(MiscStorage) == 792, "MiscStorage size");
LH_STATIC_ASSERT(offsetof(MiscStorage, shadow_sphere_radii) == 768, "MiscStorage.radii");
LH_STATIC_ASSERT(offsetof(MiscStorage, min_depth) == 784, "MiscStorage.min_depth");
LH_STATIC_ASSERT(offsetof(MiscStorage, max_depth) == 788, "MiscStorage.max_depth");
LH_STATIC_ASSERT(sizeof(PrefixSumValue) == 8, "PrefixSumValue size");
LH_STATIC_ASSERT(sizeof(DrawIndirectCommand) == 16, "DrawIndirectCommand size");
LH_STATIC_ASSERT(sizeof(DispatchIndirectCommand) == 12, "DispatchIndirectCommand size");
LH_STATIC_ASSERT(sizeof(LhTexture) == 88, "LhTexture size");

#endif /* LH_SHARED_CPU_GPU_H */
