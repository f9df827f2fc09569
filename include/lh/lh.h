/*
 * C ABI of the MI355X GPU-driven frame (liblh.so).
 *
 * The reference records its frame with Vulkan in render()
 * (src/rendering.h:4-16, src/rendering.cpp:28-548); every pass there is a
 * compute dispatch or an indirect draw that reads the shared structs through
 * buffer-device addresses and push constants
 * {u64 uniforms_address; u32 cascade_index} (src/shaders/common/bindings.glsl:23-26).
 * Here every pass is an extern "C" function taking
 *   - lh_scene: device addresses of the static scene (instances, mesh infos,
 *     meshlet tables, textures) - the reference's InstanceBuffer /
 *     MeshInfoBuffer / bindless texture table;
 *   - lh_view: device addresses of one view's working set (uniforms, misc
 *     storage, prefix-sum sectors, draw lists, visibility buffer, shadow maps,
 *     radiance) - the reference's per-frame resources;
 *   - a HIP stream (passed as void*); passes are stream ordered, so the
 *     reference's pipeline barriers become plain stream order.
 * Nothing here allocates device memory: all buffers are caller-owned, so a
 * sequence of calls can be captured into a hipGraph.  All structs hold plain
 * device pointers and sizes; no C++ or torch types cross the boundary.
 *
 * Pass numbering: pass 0 = main view, pass 1 + c = shadow cascade c (c < 4).
 */
#ifndef LH_LH_H
#define LH_LH_H

#include "shared_cpu_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lh_status;
#define LH_OK 0
#define LH_ERR_INVALID 1     /* bad argument / size out of range */
#define LH_ERR_CAPACITY 2    /* a capacity-bounded buffer overflowed (see counters) */
#define LH_ERR_HIP 3         /* a HIP runtime call failed */
#define LH_ERR_UNSUPPORTED 4 /* feature outside this build (e.g. alpha clip) */

typedef void* lh_stream; /* hipStream_t */

#define LH_NUM_PASSES 5
#define LH_NUM_CASCADES 4
#define LH_MAX_VIEW_SIDE 8192u
#define LH_MAX_BATCH_VIEWS 256u
#define LH_MAX_BATCH_SEGMENTS (LH_MAX_BATCH_VIEWS * LH_NUM_CASCADES)

/* Device counters of one view (u32 each). */
#define LH_CNT_DRAWS 0     /* + pass: surviving meshlets of the pass (may exceed draw_capacity) */
#define LH_CNT_BIG 8       /* + pass: triangles queued for the large-triangle raster */
#define LH_CNT_CLIP 16     /* + pass: triangles queued for clipping */
#define LH_CNT_OVERFLOW 24 /* bit mask of LH_OVERFLOW_* */
#define LH_CNT_FLAGS 25    /* bit mask of LH_FLAG_* */
#define LH_CNT_HIZ_CANDIDATES 26 /* LH_VIEW_HIZ: small meshlets tested against the depth pyramid */
#define LH_CNT_HIZ_PHASE1 27     /* LH_VIEW_HIZ: main-view draws rasterized before the pyramid */
#define LH_CNT_FRAGMENTS 32 /* + pass: covered samples (only with LH_VIEW_STATISTICS) */
#define LH_CNT_ATOMICS 40   /* + pass: depth atomics issued (same; main view: samples with depth > 0) */
#define LH_CNT_ALPHA_DRAWS 48 /* + pass: surviving meshlets of alpha-clip mesh infos (with LH_VIEW_HIZ: before
                                 the occlusion test); zero skips the alpha-clip raster of the pass */
#define LH_CNT_TILE_ITEMS 56   /* tiled main-view raster: work items (tile, run of binned meshlets) of the last launch */
#define LH_CNT_TILE_NEXT 57     /* tiled main-view raster: work-item queue of the last launch (items handed
                                  out beyond the grid's first) */
#define LH_CNT_TILE_DIRECT 58  /* tiled main-view raster: meshlets left to the per-meshlet raster (beyond
                                  bin_capacity, or a screen rectangle over LH_TILE_MAX_RECT tiles) */
#define LH_CNT_BACKWARD_TICKET 60 /* lh_resolve_backward: blocks done; zero between calls (the view reset
                                     zeroes it, the call's last block sets it back to zero) */
#define LH_NUM_COUNTERS 64
#define LH_OVERFLOW_INSTANCES 0x1u
#define LH_OVERFLOW_DRAWS 0x2u
#define LH_OVERFLOW_BIG 0x4u
#define LH_OVERFLOW_CLIP 0x8u
#define LH_FLAG_EMPTY_VIEW 0x1u /* no geometry: depth bounds stayed at their reset values */

/* lh_view.flags */
#define LH_VIEW_STATISTICS 0x1u /* single-view raster also counts LH_CNT_FRAGMENTS / LH_CNT_ATOMICS
                                   (slower; batched launches ignore it) */
#define LH_VIEW_HIZ 0x2u        /* two-phase occlusion culling of the main view (single view only):
                                   phase 1 rasterizes the meshlets with a large screen footprint and
                                   those visible in this view's last frame (lh_view.history); the rest
                                   are tested against a depth pyramid of phase 1 and skipped when
                                   hidden.  The final keys, shadows and radiance are those of the frame
                                   without it; draws[0] then lists only the meshlets actually
                                   rasterized, and the candidates are staged in draws[1] (free until
                                   the shadow passes). */

/* Raster scratch inside lh_view.scratch, each part 256-byte aligned, sizes from
 * lh_scratch_bytes(view):
 *   clip queue     clip_capacity x 8 B
 *   big queue      big_capacity x 48 B
 *   tiled main-view raster (only when bin_capacity > 0):
 *     bin headers  bin_capacity x 16 B      (screen rectangle, triangle count, meshlet id per draw)
 *     bin cache    bin_capacity x 1280 B    (snapped vertices and front-facing triangles per draw)
 *     tile counts  2 x tiles x 4 B          (tiles = ceil(width / 64) * ceil(height / 64))
 *     items        (tiles + entries / 8) x 12 B
 *     entries      LH_TILE_MAX_RECT * bin_capacity x 4 B  ((meshlet, tile) pairs)
 *     direct list  draw_capacity x 4 B      (draws left to the per-meshlet raster) */
#define LH_CLIP_ENTRY_BYTES 8u
#define LH_BIG_ENTRY_BYTES 48u
#define LH_TILE_SIDE 64u       /* pixels per side of a raster tile (one LDS tile of keys) */
#define LH_TILE_MAX_RECT 16u   /* binned meshlets cover at most this many tiles (larger ones take the
                                  per-meshlet path) */
#define LH_BIN_CACHE_BYTES 1280u

typedef struct lh_scene {
    const Instance* instances;              /* [num_instances] */
    const MeshInfo* mesh_infos;             /* [num_mesh_infos]; Instance.mesh_info_address points here */
    const LhMaterialFactors* factors;       /* [num_mesh_infos] */
    const uint32_t* instance_meshlet_base;  /* [num_instances] exclusive scan of num_meshlets */
    const MeshletReference* meshlet_table;  /* [total_meshlets] stable meshlet id -> (instance, meshlet) */
    const LhTexture* textures;              /* [num_textures] */
    uint32_t num_instances;
    uint32_t num_mesh_infos;
    uint32_t total_meshlets;
    uint32_t num_textures;
    uint32_t flags;                         /* LH_SCENE_* */
} lh_scene;
/* The caller states that no mesh info has MESH_INFO_FLAGS_ALPHA_CLIP: the
 * alpha-clip raster launches (double-sided, alpha test per sample) are
 * skipped.  Without it they run and return at once when no alpha-clip
 * meshlet survived the culls; with it, alpha-clip meshlets are not drawn. */
#define LH_SCENE_OPAQUE_ONLY 0x1u

typedef struct lh_view {
    const Uniforms* uniforms;   /* device copy of this view's uniforms */
    MiscStorage* misc;          /* cascade matrices + depth bounds */
    uint64_t* prefix_sums;      /* LH_NUM_PASSES sectors of (1 + instance_capacity) u64 words:
                                   word 0 = counter (count << 32 | sum), then PrefixSumValue[] */
    MeshletReference* draws;    /* LH_NUM_PASSES lists of draw_capacity entries */
    uint32_t* counters;         /* LH_NUM_COUNTERS */
    uint64_t* visbuffer;        /* width * height keys */
    uint32_t* shadow_maps;      /* LH_NUM_CASCADES * shadow_size^2 float bits */
    float* radiance;            /* width * height * 4 (RGBA, FP32) */
    uint32_t* ldr;              /* width * height BGRA8 after tonemap; may be NULL */
    void* scratch;              /* raster queues, see LH_*_ENTRY_BYTES */
    uint64_t scratch_bytes;
    uint32_t instance_capacity;
    uint32_t draw_capacity;
    uint32_t width;
    uint32_t height;
    uint32_t shadow_size;       /* cascade side (reference: 1024, src/main.cpp:385-409) */
    uint32_t big_capacity;
    uint32_t clip_capacity;
    uint32_t flags;             /* LH_VIEW_* bits */
    uint32_t* history;          /* LH_VIEW_HIZ: one bit per stable meshlet id (total_meshlets bits),
                                   set for the meshlets that own a pixel of the last frame of this
                                   view; may be NULL.  Zero it when the view is created. */
    const float* tonemap_lut;   /* optional display LUT (tony_mc_mapface, display_transform.comp:5-17):
                                   LH_TONEMAP_LUT_DIMS^3 RGBA float entries, red fastest, indexed by the
                                   stimulus encoded as x / (x + 1); NULL: that encoding is the curve */
    uint32_t bin_capacity;      /* tiled main-view raster: draws binned into 64x64 screen tiles (the first
                                   bin_capacity surviving meshlets; later ones take the per-meshlet path);
                                   0 = per-meshlet raster only.  Keys are identical either way. */
    uint32_t skybox_size;       /* side of the skybox faces; 0 = constant sky */
    const float* skybox;        /* optional skybox cubemap (render_geometry.comp:86-101): 6 faces in the order
                                   +X, -X, +Y, -Y, +Z, -Z, each skybox_size^2 RGBA float texels, rows top
                                   to bottom; sampled at LOD 0 times 2^3 exposure where nothing is drawn */
} lh_view;
#define LH_TONEMAP_LUT_DIMS 48u

/* Bytes of scratch the raster needs for the view's sizes and capacities
 * (width, height, draw_capacity, big_capacity, clip_capacity, bin_capacity). */
uint64_t lh_scratch_bytes(const lh_view* view);

/* reset_buffers_a (src/shaders/compute/reset_buffers.glsl:9-29) + the
 * attachment clears of src/rendering.cpp:206-211,384: counters and prefix-sum
 * sectors to 0, min_depth = UINT_MAX, max_depth = 0, visibility buffer to
 * LH_VISBUFFER_CLEAR, shadow maps to LH_SHADOW_CLEAR. */
lh_status lh_view_reset(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* cull_instances (src/shaders/cull_instances.comp:7-27): sphere frustum test,
 * 64-bit-atomic inclusive prefix-sum append of num_meshlets into sector 0
 * (src/shaders/common/prefix_sum.glsl:17-30). */
lh_status lh_cull_instances(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* write_draw_calls (pass 0, src/shaders/write_draw_calls.comp:7-57) and
 * write_draw_calls_shadows (pass 1 + c, :61-117): per meshlet of the visible
 * instances, sphere + normal-cone culling and compaction into draws[pass]. */
lh_status lh_cull_meshlets(const lh_scene* scene, const lh_view* view, uint32_t pass, lh_stream stream);

/* Visibility-buffer raster of draws[0] (replaces visbuffer_opaque_vertex/pixel,
 * src/shaders/rasterization.glsl:38-58, drawn at src/rendering.cpp:235-242):
 * 64-bit atomicMin of depth_key << 32 | meshlet << 7 | triangle. */
lh_status lh_raster_view(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* read_depth (src/shaders/compute/read_depth.glsl:22-83): min over non-zero
 * and max of the depth bits into misc->min_depth / max_depth. */
lh_status lh_depth_bounds(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* generate_shadow_matrices (src/shaders/compute/generate_shadow_matrices.glsl:15-110). */
lh_status lh_fit_cascades(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* cull_instances_shadows (src/shaders/cull_instances.comp:31-56): all four
 * cascades, appending into sectors 1..4. */
lh_status lh_cull_instances_shadow(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* Depth-only shadow raster of draws[1 + cascade] (shadowmap_opaque_vertex,
 * src/shaders/rasterization.glsl:101-111; front faces culled, compare LESS). */
lh_status lh_raster_shadow(const lh_scene* scene, const lh_view* view, uint32_t cascade, lh_stream stream);

/* render_geometry (src/shaders/render_geometry.comp:70-266): visibility-buffer
 * resolve into FP32 radiance; when view->ldr is non-NULL the display
 * transform (src/shaders/display_transform.comp:37-81) is fused in. */
lh_status lh_resolve(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* display_transform as a separate pass: radiance -> BGRA8 ldr. */
lh_status lh_tonemap(const lh_view* view, lh_stream stream);

/* ---- differentiable resolve (no counterpart in the reference) ----
 * The resolve as a function of explicit parameters, with the visibility
 * buffer (the triangle seen by each pixel) held fixed.  Each non-NULL pointer
 * of lh_resolve_params replaces the value the scene / uniforms hold; with all
 * pointers NULL lh_resolve_ex is lh_resolve, bit for bit. */
typedef struct lh_resolve_params {
    const float* base_color_factor; /* [num_mesh_infos * 3]  (MeshInfo.base_color_factor) */
    const float* roughness_factor;  /* [num_mesh_infos]      (LhMaterialFactors) */
    const float* metallic_factor;   /* [num_mesh_infos] */
    const float* sun_dir;           /* [3]                   (Uniforms.sun_dir, used as given) */
    const float* sun_intensity;     /* [3]                   (Uniforms.sun_intensity) */
    uint32_t* frozen;               /* optional [width * height * 4]: lh_resolve_ex stores, per pixel, the
                                       quantities the adjoint holds fixed (shadow factor, back-face flip,
                                       texture LOD); lh_resolve_backward then reads them instead of
                                       recomputing them.  Pass the buffer the forward filled. */
} lh_resolve_params;
lh_status lh_resolve_ex(const lh_scene* scene, const lh_view* view, const lh_resolve_params* params,
                        lh_stream stream);

/* Gradients of sum(grad_radiance * radiance) (radiance RGB of lh_resolve_ex;
 * alpha ignored), ADDED into the non-NULL outputs.  Vertex positions and
 * normals: the gradient of the FP32 component stored at device address A of
 * the geometry arena [geometry_base, geometry_base + geometry_bytes) is added
 * at geometry[(A - geometry_base) / 4], so the gradient arena mirrors the
 * geometry arena (meshes shared by several mesh infos / instances accumulate).
 * Held fixed, like visibility: the back-face flip, the cascade choice and PCF
 * shadow factor, and the texture LOD (mip levels and their blend weight).
 * Mesh infos with 16-bit vertices, a normal texture or alpha clip are skipped
 * and flagged with LH_FLAG_BACKWARD_UNSUPPORTED in the view's counters.  The
 * view's LH_CNT_BACKWARD_TICKET counter must be zero when the call starts, as
 * the view reset and every completed call leave it; calls on one view must not
 * overlap.  The view's scratch holds the per-block partial sums (it must be
 * at least 128 KiB). */
typedef struct lh_resolve_grads {
    float* base_color_factor; /* [num_mesh_infos * 3] */
    float* roughness_factor;  /* [num_mesh_infos] */
    float* metallic_factor;   /* [num_mesh_infos] */
    float* sun_dir;           /* [3] */
    float* sun_intensity;     /* [3] */
    float* geometry;          /* [geometry_bytes / 4] */
    uint64_t geometry_base;
    uint64_t geometry_bytes;
} lh_resolve_grads;
#define LH_FLAG_BACKWARD_UNSUPPORTED 0x2u
lh_status lh_resolve_backward(const lh_scene* scene, const lh_view* view, const lh_resolve_params* params,
                              const float* grad_radiance, const lh_resolve_grads* grads, lh_stream stream);

/* The whole frame in the reference's pass order (src/rendering.cpp:79-502). */
lh_status lh_frame(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* lh_frame with the raster scratch in a device array that the library owns
 * instead of view->scratch.  No counterpart in the reference: this entry
 * point is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc); it
 * allocates nothing.  view->scratch and view->scratch_bytes are ignored;
 * lh_scratch_bytes(view) must not exceed LH_STATIC_SCRATCH_BYTES (else
 * LH_ERR_CAPACITY).  The array is part of the library's device code, one per
 * device, held until the process ends; its address is looked up once per
 * device on the first call made with that device current, and `stream` must
 * belong to the current device.  Every call on one device shares the array
 * (it keeps the last call's contents; the frame writes every word it reads
 * first): calls that overlap on one device, from two streams or threads,
 * race. */
#define LH_STATIC_SCRATCH_BYTES (64ull << 20)
lh_status lh_frame_noalloc(const lh_scene* scene, const lh_view* view, lh_stream stream);

/* lh_frame captured once into a hipGraph and replayed: one graph launch per
 * frame instead of the frame's separate kernel launches, with bit-identical
 * results.  It plays the part of the reference's per-frame command buffer
 * (src/rendering.cpp:28-548, re-recorded there every frame).  The scene and
 * view structs are captured by value at lh_frame_graph_create: the device
 * memory they point to (uniforms and camera, instances, buffers) may change
 * between runs, but changing a field of either struct needs a new graph.
 * Capture happens on a stream the graph owns (default priority); runs may
 * use any stream of the same device. */
typedef struct lh_frame_graph lh_frame_graph;
lh_status lh_frame_graph_create(const lh_scene* scene, const lh_view* view, lh_frame_graph** out);
lh_status lh_frame_graph_run(lh_frame_graph* graph, lh_stream stream);
void lh_frame_graph_destroy(lh_frame_graph* graph);

/* lh_frame with per-kernel device times: an event is recorded on `stream`
 * before every kernel launch, the call synchronizes on the last one and
 * writes one entry per kernel (in launch order).  No counterpart in the
 * reference (its per-pass GPU zones are Tracy's, src/rendering.cpp:42-512). */
typedef struct lh_kernel_time {
    char name[40];
    float ms;
} lh_kernel_time;
lh_status lh_frame_profiled(const lh_scene* scene, const lh_view* view, lh_stream stream, lh_kernel_time* out,
                            uint32_t capacity, uint32_t* count);

/* ---- batched views (no counterpart in the reference, which renders one view
 * per frame, src/rendering.cpp:28-548) ----
 * A batch is a fixed set of up to LH_MAX_BATCH_VIEWS views over one scene,
 * each view a camera with its own four cascades and buffers (caller-owned, as
 * for lh_frame).  lh_batch_create copies the view table to the device (the
 * only allocation; nothing is allocated per frame).  lh_batch_run executes one
 * frame of every view either as the per-view loop (LH_BATCH_LOOP: lh_frame per
 * view, in order) or batched (LH_BATCH_BATCHED: each pass launched once for
 * all views, flattened, load-balanced grids); with use_graph != 0 the sequence
 * is captured into a hipGraph on first use and replayed.  Both modes produce
 * bit-identical keys, shadow depths and radiance for every view. */
typedef struct lh_batch lh_batch;
#define LH_BATCH_LOOP 0u
#define LH_BATCH_BATCHED 1u
lh_status lh_batch_create(const lh_scene* scene, const lh_view* views, uint32_t n_views, lh_batch** out);
lh_status lh_batch_run(lh_batch* batch, uint32_t mode, uint32_t use_graph, lh_stream stream);
/* One un-captured frame of the batch with per-kernel device times (see lh_frame_profiled). */
lh_status lh_batch_profiled(lh_batch* batch, uint32_t mode, lh_stream stream, lh_kernel_time* out,
                            uint32_t capacity, uint32_t* count);
void lh_batch_destroy(lh_batch* batch);

/* Copies the view's counters to host memory (synchronizes the stream) and
 * returns LH_ERR_CAPACITY if any overflow bit is set. */
lh_status lh_view_status(const lh_view* view, uint32_t* host_counters, lh_stream stream);

/* Launch tunables of the culling and raster kernels (no counterpart in the
 * reference, whose raster is fixed-function).  Results are independent of
 * them; the autotuner picks them per resolution class.  Process-global. */
typedef struct lh_tuning {
    uint32_t small_extent_px;  /* triangles whose vertex bbox fits in this many pixels (1..32) are
                                  rasterized by one lane */
    uint32_t meshlet_blocks;   /* persistent grid of the per-meshlet raster (one wave per block) */
    uint32_t big_blocks;       /* persistent grid of the large-triangle raster */
    uint32_t clip_blocks;      /* persistent grid of the clipping kernel */
    uint32_t cull_blocks;      /* persistent grid of the meshlet cull */
    uint32_t medium_extent_px; /* up to this extent (small_extent_px..64) the whole wave rasterizes
                                  the triangle in place; larger ones go to the big-triangle queue */
    uint32_t tile_item_entries; /* tiled main-view raster: binned meshlets per work item (8..512; 0 = 64) */
    uint32_t tile_blocks;       /* tiled main-view raster: persistent grid of 512-thread blocks (0 = 1024) */
    uint32_t tile_raster;       /* 1: the main view goes through the tiled raster (when its bin_capacity
                                   is non-zero); 0: the per-meshlet raster.  The tiled raster wins when
                                   meshlets cover many pixels each; with a few pixels per meshlet (small
                                   views of dense scenes) the per-meshlet raster is cheaper. */
    uint32_t shadow_small_extent_px;  /* the cascades' small_extent_px (0..32; 0 = small_extent_px): the
                                         shadow maps' triangles and raster differ from the main view's */
    uint32_t shadow_medium_extent_px; /* the cascades' medium_extent_px (0..64; 0 = medium_extent_px) */
} lh_tuning;
/* lh_set_tuning sets every class at once, lh_get_tuning reads class 1;
 * lh_set_tuning_class / lh_get_tuning_class set or read one class.  Classes by
 * main-view pixels: 0 < 0.6 MP, 1 < 2.5 MP, 2 < 9 MP, 3 larger; class 4 is used
 * by batched launches (lh_batch_run with LH_BATCH_BATCHED). */
#define LH_TUNING_CLASSES 5
lh_status lh_set_tuning(const lh_tuning* tuning);
lh_status lh_get_tuning(lh_tuning* tuning);
lh_status lh_set_tuning_class(uint32_t tuning_class, const lh_tuning* tuning);
lh_status lh_get_tuning_class(uint32_t tuning_class, lh_tuning* tuning);
uint32_t lh_tuning_class_of(uint32_t width, uint32_t height);

/* Library version string. */
const char* lh_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LH_LH_H */
