"""numpy / ctypes mirrors of include/lh/shared_cpu_gpu.h, lh_scene.h and lh.h.

Offsets follow the reference's shared header (src/shared_cpu_gpu.h:18-157)
and are checked against the C static assertions by the CPU test suite.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

MESH_INFO_FLAGS_32_BIT_INDICES = 0x01
MESH_INFO_FLAGS_ALPHA_CLIP = 0x02
MESH_INFO_FLAGS_FP32_VERTICES = 0x04
SCENE_OPAQUE_ONLY = 0x1  # LH_SCENE_OPAQUE_ONLY
TONEMAP_LUT_DIMS = 48  # LH_TONEMAP_LUT_DIMS
TEXTURE_RGBA8, TEXTURE_BC1, TEXTURE_BC7 = 0, 1, 2  # LhTexture.format storage
TEXTURE_FORMAT_MASK, TEXTURE_SRGB, TEXTURE_RGB = 0xFF, 0x100, 0x200  # flags: sRGB-encoded RGB, alpha read as 1
# Uniforms.debug (src/shared_cpu_gpu.h:46-51)
(UNIFORMS_DEBUG_OFF, UNIFORMS_DEBUG_CASCADES, UNIFORMS_DEBUG_TRIANGLE_INDEX, UNIFORMS_DEBUG_INSTANCE_INDEX,
 UNIFORMS_DEBUG_SHADER_CLOCK, UNIFORMS_DEBUG_NORMALS) = range(6)
UNUSED_TEXTURE_INDEX = 0xFFFF
MAX_MESHLET_UNIQUE_VERTICES = 64
MAX_MESHLET_TRIANGLES = 124
LH_VISBUFFER_CLEAR = 0xFFFFFFFFFFFFFFFF
LH_SHADOW_CLEAR = 0x3F800000
LH_KEY_TRI_BITS = 7
LH_KEY_MESHLET_BITS = 25

PRESET_SMALL, PRESET_LARGE, PRESET_STRESS = 1, 2, 3
VERTEX_FORMAT_FP32, VERTEX_FORMAT_QUANTIZED = 0, 1  # LH_VERTEX_FORMAT_* (lh_scene.h)


def _dt(fields, itemsize):
    names, formats, offsets = zip(*fields)
    return np.dtype({"names": list(names), "formats": list(formats), "offsets": list(offsets), "itemsize": itemsize})


MAT4 = ("<f4", (4, 4))  # [column][row]
MAT3 = ("<f4", (3, 3))

MESH_INFO = _dt(
    [
        ("positions", "<u8", 0), ("indices", "<u8", 8), ("normals", "<u8", 16), ("uvs", "<u8", 24),
        ("micro_indices", "<u8", 32), ("meshlets", "<u8", 40), ("num_meshlets", "<u2", 48), ("flags", "u1", 50),
        ("bounding_sphere", ("<f4", 4), 52), ("texture_scale", ("<f4", 2), 68), ("texture_offset", ("<f4", 2), 76),
        ("base_color_texture_index", "<u2", 84), ("metallic_roughness_texture_index", "<u2", 86),
        ("normal_texture_index", "<u2", 88), ("base_color_factor", ("<f4", 3), 92),
    ],
    104,
)
INSTANCE = _dt([("transform", MAT4, 0), ("mesh_info_address", "<u8", 64), ("normal_transform", MAT3, 72)], 112)
MESHLET = _dt(
    [
        ("cone_apex", ("<f4", 3), 0), ("cone_axis", ("<f4", 3), 12), ("cone_cutoff", "<f4", 24),
        ("bounding_sphere", ("<f4", 4), 28), ("triangle_offset", "<u4", 44), ("index_offset", "<u4", 48),
        ("triangle_count", "u1", 52), ("index_count", "u1", 53),
    ],
    56,
)
MESHLET_REFERENCE = _dt([("instance_index", "<u4", 0), ("meshlet_index", "<u2", 4)], 8)
PREFIX_SUM_VALUE = _dt([("index", "<u4", 0), ("sum", "<u4", 4)], 8)
MATERIAL_FACTORS = _dt([("roughness_factor", "<f4", 0), ("metallic_factor", "<f4", 4)], 8)
TEXTURE = _dt(
    [("texels", "<u8", 0), ("width", "<u4", 8), ("height", "<u4", 12), ("levels", "<u4", 16), ("format", "<u4", 20),
     ("level_offset", ("<u4", 16), 24)],
    88,
)
UNIFORMS = _dt(
    [
        ("combined_perspective_view", MAT4, 0), ("inv_perspective_view", MAT4, 64), ("view", MAT4, 128),
        ("initial_view", MAT4, 192), ("perspective", MAT4, 256), ("perspective_inverse", MAT4, 320),
        ("view_inverse", MAT4, 384), ("meshlet_references", "<u8", 448), ("instances", "<u8", 456),
        ("draw_calls", "<u8", 464), ("misc_storage", "<u8", 472), ("num_meshlets_prefix_sum", "<u8", 480),
        ("dispatches", "<u8", 488), ("camera_pos", ("<f4", 3), 496), ("sun_dir", ("<f4", 3), 508),
        ("sun_intensity", ("<f4", 3), 520), ("num_instances", "<u4", 532), ("window_size", ("<u4", 2), 536),
        ("shadow_cam_distance", "<f4", 544), ("cascade_split_pow", "<f4", 548), ("debug", "<i4", 552),
        ("debug_shadowmaps", "u1", 556),
    ],
    560,
)
MISC_STORAGE = _dt(
    [
        ("shadow_matrices", ("<f4", (4, 4, 4)), 0), ("uv_space_shadow_matrices", ("<f4", (4, 4, 4)), 256),
        ("shadow_view_matrices", ("<f4", (4, 4, 4)), 512), ("shadow_sphere_radii", ("<f4", 4), 768),
        ("min_depth", "<u4", 784), ("max_depth", "<u4", 788),
    ],
    792,
)


class Vec2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def tolist(self):
        return [self.x, self.y, self.z]


class Camera(C.Structure):
    """lh_camera (include/lh/lh_scene.h)."""

    _fields_ = [
        ("position", Vec3), ("target", Vec3), ("fov_y_rad", C.c_float), ("near_plane", C.c_float),
        ("width", C.c_uint32), ("height", C.c_uint32), ("sun_dir", Vec3), ("sun_intensity", Vec3),
        ("shadow_cam_distance", C.c_float), ("cascade_split_pow", C.c_float),
    ]


class PresetConfig(C.Structure):
    _fields_ = [("preset", C.c_uint32), ("num_instances", C.c_uint32), ("seed", C.c_uint64),
                ("texture_size", C.c_uint32), ("mesh_tri_scale_pct", C.c_uint32), ("vertex_format", C.c_uint32),
                ("texture_format", C.c_uint32), ("cache_dir", C.c_char_p)]


class SceneArrays(C.Structure):
    _fields_ = [
        ("geometry", C.c_void_p), ("geometry_bytes", C.c_uint64), ("mesh_infos", C.c_void_p),
        ("num_mesh_infos", C.c_uint32), ("factors", C.c_void_p), ("instances", C.c_void_p),
        ("num_instances", C.c_uint32), ("instance_meshlet_base", C.c_void_p), ("meshlet_table", C.c_void_p),
        ("total_meshlets", C.c_uint32), ("texels", C.c_void_p), ("texel_bytes", C.c_uint64),
        ("textures", C.c_void_p), ("num_textures", C.c_uint32), ("num_meshes", C.c_uint32),
        ("total_triangles", C.c_uint64), ("unique_triangles", C.c_uint64),
    ]


class LhScene(C.Structure):
    """lh_scene (include/lh/lh.h): static scene, device addresses."""

    _fields_ = [
        ("instances", C.c_uint64), ("mesh_infos", C.c_uint64), ("factors", C.c_uint64),
        ("instance_meshlet_base", C.c_uint64), ("meshlet_table", C.c_uint64), ("textures", C.c_uint64),
        ("num_instances", C.c_uint32), ("num_mesh_infos", C.c_uint32), ("total_meshlets", C.c_uint32),
        ("num_textures", C.c_uint32), ("flags", C.c_uint32),
    ]


class LhView(C.Structure):
    """lh_view (include/lh/lh.h): one view's working set, device addresses."""

    _fields_ = [
        ("uniforms", C.c_uint64), ("misc", C.c_uint64), ("prefix_sums", C.c_uint64), ("draws", C.c_uint64),
        ("counters", C.c_uint64), ("visbuffer", C.c_uint64), ("shadow_maps", C.c_uint64),
        ("radiance", C.c_uint64), ("ldr", C.c_uint64), ("scratch", C.c_uint64), ("scratch_bytes", C.c_uint64),
        ("instance_capacity", C.c_uint32), ("draw_capacity", C.c_uint32), ("width", C.c_uint32),
        ("height", C.c_uint32), ("shadow_size", C.c_uint32), ("big_capacity", C.c_uint32),
        ("clip_capacity", C.c_uint32), ("flags", C.c_uint32), ("history", C.c_uint64), ("tonemap_lut", C.c_uint64),
        ("bin_capacity", C.c_uint32), ("skybox_size", C.c_uint32), ("skybox", C.c_uint64),
    ]


NUM_PASSES = 5
NUM_CASCADES = 4
NUM_COUNTERS = 64
CNT_DRAWS, CNT_BIG, CNT_CLIP, CNT_OVERFLOW, CNT_FLAGS = 0, 8, 16, 24, 25
CNT_FRAGMENTS, CNT_ATOMICS = 32, 40  # + pass, with VIEW_STATISTICS
CNT_ALPHA_DRAWS = 48  # + pass: surviving alpha-clip meshlets
VIEW_STATISTICS = 0x1
VIEW_HIZ = 0x2
CNT_HIZ_CANDIDATES, CNT_HIZ_PHASE1 = 26, 27
FLAG_EMPTY_VIEW, FLAG_BACKWARD_UNSUPPORTED = 0x1, 0x2
CLIP_ENTRY_BYTES, BIG_ENTRY_BYTES = 8, 48
LH_OK, LH_ERR_INVALID, LH_ERR_CAPACITY, LH_ERR_HIP, LH_ERR_UNSUPPORTED = 0, 1, 2, 3, 4


TILE_SIDE, TILE_MAX_RECT, BIN_CACHE_BYTES = 64, 16, 1280
STATIC_SCRATCH_BYTES = 64 << 20  # LH_STATIC_SCRATCH_BYTES: the scratch array of lh_frame_noalloc
TUNING_CLASSES = 5  # LH_TUNING_CLASSES
CNT_TILE_ITEMS, CNT_TILE_DIRECT = 56, 58
CNT_BACKWARD_TICKET = 60  # LH_CNT_BACKWARD_TICKET: zero between lh_resolve_backward calls


def scratch_layout(width: int, height: int, draw_capacity: int, big_capacity: int, clip_capacity: int,
                   bin_capacity: int) -> dict:
    """Byte offsets of the parts of lh_view.scratch (include/lh/lh.h): raster
    queues, then the tiled raster's bins when bin_capacity > 0, each part
    256-byte aligned; "end" is lh_scratch_bytes."""
    a = lambda n: (n + 255) & ~255  # noqa: E731
    parts = [("clip", clip_capacity * CLIP_ENTRY_BYTES), ("big", big_capacity * BIG_ENTRY_BYTES)]
    out = {}
    if bin_capacity:
        tiles_x, tiles_y = (width + TILE_SIDE - 1) // TILE_SIDE, (height + TILE_SIDE - 1) // TILE_SIDE
        entries = TILE_MAX_RECT * bin_capacity
        items = tiles_x * tiles_y + entries // 8
        parts += [("hdr", bin_capacity * 16), ("cache", bin_capacity * BIN_CACHE_BYTES),
                  ("tile_count", tiles_x * tiles_y * 4), ("tile_cursor", tiles_x * tiles_y * 4),
                  ("items", items * 12), ("entries", entries * 4), ("direct", draw_capacity * 4)]
        out.update(tiles_x=tiles_x, tiles_y=tiles_y, entry_capacity=entries, item_capacity=items)
    o = 0
    for name, nb in parts:
        out[name] = o
        o += a(nb)
    out["end"] = o
    return out


def scratch_bytes(width: int, height: int, draw_capacity: int, big_capacity: int, clip_capacity: int,
                  bin_capacity: int) -> int:
    """lh_scratch_bytes (include/lh/lh.h)."""
    return scratch_layout(width, height, draw_capacity, big_capacity, clip_capacity, bin_capacity)["end"]
