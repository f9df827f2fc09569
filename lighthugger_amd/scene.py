"""Python handle on the C++ scene producer (liblh_scene.so, include/lh/lh_scene.h).

The producer replaces the reference's asset ingestion for this path
(src/resources/mesh_loading.cpp:241-738, src/resources/meshlets.cpp:3-140);
this module only moves its flat arrays into numpy.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import abi

_PKG = Path(__file__).resolve().parent
_lib = None


def scene_lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = _PKG / "liblh_scene.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing; run __graft_entry__.build() first")
        lib = C.CDLL(str(path))
        lib.lh_scene_create.restype = C.c_void_p
        lib.lh_scene_destroy.argtypes = [C.c_void_p]
        lib.lh_scene_add_mesh.restype = C.c_int32
        lib.lh_scene_add_mesh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                          C.c_uint32]
        lib.lh_scene_add_texture.restype = C.c_int32
        lib.lh_scene_add_texture.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
        lib.lh_scene_add_texture_encoded.restype = C.c_int32
        lib.lh_scene_add_texture_encoded.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                     C.c_void_p, C.c_uint64]
        lib.lh_scene_add_material.restype = C.c_int32
        lib.lh_scene_add_material.argtypes = [C.c_void_p, abi.Vec3, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                                              C.c_uint32, abi.Vec2, abi.Vec2]
        lib.lh_scene_add_instance.restype = C.c_int32
        lib.lh_scene_add_instance.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        lib.lh_scene_set_material_alpha_clip.restype = C.c_int32
        lib.lh_scene_set_material_alpha_clip.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        lib.lh_scene_set_texture_format.restype = C.c_int32
        lib.lh_scene_set_texture_format.argtypes = [C.c_void_p, C.c_uint32]
        lib.lh_scene_set_cache_dir.restype = C.c_int32
        lib.lh_scene_set_cache_dir.argtypes = [C.c_void_p, C.c_char_p]
        lib.lh_scene_set_vertex_format.restype = C.c_int32
        lib.lh_scene_set_vertex_format.argtypes = [C.c_void_p, C.c_uint32]
        lib.lh_scene_finish.restype = C.c_int32
        lib.lh_scene_finish.argtypes = [C.c_void_p]
        lib.lh_scene_build_preset.restype = C.c_void_p
        lib.lh_scene_build_preset.argtypes = [C.POINTER(abi.PresetConfig)]
        lib.lh_preset_camera.restype = C.c_int32
        lib.lh_preset_camera.argtypes = [C.c_uint32, C.POINTER(abi.Camera)]
        lib.lh_batch_cameras.restype = C.c_int32
        lib.lh_batch_cameras.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(abi.Camera)]
        lib.lh_scene_get.restype = C.c_int32
        lib.lh_scene_get.argtypes = [C.c_void_p, C.POINTER(abi.SceneArrays)]
        lib.lh_scene_relocate.restype = C.c_int32
        lib.lh_scene_relocate.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        lib.lh_make_uniforms.restype = C.c_int32
        lib.lh_make_uniforms.argtypes = [C.POINTER(abi.Camera), C.c_uint32, C.c_void_p]
        lib.lh_make_transform.restype = None
        lib.lh_make_transform.argtypes = [abi.Vec3, C.c_float, C.c_float, C.c_void_p]
        _lib = lib
    return _lib


def _copy(ptr: int, dtype, count: int) -> np.ndarray:
    dtype = np.dtype(dtype)
    if count == 0:
        return np.zeros(0, dtype)
    buf = (C.c_uint8 * (dtype.itemsize * count)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def preset_camera(preset: int) -> abi.Camera:
    cam = abi.Camera()
    if scene_lib().lh_preset_camera(preset, C.byref(cam)) != 0:
        raise ValueError(f"unknown preset {preset}")
    return cam


def batch_cameras(seed: int, count: int, min_side: int = 256, max_side: int = 2048) -> list[abi.Camera]:
    arr = (abi.Camera * count)()
    if scene_lib().lh_batch_cameras(seed, count, min_side, max_side, arr) != 0:
        raise ValueError("bad batch camera arguments")
    return [arr[i] for i in range(count)]


def make_camera(position, target, width, height, *, fov_deg=60.0, near=0.1, sun_dir=(0.35, 0.85, 0.40),
                sun_intensity=(1.0, 1.0, 1.0), shadow_cam_distance=1024.0, cascade_split_pow=3.0) -> abi.Camera:
    s = np.asarray(sun_dir, np.float64)
    s = s / np.linalg.norm(s)
    return abi.Camera(abi.Vec3(*map(float, position)), abi.Vec3(*map(float, target)), math.radians(fov_deg), near,
                      width, height, abi.Vec3(*map(float, s)), abi.Vec3(*map(float, sun_intensity)),
                      shadow_cam_distance, cascade_split_pow)


def make_uniforms(cam: abi.Camera, num_instances: int) -> np.ndarray:
    u = np.zeros(1, abi.UNIFORMS)
    if scene_lib().lh_make_uniforms(C.byref(cam), num_instances, u.ctypes.data) != 0:
        raise ValueError("bad camera")
    return u


def make_transform(translation, yaw: float, scale: float) -> np.ndarray:
    m = np.zeros((4, 4), np.float32)
    scene_lib().lh_make_transform(abi.Vec3(*map(float, translation)), float(yaw), float(scale), m.ctypes.data)
    return m


@dataclass
class EncodedTexture:
    """A texture as a file stores it: `levels` holds the bytes of each mip
    level (level 0 first) in `format` (abi.TEXTURE_RGBA8 / BC1 / BC7, optionally
    | abi.TEXTURE_SRGB); block formats hold ceil(w/4) x ceil(h/4) blocks."""

    format: int
    width: int
    height: int
    levels: list


@dataclass
class Material:
    base_color: tuple = (1.0, 1.0, 1.0)
    roughness: float = 1.0
    metallic: float = 1.0
    base_color_texture: int = abi.UNUSED_TEXTURE_INDEX
    metallic_roughness_texture: int = abi.UNUSED_TEXTURE_INDEX
    normal_texture: int = abi.UNUSED_TEXTURE_INDEX
    texture_scale: tuple = (1.0, 1.0)
    texture_offset: tuple = (0.0, 0.0)
    alpha_clip: bool = False


class HostScene:
    """A finished scene in host memory (offsets in place of device addresses)."""

    def __init__(self, handle: int):
        if not handle:
            raise RuntimeError("scene build failed")
        self._h = C.c_void_p(handle)
        a = abi.SceneArrays()
        if scene_lib().lh_scene_get(self._h, C.byref(a)) != 0:
            raise RuntimeError("scene not finished")
        self.geometry = _copy(a.geometry, np.uint8, a.geometry_bytes)
        self.mesh_infos = _copy(a.mesh_infos, abi.MESH_INFO, a.num_mesh_infos)
        self.factors = _copy(a.factors, abi.MATERIAL_FACTORS, a.num_mesh_infos)
        self.instances = _copy(a.instances, abi.INSTANCE, a.num_instances)
        self.instance_meshlet_base = _copy(a.instance_meshlet_base, np.uint32, a.num_instances)
        self.meshlet_table = _copy(a.meshlet_table, abi.MESHLET_REFERENCE, a.total_meshlets)
        self.texels = _copy(a.texels, np.uint8, a.texel_bytes)
        self.textures = _copy(a.textures, abi.TEXTURE, a.num_textures)
        self.num_meshes = a.num_meshes
        self.total_triangles = a.total_triangles
        self.unique_triangles = a.unique_triangles

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.lh_scene_destroy(self._h)
            self._h = None

    @property
    def num_instances(self) -> int:
        return len(self.instances)

    @property
    def total_meshlets(self) -> int:
        return len(self.meshlet_table)

    @classmethod
    def preset(cls, preset: int, *, num_instances=0, seed=0, texture_size=0, tri_scale_pct=0,
               vertex_format=abi.VERTEX_FORMAT_FP32, texture_format=abi.TEXTURE_RGBA8, cache_dir=None) -> "HostScene":
        """cache_dir: an existing directory for the built-mesh cache (None: no cache)."""
        cfg = abi.PresetConfig(preset, num_instances, seed, texture_size, tri_scale_pct, vertex_format, texture_format,
                               None if cache_dir is None else str(cache_dir).encode())
        return cls(scene_lib().lh_scene_build_preset(C.byref(cfg)))

    @classmethod
    def build(cls, meshes, materials, instances, textures=(), vertex_format=abi.VERTEX_FORMAT_FP32,
              texture_format=abi.TEXTURE_RGBA8, cache_dir=None) -> "HostScene":
        """meshes: [(positions[N,3], normals[N,3] | None, uvs[N,2] | None, indices[T,3])];
        materials: [Material]; instances: [(mesh, material, transform[4,4] column-major)];
        textures: [uint8 RGBA array (H, W, 4) | EncodedTexture]; vertex_format: abi.VERTEX_FORMAT_*;
        texture_format: abi.TEXTURE_* (| abi.TEXTURE_SRGB), the storage of the RGBA textures;
        cache_dir: directory of the built-mesh cache or None."""
        lib = scene_lib()
        h = lib.lh_scene_create()
        try:
            if lib.lh_scene_set_vertex_format(h, vertex_format) != 0:
                raise ValueError(f"unknown vertex format {vertex_format}")
            if lib.lh_scene_set_texture_format(h, texture_format) != 0:
                raise ValueError(f"unknown texture format {texture_format}")
            if cache_dir is not None and lib.lh_scene_set_cache_dir(h, str(cache_dir).encode()) != 0:
                raise ValueError(f"cache directory {cache_dir} refused")
            for pos, nrm, uv, idx in meshes:
                pos = np.ascontiguousarray(pos, np.float32)
                idx = np.ascontiguousarray(idx, np.uint32)
                nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
                uv = None if uv is None else np.ascontiguousarray(uv, np.float32)
                r = lib.lh_scene_add_mesh(h, pos.ctypes.data, None if nrm is None else nrm.ctypes.data,
                                          None if uv is None else uv.ctypes.data, len(pos), idx.ctypes.data,
                                          len(idx))
                if r < 0:
                    raise ValueError("lh_scene_add_mesh failed")
            for tex in textures:
                if isinstance(tex, EncodedTexture):
                    data = np.ascontiguousarray(np.concatenate([np.frombuffer(bytes(lv), np.uint8)
                                                                for lv in tex.levels]))
                    if lib.lh_scene_add_texture_encoded(h, tex.format, tex.width, tex.height, len(tex.levels),
                                                        data.ctypes.data, data.nbytes) < 0:
                        raise ValueError("lh_scene_add_texture_encoded refused the texture")
                    continue
                tex = np.ascontiguousarray(tex, np.uint8)
                if lib.lh_scene_add_texture(h, tex.ctypes.data, tex.shape[1], tex.shape[0]) < 0:
                    raise ValueError("lh_scene_add_texture failed")
            for m in materials:
                r = lib.lh_scene_add_material(h, abi.Vec3(*m.base_color), m.roughness, m.metallic,
                                              m.base_color_texture, m.metallic_roughness_texture, m.normal_texture,
                                              abi.Vec2(*m.texture_scale), abi.Vec2(*m.texture_offset))
                if r < 0:
                    raise ValueError("lh_scene_add_material failed")
                if m.alpha_clip and lib.lh_scene_set_material_alpha_clip(h, r, 1) != 0:
                    raise ValueError("lh_scene_set_material_alpha_clip failed")
            for mesh, mat, tr in instances:
                tr = np.ascontiguousarray(tr, np.float32)
                if lib.lh_scene_add_instance(h, mesh, mat, tr.ctypes.data) < 0:
                    raise ValueError("lh_scene_add_instance failed")
            if lib.lh_scene_finish(h) != 0:
                raise ValueError("lh_scene_finish failed")
        except Exception:
            lib.lh_scene_destroy(h)
            raise
        return cls(h)

    def relocate(self, geometry_base: int, mesh_info_base: int, texel_base: int):
        mi = np.zeros(len(self.mesh_infos), abi.MESH_INFO)
        inst = np.zeros(len(self.instances), abi.INSTANCE)
        tex = np.zeros(max(1, len(self.textures)), abi.TEXTURE)
        if scene_lib().lh_scene_relocate(self._h, geometry_base, mesh_info_base, texel_base, mi.ctypes.data,
                                         inst.ctypes.data, tex.ctypes.data) != 0:
            raise RuntimeError("relocate failed")
        return mi, inst, tex[: len(self.textures)]

    # -- convenience views over the geometry arena (offsets are relative) --
    def meshlets(self, mesh_info: int) -> np.ndarray:
        mi = self.mesh_infos[mesh_info]
        off, n = int(mi["meshlets"]), int(mi["num_meshlets"])
        return np.frombuffer(self.geometry[off: off + n * 56].tobytes(), abi.MESHLET)

    def _fp32(self, mesh_info: int) -> bool:
        return bool(int(self.mesh_infos[mesh_info]["flags"]) & abi.MESH_INFO_FLAGS_FP32_VERTICES)

    def _attr(self, mesh_info: int, field: str, count: int, comps: int, qtype) -> np.ndarray:
        """A vertex attribute as the kernels read it, as float32: FP32 as
        stored, quantized values converted (u16 grid coordinates, i8 normals)."""
        off = int(self.mesh_infos[mesh_info][field])
        if self._fp32(mesh_info):
            return self.geometry[off: off + 4 * comps * count].view(np.float32).reshape(-1, comps)
        n = np.dtype(qtype).itemsize * comps * count
        return self.geometry[off: off + n].view(qtype).reshape(-1, comps).astype(np.float32)

    def positions(self, mesh_info: int, count: int) -> np.ndarray:
        return self._attr(mesh_info, "positions", count, 3, np.uint16)

    def normals(self, mesh_info: int, count: int) -> np.ndarray:
        return self._attr(mesh_info, "normals", count, 3, np.int8)

    def uvs(self, mesh_info: int, count: int) -> np.ndarray:
        return self._attr(mesh_info, "uvs", count, 2, np.uint16)

    def remap(self, mesh_info: int, count: int) -> np.ndarray:
        off = int(self.mesh_infos[mesh_info]["indices"])
        if int(self.mesh_infos[mesh_info]["flags"]) & abi.MESH_INFO_FLAGS_32_BIT_INDICES:
            return self.geometry[off: off + 4 * count].view(np.uint32)
        return self.geometry[off: off + 2 * count].view(np.uint16).astype(np.uint32)

    def micro(self, mesh_info: int, count: int) -> np.ndarray:
        off = int(self.mesh_infos[mesh_info]["micro_indices"])
        return self.geometry[off: off + count]

    def instance_triangles(self) -> np.ndarray:
        """Triangles per instance."""
        tri = np.zeros(len(self.mesh_infos), np.int64)
        for i in range(len(self.mesh_infos)):
            ml = self.meshlets(i)
            tri[i] = int(ml["triangle_count"].astype(np.int64).sum())
        return tri[self.instances["mesh_info_address"].astype(np.int64)]
