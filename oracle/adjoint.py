"""FP64 resolve and its dual-number gradient (oracle/oracle64.cpp).  TEST
INFRASTRUCTURE ONLY: the checker of lh_resolve_backward.

Parameters are a dict of float64 arrays:
  base_color (K, 3), roughness (K,), metallic (K,), sun_dir (3,),
  sun_intensity (3,), geometry (n_words,) = the geometry arena as doubles.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as oracle_lib

FROZEN = np.dtype([("kind", "<u4"), ("flip", "<f4"), ("shadow", "<f4"), ("lod_l0", "<u4", 2), ("lod_l1", "<u4", 2),
                   ("lod_f", "<f4", 2)])
assert FROZEN.itemsize == 36


class Params64(C.Structure):
    _fields_ = [("base_color_factor", C.c_void_p), ("roughness_factor", C.c_void_p), ("metallic_factor", C.c_void_p),
                ("sun_dir", C.c_void_p), ("sun_intensity", C.c_void_p), ("geometry", C.c_void_p),
                ("geometry_base", C.c_uint64), ("geometry_bytes", C.c_uint64)]


def scene_params(scene_data, view_data) -> dict:
    """The differentiable parameters as the scene and uniforms hold them."""
    hs = scene_data.scene
    u = view_data.uniforms
    return {
        "base_color": np.asarray(hs.mesh_infos["base_color_factor"], np.float64).reshape(-1, 3),
        "roughness": hs.factors["roughness_factor"].astype(np.float64),
        "metallic": hs.factors["metallic_factor"].astype(np.float64),
        "sun_dir": np.asarray(u["sun_dir"], np.float64).ravel(),
        "sun_intensity": np.asarray(u["sun_intensity"], np.float64).ravel(),
        "geometry": hs.geometry[: hs.geometry.nbytes // 4 * 4].view(np.float32).astype(np.float64),
    }


def _struct(scene_data, p: dict) -> tuple[Params64, list]:
    keep = [np.ascontiguousarray(p[k], np.float64) for k in
            ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity", "geometry")]
    s = Params64(*[a.ctypes.data for a in keep], scene_data.geometry_base, scene_data.geometry_bytes)
    return s, keep


def frozen(scene_data, view_data) -> np.ndarray:
    out = np.zeros(view_data.cfg.width * view_data.cfg.height, FROZEN)
    st = oracle_lib().orc_resolve_frozen(scene_data.address, view_data.address, out.ctypes.data)
    if st:
        raise RuntimeError(f"orc_resolve_frozen failed: {st}")
    return out


def resolve64(scene_data, view_data, p: dict, fz: np.ndarray) -> np.ndarray:
    s, keep = _struct(scene_data, p)
    out = np.zeros((view_data.cfg.height, view_data.cfg.width, 3), np.float64)
    st = oracle_lib().orc64_resolve(scene_data.address, view_data.address, C.byref(s), fz.ctypes.data, out.ctypes.data)
    if st:
        raise RuntimeError(f"orc64_resolve failed: {st}")
    return out


def grad64(scene_data, view_data, p: dict, fz: np.ndarray, grad_radiance: np.ndarray) -> dict:
    s, keep = _struct(scene_data, p)
    g = {k: np.zeros_like(np.asarray(p[k], np.float64)) for k in p}
    gr = np.ascontiguousarray(grad_radiance, np.float32)
    assert gr.shape == (view_data.cfg.height, view_data.cfg.width, 4)
    st = oracle_lib().orc64_resolve_grad(scene_data.address, view_data.address, C.byref(s), fz.ctypes.data,
                                         gr.ctypes.data, g["base_color"].ctypes.data, g["roughness"].ctypes.data,
                                         g["metallic"].ctypes.data, g["sun_dir"].ctypes.data,
                                         g["sun_intensity"].ctypes.data, g["geometry"].ctypes.data)
    if st:
        raise RuntimeError(f"orc64_resolve_grad failed: {st}")
    return g
