"""ctypes binding of the scalar oracle (oracle/liblh_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path (liblh.so)
never loads it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_HERE = Path(__file__).resolve().parent
_lib = None

_VIEW_FNS = ["orc_view_reset", "orc_cull_instances", "orc_raster_view", "orc_depth_bounds", "orc_fit_cascades",
             "orc_cull_instances_shadow", "orc_resolve", "orc_frame"]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = _HERE / "liblh_oracle.so"
        if not path.exists():
            from . import build  # noqa: PLC0415

            build.build()
        _lib = bind(C.CDLL(str(path)))
    return _lib


def bind(L: C.CDLL) -> C.CDLL:
    """Declares the argument and result types of the oracle's exports on L."""
    for name in _VIEW_FNS:
        getattr(L, name).restype = C.c_int32
        getattr(L, name).argtypes = [C.c_void_p, C.c_void_p]
    L.orc_cull_meshlets.restype = C.c_int32
    L.orc_cull_meshlets.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    L.orc_raster_shadow.restype = C.c_int32
    L.orc_raster_shadow.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    L.orc_tonemap.restype = C.c_int32
    L.orc_tonemap.argtypes = [C.c_void_p]
    L.orc_raster_clip_triangle.restype = C.c_uint64
    L.orc_raster_clip_triangle.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p]
    L.orc_cull_sphere.restype = C.c_int
    L.orc_cull_sphere.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_cull_sphere_shadow.restype = C.c_int
    L.orc_cull_sphere_shadow.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
    L.orc_cull_cone.restype = C.c_int
    L.orc_cull_cone.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    L.orc_calc_full_bary.restype = None
    L.orc_calc_full_bary.argtypes = [C.c_void_p] * 6
    L.orc_brdf.restype = None
    L.orc_brdf.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_int,
                           C.c_void_p]
    L.orc_srgb_oetf.restype = C.c_float
    L.orc_srgb_oetf.argtypes = [C.c_float]
    L.orc_frame_threads.restype = C.c_int32
    L.orc_frame_threads.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    L.orc_bc1_texel.restype = C.c_uint32
    L.orc_bc1_texel.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
    L.orc_hiz_bound_violations.restype = C.c_int64
    L.orc_hiz_bound_violations.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_resolve_frozen.restype = C.c_int32
    L.orc_resolve_frozen.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc64_resolve.restype = C.c_int32
    L.orc64_resolve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc64_resolve_grad.restype = C.c_int32
    L.orc64_resolve_grad.argtypes = [C.c_void_p] * 5 + [C.c_void_p] * 6
    L.orc_bc7_block.restype = None
    L.orc_bc7_block.argtypes = [C.c_void_p, C.c_void_p]
    L.orc_srgb_to_linear.restype = C.c_float
    L.orc_srgb_to_linear.argtypes = [C.c_uint32]
    L.orc_shadow_tap.restype = C.c_float
    L.orc_shadow_tap.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_float]
    L.orc_shadow_factor.restype = C.c_float
    L.orc_shadow_factor.argtypes = [C.c_void_p] * 4
    L.orc_texture_lod.restype = None
    L.orc_texture_lod.argtypes = [C.c_void_p] * 5
    L.orc_bilinear.restype = None
    L.orc_bilinear.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    L.orc_sample_texture.restype = None
    L.orc_sample_texture.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    L.orc_normal_map.restype = None
    L.orc_normal_map.argtypes = [C.c_void_p] * 5
    L.orc_sky_radiance.restype = None
    L.orc_sky_radiance.argtypes = [C.c_void_p] * 3
    L.orc_temperature.restype = None
    L.orc_temperature.argtypes = [C.c_float, C.c_void_p]
    return L


def check(status: int, what: str = "oracle") -> None:
    if status not in (0, 2):
        raise RuntimeError(f"{what} failed with status {status}")


def frame(scene, view) -> int:
    """Runs the whole frame on the CPU: scene/view are SceneData/ViewData in HostMemory."""
    return lib().orc_frame(scene.address, view.address)
