"""ctypes binding of liblh.so (include/lh/lh.h) and a small renderer facade.

The product path: every call goes through the C ABI into HIP kernels.  There
is no fallback; a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import abi
from .frame import DeviceMemory, SceneData, ViewConfig, ViewData
from .scene import HostScene, make_uniforms

_PKG = Path(__file__).resolve().parent
_lib = None

# (name, extra argument types after (scene, view)); None = (view, stream) only
_PASSES = {
    "lh_view_reset": [], "lh_cull_instances": [], "lh_cull_meshlets": [C.c_uint32], "lh_raster_view": [],
    "lh_depth_bounds": [], "lh_fit_cascades": [], "lh_cull_instances_shadow": [], "lh_raster_shadow": [C.c_uint32],
    "lh_resolve": [], "lh_frame": [],
}


class Tuning(C.Structure):
    """lh_tuning (include/lh/lh.h)."""

    _fields_ = [("small_extent_px", C.c_uint32), ("meshlet_blocks", C.c_uint32), ("big_blocks", C.c_uint32),
                ("clip_blocks", C.c_uint32), ("cull_blocks", C.c_uint32)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = _PKG / "liblh.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: the HIP library must be built (__graft_entry__.build())")
        L = C.CDLL(str(path))
        for name, extra in _PASSES.items():
            fn = getattr(L, name)
            fn.restype = C.c_int32
            fn.argtypes = [C.c_void_p, C.c_void_p, *extra, C.c_void_p]
        L.lh_tonemap.restype = C.c_int32
        L.lh_tonemap.argtypes = [C.c_void_p, C.c_void_p]
        L.lh_view_status.restype = C.c_int32
        L.lh_view_status.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.lh_scratch_bytes.restype = C.c_uint64
        L.lh_scratch_bytes.argtypes = [C.c_uint32, C.c_uint32]
        L.lh_set_tuning.restype = C.c_int32
        L.lh_set_tuning.argtypes = [C.POINTER(Tuning)]
        L.lh_get_tuning.restype = C.c_int32
        L.lh_get_tuning.argtypes = [C.POINTER(Tuning)]
        L.lh_version.restype = C.c_char_p
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != abi.LH_OK:
        raise RuntimeError(f"{what} failed with lh_status {status}")


def current_stream() -> int:
    import torch  # noqa: PLC0415

    return torch.cuda.current_stream().cuda_stream


def get_tuning() -> Tuning:
    t = Tuning()
    check(lib().lh_get_tuning(C.byref(t)), "lh_get_tuning")
    return t


def set_tuning(t: Tuning) -> None:
    check(lib().lh_set_tuning(C.byref(t)), "lh_set_tuning")


class Renderer:
    """A scene resident on the GPU plus one view's working set."""

    def __init__(self, scene: HostScene, cfg: ViewConfig, device="cuda:0"):
        self.mem = DeviceMemory(device)
        self.scene = SceneData(scene, self.mem)
        self.view = ViewData(cfg, self.scene, self.mem)
        self.host_scene = scene

    def set_camera(self, cam: abi.Camera) -> None:
        self.view.set_uniforms(make_uniforms(cam, self.host_scene.num_instances), self.scene)

    def call(self, name: str, *extra) -> None:
        fn = getattr(lib(), name)
        check(fn(self.scene.address, self.view.address, *extra, current_stream()), name)

    def frame(self) -> None:
        self.call("lh_frame")

    def status(self) -> np.ndarray:
        out = np.zeros(abi.NUM_COUNTERS, np.uint32)
        lib().lh_view_status(self.view.address, out.ctypes.data, current_stream())
        return out
