"""ctypes binding of liblh.so (include/lh/lh.h) and a small renderer facade.

The product path: every call goes through the C ABI into HIP kernels.  There
is no fallback; a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os
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
                ("clip_blocks", C.c_uint32), ("cull_blocks", C.c_uint32), ("medium_extent_px", C.c_uint32),
                ("tile_item_entries", C.c_uint32), ("tile_blocks", C.c_uint32), ("tile_raster", C.c_uint32),
                ("shadow_small_extent_px", C.c_uint32), ("shadow_medium_extent_px", C.c_uint32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class KernelTime(C.Structure):
    """lh_kernel_time (include/lh/lh.h)."""

    _fields_ = [("name", C.c_char * 40), ("ms", C.c_float)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # LH_LIBRARY selects another build of the same library (e.g. a
        # different occupancy target); default: the in-tree build.
        path = Path(os.environ.get("LH_LIBRARY", _PKG / "liblh.so"))
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
        L.lh_scratch_bytes.argtypes = [C.c_void_p]
        L.lh_set_tuning.restype = C.c_int32
        L.lh_set_tuning.argtypes = [C.POINTER(Tuning)]
        L.lh_get_tuning.restype = C.c_int32
        L.lh_get_tuning.argtypes = [C.POINTER(Tuning)]
        L.lh_version.restype = C.c_char_p
        L.lh_frame_noalloc.restype = C.c_int32
        L.lh_frame_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.lh_frame_graph_create.restype = C.c_int32
        L.lh_frame_graph_create.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.lh_frame_graph_run.restype = C.c_int32
        L.lh_frame_graph_run.argtypes = [C.c_void_p, C.c_void_p]
        L.lh_frame_graph_destroy.restype = None
        L.lh_frame_graph_destroy.argtypes = [C.c_void_p]
        L.lh_frame_profiled.restype = C.c_int32
        L.lh_frame_profiled.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(KernelTime), C.c_uint32,
                                        C.POINTER(C.c_uint32)]
        L.lh_batch_create.restype = C.c_int32
        L.lh_batch_create.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]
        L.lh_batch_run.restype = C.c_int32
        L.lh_batch_run.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.lh_batch_profiled.restype = C.c_int32
        L.lh_batch_profiled.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(KernelTime), C.c_uint32,
                                        C.POINTER(C.c_uint32)]
        L.lh_batch_destroy.restype = None
        L.lh_batch_destroy.argtypes = [C.c_void_p]
        L.lh_set_tuning_class.restype = C.c_int32
        L.lh_set_tuning_class.argtypes = [C.c_uint32, C.POINTER(Tuning)]
        L.lh_get_tuning_class.restype = C.c_int32
        L.lh_get_tuning_class.argtypes = [C.c_uint32, C.POINTER(Tuning)]
        L.lh_tuning_class_of.restype = C.c_uint32
        L.lh_tuning_class_of.argtypes = [C.c_uint32, C.c_uint32]
        _lib = L
        if not os.environ.get("LH_NO_TUNING_FILE"):
            from .autotune import apply_file  # noqa: PLC0415

            apply_file()
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


def get_tuning_classes() -> list[Tuning]:
    """Every class's tuning (lh_get_tuning_class), e.g. to restore them later."""
    out = []
    for c in range(abi.TUNING_CLASSES):
        t = Tuning()
        check(lib().lh_get_tuning_class(c, C.byref(t)), "lh_get_tuning_class")
        out.append(t)
    return out


def set_tuning_classes(ts: list[Tuning]) -> None:
    for c, t in enumerate(ts):
        check(lib().lh_set_tuning_class(c, C.byref(t)), "lh_set_tuning_class")


class Renderer:
    """A scene resident on the GPU plus one view's working set."""

    def __init__(self, scene: HostScene, cfg: ViewConfig, device="cuda:0"):
        self.mem = DeviceMemory(device)
        self.scene = SceneData(scene, self.mem)
        self.view = ViewData(cfg, self.scene, self.mem)
        self.host_scene = scene
        self._graph = None

    def set_camera(self, cam: abi.Camera) -> None:
        self.view.set_uniforms(make_uniforms(cam, self.host_scene.num_instances), self.scene)

    def call(self, name: str, *extra) -> None:
        fn = getattr(lib(), name)
        check(fn(self.scene.address, self.view.address, *extra, current_stream()), name)

    def frame(self) -> None:
        self.call("lh_frame")

    def frame_graph(self) -> None:
        """lh_frame replayed from a hipGraph (lh_frame_graph), captured on the
        first call.  The camera may change between calls (set_camera writes
        the device uniforms the graph reads)."""
        if self._graph is None:
            h = C.c_void_p()
            check(lib().lh_frame_graph_create(self.scene.address, self.view.address, C.byref(h)),
                  "lh_frame_graph_create")
            self._graph = h
        check(lib().lh_frame_graph_run(self._graph, current_stream()), "lh_frame_graph_run")

    def close_graph(self) -> None:
        if self._graph is not None:
            lib().lh_frame_graph_destroy(self._graph)
            self._graph = None

    def frame_profiled(self) -> list[tuple[str, float]]:
        """One frame with per-kernel device times (ms), in launch order."""
        buf = (KernelTime * 160)()
        n = C.c_uint32(0)
        check(lib().lh_frame_profiled(self.scene.address, self.view.address, current_stream(), buf, 160,
                                      C.byref(n)), "lh_frame_profiled")
        return [(buf[i].name.decode(), float(buf[i].ms)) for i in range(n.value)]

    def status(self) -> np.ndarray:
        out = np.zeros(abi.NUM_COUNTERS, np.uint32)
        lib().lh_view_status(self.view.address, out.ctypes.data, current_stream())
        return out


BATCH_LOOP, BATCH_BATCHED = 0, 1


class BatchRenderer:
    """Many views (cameras, each with its own cascades and buffers) over one
    GPU-resident scene, run through lh_batch_run: per-view loop or batched,
    direct launches or a replayed hipGraph."""

    def __init__(self, scene: HostScene, cameras, *, shadow_sizes=None, device="cuda:0", big_capacity=1 << 18,
                 clip_capacity=1 << 16, draw_capacity=0):
        self.mem = DeviceMemory(device)
        self.host_scene = scene
        self.scene = SceneData(scene, self.mem)
        self.views = []
        for i, cam in enumerate(cameras):
            s = shadow_sizes[i] if shadow_sizes is not None else max(cam.width, cam.height)
            vd = ViewData(ViewConfig(cam.width, cam.height, s, big_capacity=big_capacity,
                                     clip_capacity=clip_capacity, draw_capacity=draw_capacity), self.scene, self.mem)
            vd.set_uniforms(make_uniforms(cam, scene.num_instances), self.scene)
            self.views.append(vd)
        self._table = (abi.LhView * len(self.views))(*[v.c for v in self.views])
        self.handle = C.c_void_p()
        check(lib().lh_batch_create(self.scene.address, self._table, len(self.views), C.byref(self.handle)),
              "lh_batch_create")

    def run(self, mode: int = BATCH_BATCHED, graph: bool = False) -> None:
        check(lib().lh_batch_run(self.handle, mode, 1 if graph else 0, current_stream()), "lh_batch_run")

    def profiled(self, mode: int = BATCH_BATCHED) -> list[tuple[str, float]]:
        buf = (KernelTime * 4096)()
        n = C.c_uint32(0)
        check(lib().lh_batch_profiled(self.handle, mode, current_stream(), buf, 4096, C.byref(n)),
              "lh_batch_profiled")
        return [(buf[i].name.decode(), float(buf[i].ms)) for i in range(n.value)]

    @property
    def pixels(self) -> int:
        return sum(v.cfg.width * v.cfg.height for v in self.views)

    def close(self) -> None:
        if self.handle:
            lib().lh_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
