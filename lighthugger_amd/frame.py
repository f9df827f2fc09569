"""Places a scene and per-view working sets in memory and builds the C structs.

Two memory kinds share one code path:
  - DeviceMemory: torch CUDA (HIP) tensors, used for the HIP library;
  - HostMemory: numpy arrays, used for CPU consumers of the same C structs.
All buffers of a view live in one allocation carved at 256-byte boundaries,
matching the reference's per-frame resources (src/main.cpp:442-532) with the
capacities made parameters.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from .scene import HostScene

ALIGN = 256


def _align(n: int) -> int:
    return (n + ALIGN - 1) & ~(ALIGN - 1)


def _scene_flags(mesh_infos: np.ndarray) -> int:
    """LH_SCENE_OPAQUE_ONLY unless some mesh info is alpha-clipped."""
    alpha = (np.asarray(mesh_infos["flags"]).astype(np.uint32) & abi.MESH_INFO_FLAGS_ALPHA_CLIP).any()
    return 0 if alpha else abi.SCENE_OPAQUE_ONLY


class HostMemory:
    kind = "host"

    def alloc(self, nbytes: int):
        buf = np.zeros(max(nbytes, ALIGN) + ALIGN, np.uint8)
        addr = buf.ctypes.data
        off = (-addr) % ALIGN
        return buf, addr + off, off

    def write(self, handle, offset: int, data: np.ndarray) -> None:
        b = np.ascontiguousarray(data).view(np.uint8).ravel()
        handle[0][handle[2] + offset: handle[2] + offset + b.size] = b

    def read(self, handle, offset: int, nbytes: int) -> np.ndarray:
        return handle[0][handle[2] + offset: handle[2] + offset + nbytes].copy()

    def synchronize(self) -> None:
        pass


class DeviceMemory:
    kind = "device"

    def __init__(self, device="cuda:0"):
        import torch  # noqa: PLC0415

        self.torch = torch
        self.device = torch.device(device)

    def alloc(self, nbytes: int):
        t = self.torch.empty(max(nbytes, ALIGN), dtype=self.torch.uint8, device=self.device)
        return t, t.data_ptr(), 0

    def write(self, handle, offset: int, data: np.ndarray) -> None:
        b = np.ascontiguousarray(data).view(np.uint8).ravel()
        if b.size:
            handle[0][offset: offset + b.size].copy_(self.torch.from_numpy(b.copy()))

    def read(self, handle, offset: int, nbytes: int) -> np.ndarray:
        return handle[0][offset: offset + nbytes].cpu().numpy()

    def synchronize(self) -> None:
        self.torch.cuda.synchronize(self.device)


class SceneData:
    """A HostScene resident in `mem` plus its lh_scene struct."""

    def __init__(self, scene: HostScene, mem):
        self.scene, self.mem = scene, mem
        n_mi, n_in, n_tex = len(scene.mesh_infos), scene.num_instances, max(1, len(scene.textures))
        layout = [
            ("geometry", scene.geometry.nbytes), ("mesh_infos", n_mi * 104), ("factors", n_mi * 8),
            ("instances", n_in * 112), ("meshlet_base", n_in * 4), ("meshlet_table", scene.total_meshlets * 8),
            ("texels", scene.texels.nbytes), ("textures", n_tex * 88),
        ]
        self.offsets, total = {}, 0
        for name, nb in layout:
            self.offsets[name] = total
            total += _align(nb)
        self.handle = mem.alloc(total)
        self.base = self.handle[1]
        o = {k: self.base + v for k, v in self.offsets.items()}
        mi, inst, tex = scene.relocate(o["geometry"], o["mesh_infos"], o["texels"])
        self.mesh_infos, self.instances_reloc, self.textures_reloc = mi, inst, tex
        mem.write(self.handle, self.offsets["geometry"], scene.geometry)
        mem.write(self.handle, self.offsets["mesh_infos"], mi)
        mem.write(self.handle, self.offsets["factors"], scene.factors)
        mem.write(self.handle, self.offsets["instances"], inst)
        mem.write(self.handle, self.offsets["meshlet_base"], scene.instance_meshlet_base)
        mem.write(self.handle, self.offsets["meshlet_table"], scene.meshlet_table)
        mem.write(self.handle, self.offsets["texels"], scene.texels)
        if len(tex):
            mem.write(self.handle, self.offsets["textures"], tex)
        self.c = abi.LhScene(o["instances"], o["mesh_infos"], o["factors"], o["meshlet_base"], o["meshlet_table"],
                             o["textures"], n_in, n_mi, scene.total_meshlets, len(scene.textures),
                             _scene_flags(scene.mesh_infos))

    def set_factors(self, factors: np.ndarray) -> None:
        self.mem.write(self.handle, self.offsets["factors"], factors.astype(abi.MATERIAL_FACTORS))

    def set_mesh_infos(self, mesh_infos: np.ndarray) -> None:
        self.mem.write(self.handle, self.offsets["mesh_infos"], mesh_infos)
        self.c.flags = _scene_flags(mesh_infos)

    def write_geometry(self, offset: int, data: np.ndarray) -> None:
        self.mem.write(self.handle, self.offsets["geometry"] + offset, data)

    @property
    def geometry_base(self) -> int:
        """Address of the geometry arena (vertex positions, normals, uvs, meshlets)."""
        return self.base + self.offsets["geometry"]

    @property
    def geometry_bytes(self) -> int:
        return int(self.scene.geometry.nbytes)

    def geometry_words(self):
        """The geometry arena as 32-bit float words, sharing the scene's memory
        (a torch tensor on the device, a numpy array on the host)."""
        o, n = self.offsets["geometry"], self.geometry_bytes // 4 * 4
        t = self.handle[0][self.handle[2] + o: self.handle[2] + o + n]
        return t.view(np.float32) if self.mem.kind == "host" else t.view(self.mem.torch.float32)

    @property
    def address(self):
        return C.byref(self.c)


@dataclass
class ViewConfig:
    width: int
    height: int
    shadow_size: int = 1024
    instance_capacity: int = 0  # 0 = number of instances
    draw_capacity: int = 0      # 0 = total meshlets of the scene (cannot overflow)
    big_capacity: int = 1 << 20
    clip_capacity: int = 1 << 18
    ldr: bool = True
    statistics: bool = False  # LH_VIEW_STATISTICS: count covered samples and depth atomics
    hiz: bool = False  # LH_VIEW_HIZ: two-phase occlusion culling of the main view
    bin_capacity: int = -1  # tiled main-view raster: draws binned per launch; -1 = draw capacity, 0 = off


class ViewData:
    """One view's working set (lh_view) in `mem`."""

    def __init__(self, cfg: ViewConfig, scene: SceneData, mem):
        self.cfg, self.mem = cfg, mem
        icap = cfg.instance_capacity or max(1, scene.scene.num_instances)
        dcap = cfg.draw_capacity or max(1, scene.scene.total_meshlets)
        bcap = dcap if cfg.bin_capacity < 0 else min(cfg.bin_capacity, dcap)
        W, H, S = cfg.width, cfg.height, cfg.shadow_size
        self.icap, self.dcap, self.bcap = icap, dcap, bcap
        layout = [
            ("uniforms", 560), ("misc", 792), ("prefix", abi.NUM_PASSES * (1 + icap) * 8),
            ("draws", abi.NUM_PASSES * dcap * 8), ("counters", abi.NUM_COUNTERS * 4), ("visbuffer", W * H * 8),
            ("shadow", abi.NUM_CASCADES * S * S * 4), ("radiance", W * H * 16), ("ldr", W * H * 4 if cfg.ldr else 0),
            ("scratch", abi.scratch_bytes(W, H, dcap, cfg.big_capacity, cfg.clip_capacity, bcap)),
            ("history", (scene.scene.total_meshlets + 31) // 32 * 4),
        ]
        self.offsets, self.sizes, total = {}, {}, 0
        for name, nb in layout:
            self.offsets[name], self.sizes[name] = total, nb
            total += _align(nb)
        self.handle = mem.alloc(total)
        b = self.handle[1]
        a = {k: b + v for k, v in self.offsets.items()}
        self.c = abi.LhView(
            a["uniforms"], a["misc"], a["prefix"], a["draws"], a["counters"], a["visbuffer"], a["shadow"],
            a["radiance"], a["ldr"] if cfg.ldr else 0, a["scratch"], self.sizes["scratch"], icap, dcap, W, H, S,
            cfg.big_capacity, cfg.clip_capacity,
            (abi.VIEW_STATISTICS if cfg.statistics else 0) | (abi.VIEW_HIZ if cfg.hiz else 0), a["history"], 0,
            bcap, 0, 0,
        )
        mem.write(self.handle, self.offsets["history"], np.zeros(self.sizes["history"], np.uint8))
        self.uniforms = None

    @property
    def address(self):
        return C.byref(self.c)

    def set_uniforms(self, uniforms: np.ndarray, scene: SceneData) -> None:
        u = uniforms.copy()
        off = {k: self.handle[1] + v for k, v in self.offsets.items()}
        u["meshlet_references"] = scene.c.meshlet_table
        u["instances"] = scene.c.instances
        u["draw_calls"] = off["draws"]
        u["misc_storage"] = off["misc"]
        u["num_meshlets_prefix_sum"] = off["prefix"]
        u["dispatches"] = 0
        self.uniforms = u
        self.mem.write(self.handle, self.offsets["uniforms"], u)

    # ---- readers ----
    def _read(self, name, dtype, count=None):
        nb = self.sizes[name] if count is None else np.dtype(dtype).itemsize * count
        return self.mem.read(self.handle, self.offsets[name], nb).view(dtype)

    def counters(self) -> np.ndarray:
        return self._read("counters", np.uint32)

    def misc(self) -> np.ndarray:
        return self._read("misc", abi.MISC_STORAGE)[0]

    def write_misc(self, misc: np.ndarray) -> None:
        self.mem.write(self.handle, self.offsets["misc"], np.asarray(misc, abi.MISC_STORAGE).reshape(1))

    def visbuffer(self) -> np.ndarray:
        return self._read("visbuffer", np.uint64).reshape(self.cfg.height, self.cfg.width)

    def write_visbuffer(self, vis: np.ndarray) -> None:
        self.mem.write(self.handle, self.offsets["visbuffer"], vis.astype(np.uint64))

    def shadow_maps(self) -> np.ndarray:
        S = self.cfg.shadow_size
        return self._read("shadow", np.uint32).reshape(abi.NUM_CASCADES, S, S)

    def write_shadow_maps(self, maps: np.ndarray) -> None:
        self.mem.write(self.handle, self.offsets["shadow"], maps.astype(np.uint32))

    def radiance(self) -> np.ndarray:
        return self._read("radiance", np.float32).reshape(self.cfg.height, self.cfg.width, 4)

    def ldr(self) -> np.ndarray:
        return self._read("ldr", np.uint32).reshape(self.cfg.height, self.cfg.width)

    def prefix(self, p: int):
        """(count, total, PrefixSumValue[count clamped])"""
        words = self._read("prefix", np.uint64).reshape(abi.NUM_PASSES, 1 + self.icap)[p]
        count, total = int(words[0] >> 32), int(words[0] & 0xFFFFFFFF)
        vals = words[1:].view(abi.PREFIX_SUM_VALUE)[: min(count, self.icap)]
        return count, total, vals.copy()

    def draws(self, p: int) -> np.ndarray:
        n = min(int(self.counters()[abi.CNT_DRAWS + p]), self.dcap)
        d = self._read("draws", abi.MESHLET_REFERENCE).reshape(abi.NUM_PASSES, self.dcap)[p]
        return d[:n].copy()


def draw_keys(draws: np.ndarray) -> np.ndarray:
    """Sorted (instance << 16 | meshlet) keys of a draw list, for set comparison."""
    k = (draws["instance_index"].astype(np.uint64) << np.uint64(16)) | draws["meshlet_index"].astype(np.uint64)
    return np.sort(k)
