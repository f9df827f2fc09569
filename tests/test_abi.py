"""The C-ABI boundary: struct layouts and exported symbols (no GPU needed).

Layouts: a probe compiled from include/lh/*.h prints sizeof/offsetof of every
shared struct; the numpy / ctypes mirrors must agree (and the header's static
assertions pin the reference's offsets, src/shared_cpu_gpu.h:18-157).
Exports: liblh.so and liblh_scene.so load without a GPU and export every
function their headers declare.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.lib import Tuning

ROOT = Path(__file__).resolve().parents[1]

PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "lh/lh.h"
#include "lh/lh_scene.h"
#define S(T) printf(#T " size %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(MeshInfo) S(Instance) S(Meshlet) S(MeshletReference) S(Uniforms) S(MiscStorage) S(PrefixSumValue)
  S(LhTexture) S(LhMaterialFactors) S(lh_scene) S(lh_view) S(lh_camera) S(lh_scene_arrays) S(lh_tuning)
  O(MeshInfo, base_color_factor) O(Instance, normal_transform) O(Meshlet, index_count)
  O(Uniforms, debug_shadowmaps) O(MiscStorage, max_depth) O(LhTexture, level_offset)
  O(lh_scene, num_instances) O(lh_scene, num_textures) O(lh_scene, flags) O(lh_view, scratch_bytes) O(lh_view, instance_capacity)
  O(lh_view, flags) O(lh_view, history) O(lh_view, bin_capacity) O(lh_camera, shadow_cam_distance) O(lh_scene_arrays, unique_triangles)
  return 0;
}
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    src, exe = d / "probe.c", d / "probe"
    src.write_text(PROBE)
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.rsplit(" ", 1) for line in out.splitlines())}


def test_struct_sizes(probe):
    assert probe["MeshInfo size"] == abi.MESH_INFO.itemsize == 104
    assert probe["Instance size"] == abi.INSTANCE.itemsize == 112
    assert probe["Meshlet size"] == abi.MESHLET.itemsize == 56
    assert probe["MeshletReference size"] == abi.MESHLET_REFERENCE.itemsize == 8
    assert probe["Uniforms size"] == abi.UNIFORMS.itemsize == 560
    assert probe["MiscStorage size"] == abi.MISC_STORAGE.itemsize == 792
    assert probe["PrefixSumValue size"] == abi.PREFIX_SUM_VALUE.itemsize == 8
    assert probe["LhTexture size"] == abi.TEXTURE.itemsize
    assert probe["LhMaterialFactors size"] == abi.MATERIAL_FACTORS.itemsize
    assert probe["lh_scene size"] == C.sizeof(abi.LhScene)
    assert probe["lh_view size"] == C.sizeof(abi.LhView)
    assert probe["lh_camera size"] == C.sizeof(abi.Camera)
    assert probe["lh_scene_arrays size"] == C.sizeof(abi.SceneArrays)
    assert probe["lh_tuning size"] == C.sizeof(Tuning)


def test_field_offsets(probe):
    assert probe["MeshInfo.base_color_factor"] == abi.MESH_INFO.fields["base_color_factor"][1]
    assert probe["Instance.normal_transform"] == abi.INSTANCE.fields["normal_transform"][1]
    assert probe["Meshlet.index_count"] == abi.MESHLET.fields["index_count"][1]
    assert probe["Uniforms.debug_shadowmaps"] == abi.UNIFORMS.fields["debug_shadowmaps"][1]
    assert probe["MiscStorage.max_depth"] == abi.MISC_STORAGE.fields["max_depth"][1]
    assert probe["LhTexture.level_offset"] == abi.TEXTURE.fields["level_offset"][1]
    assert probe["lh_scene.num_instances"] == abi.LhScene.num_instances.offset
    assert probe["lh_scene.num_textures"] == abi.LhScene.num_textures.offset
    assert probe["lh_scene.flags"] == abi.LhScene.flags.offset
    assert probe["lh_view.scratch_bytes"] == abi.LhView.scratch_bytes.offset
    assert probe["lh_view.instance_capacity"] == abi.LhView.instance_capacity.offset
    assert probe["lh_view.flags"] == abi.LhView.flags.offset
    assert probe["lh_view.history"] == abi.LhView.history.offset
    assert probe["lh_view.bin_capacity"] == abi.LhView.bin_capacity.offset
    assert probe["lh_camera.shadow_cam_distance"] == abi.Camera.shadow_cam_distance.offset
    assert probe["lh_scene_arrays.unique_triangles"] == abi.SceneArrays.unique_triangles.offset


def _declared(header: Path):
    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(lh_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("lib_name,header", [("liblh.so", "lh.h"), ("liblh_scene.so", "lh_scene.h")])
def test_library_exports_every_declared_function(lib_name, header):
    names = _declared(ROOT / "include" / "lh" / header)
    assert len(names) >= 8
    lib = C.CDLL(str(ROOT / "lighthugger_amd" / lib_name))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_entry_points_without_gpu():
    from lighthugger_amd.lib import get_tuning, get_tuning_classes, lib, set_tuning_classes

    L = lib()
    assert L.lh_version().startswith(b"lighthugger_amd")
    v = abi.LhView()
    v.big_capacity, v.clip_capacity = 10, 100
    assert L.lh_scratch_bytes(C.byref(v)) == abi.scratch_bytes(0, 0, 0, 10, 100, 0) == 1024 + 512
    v.width, v.height, v.draw_capacity, v.bin_capacity = 1920, 1080, 5000, 3000
    assert L.lh_scratch_bytes(C.byref(v)) == abi.scratch_bytes(1920, 1080, 5000, 10, 100, 3000)
    assert abi.scratch_bytes(1920, 1080, 5000, 10, 100, 3000) > 3000 * abi.BIN_CACHE_BYTES
    t = get_tuning()
    assert 1 <= t.small_extent_px <= 32 and t.tile_raster in (0, 1)
    bad = Tuning(0, 1, 1, 1, 1, 8)
    assert L.lh_set_tuning(C.byref(bad)) == abi.LH_ERR_INVALID
    bad = Tuning(8, 1, 1, 1, 1, 4)  # medium below small
    assert L.lh_set_tuning(C.byref(bad)) == abi.LH_ERR_INVALID
    bad = Tuning(8, 1, 1, 1, 1, 8, 0, 0, 2)  # tile_raster is 0 or 1
    assert L.lh_set_tuning(C.byref(bad)) == abi.LH_ERR_INVALID
    bad = Tuning(8, 1, 1, 1, 1, 8, 0, 0, 1, 33)  # the cascades' small extent is at most 32
    assert L.lh_set_tuning(C.byref(bad)) == abi.LH_ERR_INVALID
    bad = Tuning(8, 1, 1, 1, 1, 8, 0, 0, 1, 6, 4)  # the cascades' medium extent below their small one
    assert L.lh_set_tuning(C.byref(bad)) == abi.LH_ERR_INVALID
    bad = Tuning(8, 1, 1, 1, 1, 16, 0, 0, 1, 0, 4)  # ... or below the main view's small one it inherits
    assert L.lh_set_tuning(C.byref(bad)) == abi.LH_ERR_INVALID
    classes = get_tuning_classes()
    set_tuning_classes(classes)
    assert [c.as_dict() for c in get_tuning_classes()] == [c.as_dict() for c in classes]
    # argument validation happens before any HIP call
    view = abi.LhView()
    scene = abi.LhScene()
    assert L.lh_frame(C.byref(scene), C.byref(view), None) == abi.LH_ERR_INVALID
    assert L.lh_cull_meshlets(C.byref(scene), C.byref(view), 7, None) == abi.LH_ERR_INVALID


def test_oracle_exports():
    import oracle

    names = re.findall(r"\b(orc\w*)\s*\(", (ROOT / "oracle" / "oracle.h").read_text())
    L = oracle.lib()
    assert all(hasattr(L, n) for n in names) and len(names) >= 10
    assert np.float32(L.orc_srgb_oetf(0.0)) == 0.0
