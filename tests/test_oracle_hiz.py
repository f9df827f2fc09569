"""The occlusion test's screen bound (two-phase Hi-Z of the main view) is
conservative: for every surviving main-view meshlet of the large scene, the
pixel box of each of its triangles lies inside the meshlet's bound rectangle
and no vertex is nearer than the bound's depth.  Brute force over every
triangle, through the oracle's own projection and snapping (the rule the
device code restates in lighthugger_amd/csrc/hiz.h)."""
import ctypes as C

import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, make_camera, make_uniforms, preset_camera
from oracle import lib as oracle_lib
from oracle.compare import oracle_view


@pytest.mark.parametrize("size", [(480, 270), (97, 61)])
def test_hiz_bound_contains_every_triangle(size):
    W, H = size
    scene = HostScene.preset(abi.PRESET_LARGE, texture_size=16)
    base = preset_camera(abi.PRESET_LARGE)
    cam = make_camera(base.position.tolist(), base.target.tolist(), W, H)
    sd, vd = oracle_view(scene, ViewConfig(W, H, 64), make_uniforms(cam, scene.num_instances))
    L = oracle_lib()
    L.orc_view_reset(sd.address, vd.address)
    L.orc_cull_instances(sd.address, vd.address)
    L.orc_cull_meshlets(sd.address, vd.address, 0)
    checked = C.c_uint64(0)
    bad = L.orc_hiz_bound_violations(sd.address, vd.address, C.byref(checked))
    assert checked.value > 100000
    assert bad == 0
