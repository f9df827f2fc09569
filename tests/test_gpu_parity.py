"""GPU parity: the HIP frame against the oracle, pass by pass, through the C ABI.

Bars: visible instance / meshlet sets equal as sorted sets, prefix sectors
consistent, visibility-buffer keys, shadow depth bits, depth bounds and
cascade matrices bit-exact, FP32 radiance within 1e-3 relative, BGRA8 within
one step.
"""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, make_uniforms, preset_camera
from oracle.compare import compare_frames, oracle_frame, snapshot

pytestmark = pytest.mark.gpu


def _gpu_frame(scene, cam, cfg):
    import torch

    from lighthugger_amd.lib import Renderer

    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    return snapshot(r.view)


def _run(scene, cam, cfg):
    print(f"case: {cfg.width}x{cfg.height} shadow {cfg.shadow_size}, {scene.num_instances} instances", flush=True)
    gpu = _gpu_frame(scene, cam, cfg)
    orc = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))
    rep = compare_frames(gpu, orc, scene)
    print(rep, flush=True)
    return rep


def test_small_frame_parity():
    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    rep = _run(scene, cam, ViewConfig(cam.width, cam.height, 1024))
    assert rep["covered_pixels"] > 10000


def test_small_frame_twice_identical():
    """Two frames in a row through the same buffers give identical keys."""
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    r = Renderer(scene, ViewConfig(cam.width, cam.height, 1024))
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    a = snapshot(r.view)
    r.frame()
    torch.cuda.synchronize()
    b = snapshot(r.view)
    assert np.array_equal(a["visbuffer"], b["visbuffer"])
    assert np.array_equal(a["shadow"], b["shadow"])
    assert np.array_equal(a["radiance"], b["radiance"])
