"""GPU: lh_frame_graph replays the single-view frame bit-identically.

The graph is captured once and replayed several times, with the program's own
eager work and synchronizations between the replays; every replay must equal
a direct lh_frame of the same camera in every output (keys, shadow depths,
radiance, BGRA8, counters, visible instance and meshlet lists).  A camera
change between replays (new uniforms in the same device buffer) must be seen
by the next replay.  With Hi-Z occlusion culling the reference frame is the
second of its camera, so it starts from the same depth history as the
replays."""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig, draw_keys
from lighthugger_amd.scene import HostScene, make_camera, preset_camera
from oracle.compare import snapshot

pytestmark = pytest.mark.gpu


def _snap(r):
    import torch

    torch.cuda.synchronize()
    return snapshot(r.view)


def _assert_same(a, b, what):
    for k in ("visbuffer", "shadow", "radiance", "ldr", "counters", "misc"):
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"
    for p in range(abi.NUM_PASSES):
        assert np.array_equal(draw_keys(a[f"draws{p}"]), draw_keys(b[f"draws{p}"])), f"{what}: draws{p}"
        assert a[f"prefix{p}"][:2] == b[f"prefix{p}"][:2], f"{what}: prefix{p}"


@pytest.mark.parametrize("tiled", [False, True])
@pytest.mark.parametrize("hiz", [False, True])
def test_frame_graph_replays_equal_lh_frame(hiz, tiled, request):
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_LARGE, tri_scale_pct=25, texture_size=128)
    base = preset_camera(abi.PRESET_LARGE)
    w, h = 640, 360
    cams = [make_camera(base.position.tolist(), base.target.tolist(), w, h),
            make_camera((np.array(base.position.tolist()) + [30.0, 5.0, -20.0]).tolist(), base.target.tolist(), w, h)]
    if tiled:
        request.getfixturevalue("tiled_raster")
    print(f"case: frame graph, {w}x{h}, hiz={hiz}, tiled={tiled}", flush=True)
    r = Renderer(scene, ViewConfig(w, h, 1024, ldr=True, hiz=hiz))
    try:
        for i, cam in enumerate(cams):
            r.set_camera(cam)
            r.frame()
            r.frame()  # with Hi-Z the second frame reads this camera's own depth history, as replays will
            ref = _snap(r)
            assert int((ref["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum()) > 1000
            junk = torch.empty(1 << 22, dtype=torch.uint8, device="cuda")
            for k in range(3):
                junk.fill_(k)  # eager work between replays
                r.frame_graph()
                _assert_same(ref, _snap(r), f"camera {i} replay {k}")
    finally:
        r.close_graph()
