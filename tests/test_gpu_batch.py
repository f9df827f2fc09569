"""GPU: batched multi-view frames (lh_batch_run) against the per-view loop and
the oracle.  The batch must be bit-identical to the loop for every view
(keys, shadow depths, radiance, BGRA8) and equal as sets for the visible
instance / meshlet lists; graph replays are checked replay by replay with
eager work in between."""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import draw_keys
from lighthugger_amd.scene import HostScene, batch_cameras, make_uniforms
from oracle.compare import compare_frames, oracle_frame, snapshot

pytestmark = pytest.mark.gpu

EXACT = ("visbuffer", "shadow", "radiance", "ldr")


def _snap_all(br):
    import torch

    torch.cuda.synchronize()
    return [snapshot(v) for v in br.views]


def _assert_same(a, b, what):
    for i, (x, y) in enumerate(zip(a, b)):
        for k in EXACT:
            assert np.array_equal(x[k], y[k]), f"{what}: view {i} {k} differs"
        for p in range(abi.NUM_PASSES):
            assert np.array_equal(draw_keys(x[f"draws{p}"]), draw_keys(y[f"draws{p}"])), f"{what}: view {i} draws{p}"
            assert x[f"prefix{p}"][:2] == y[f"prefix{p}"][:2], f"{what}: view {i} prefix{p}"


@pytest.fixture(scope="module")
def batch():
    from lighthugger_amd.lib import BatchRenderer

    scene = HostScene.preset(abi.PRESET_LARGE, texture_size=128)
    cams = batch_cameras(4, 6, 64, 320)
    print("batch sides", [c.width for c in cams], flush=True)
    br = BatchRenderer(scene, cams)
    yield scene, cams, br
    br.close()


def test_batched_equals_per_view_loop(batch):
    from lighthugger_amd.lib import BATCH_BATCHED, BATCH_LOOP

    scene, cams, br = batch
    br.run(BATCH_LOOP)
    loop = _snap_all(br)
    br.run(BATCH_BATCHED)
    bat = _snap_all(br)
    _assert_same(loop, bat, "batched vs loop")
    assert sum(int((s["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum()) for s in bat) > 1000


def test_batched_graph_replays(batch):
    import torch

    from lighthugger_amd.lib import BATCH_BATCHED, BATCH_LOOP

    scene, cams, br = batch
    br.run(BATCH_LOOP)
    ref = _snap_all(br)
    junk = torch.empty(1 << 24, dtype=torch.uint8, device="cuda")
    for k in range(3):
        junk.fill_(k)  # eager work between replays
        br.run(BATCH_BATCHED, graph=True)
        _assert_same(ref, _snap_all(br), f"batched graph replay {k}")
    for k in range(2):
        br.run(BATCH_LOOP, graph=True)
        _assert_same(ref, _snap_all(br), f"loop graph replay {k}")


def test_batched_views_match_oracle(batch):
    from lighthugger_amd.frame import ViewConfig
    from lighthugger_amd.lib import BATCH_BATCHED

    scene, cams, br = batch
    br.run(BATCH_BATCHED)
    snaps = _snap_all(br)
    for i in (0, int(np.argmax([c.width for c in cams]))):
        cam, vd = cams[i], br.views[i]
        orc = oracle_frame(scene, vd.cfg, make_uniforms(cam, scene.num_instances))
        print(i, compare_frames(snaps[i], orc, scene), flush=True)


def test_batch_profiled_names(batch):
    from lighthugger_amd.lib import BATCH_BATCHED

    _, _, br = batch
    names = [n for n, _ in br.profiled(BATCH_BATCHED)]
    assert names.count("raster_meshlets_view") == 1 and names.count("raster_meshlets_shadow") == 1
    assert names[-1] == "resolve"


def test_alpha_clip_batch_equals_loop_and_oracle():
    """Alpha-clip geometry in a batch: the batched launch equals the per-view
    loop bit for bit and both equal the oracle."""
    from lighthugger_amd.lib import BATCH_BATCHED, BATCH_LOOP, BatchRenderer
    from lighthugger_amd.scene import make_camera

    from .fixtures import alpha_scene

    scene, _, _ = alpha_scene()
    cams = [make_camera((0.2, 0.4, 4.0), (0, 0, 0), 96, 64), make_camera((-2.0, 1.0, 3.0), (0, 0, 0), 64, 64),
            make_camera((1.5, 0.2, -3.0), (0, 0, 0), 128, 96)]
    br = BatchRenderer(scene, cams, shadow_sizes=[128, 64, 128], big_capacity=1 << 16, clip_capacity=1 << 12)
    try:
        br.run(BATCH_LOOP)
        loop = _snap_all(br)
        br.run(BATCH_BATCHED)
        bat = _snap_all(br)
        _assert_same(loop, bat, "alpha batch vs loop")
        for i, cam in enumerate(cams):
            orc = oracle_frame(scene, br.views[i].cfg, make_uniforms(cam, scene.num_instances))
            print(i, compare_frames(bat[i], orc, scene), flush=True)
    finally:
        br.close()
