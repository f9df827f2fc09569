"""GPU parity at the full sizes of the three heavier workloads.

* Stress frame: 20,000 instances of the 50 seeded meshes (about 464M scene
  triangles), 3840x2160 with 4 cascades at 4096^2, against the threaded
  oracle: visible sets, keys and shadow depth bits bit-exact, radiance within
  1e-3 relative.
* Batched views: 64 seeded square cameras of side 256..2048 (log-uniform,
  multiples of 8), each with its own 4 cascades of the same side, i.e. 320
  views in one batch: batched launch and replayed graph bit-identical to the
  per-view loop (compared on the device), the smallest and the largest camera
  against the oracle.
* Differentiable resolve: the large scene at 18% of its triangles (about 1M
  visible-candidate triangles), 1024x1024 with 4 x 1024^2 cascades, visibility
  fixed: every gradient group within 1e-4 relative L2 of the FP64 adjoint.
"""
import os

import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, batch_cameras, make_camera, make_uniforms, preset_camera
from oracle.compare import compare_frames, oracle_frame, snapshot

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
REL_L2 = 1e-4


def test_stress_frame_parity_full_size():
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_STRESS)
    cam = preset_camera(abi.PRESET_STRESS)
    cfg = ViewConfig(cam.width, cam.height, 4096, big_capacity=1 << 21, clip_capacity=1 << 18)
    print(f"case: stress {cam.width}x{cam.height} + 4x4096^2, {scene.num_instances} instances, "
          f"{scene.total_triangles} triangles", flush=True)
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    gpu = snapshot(r.view)
    del r
    orc = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances), threads=THREADS)
    rep = compare_frames(gpu, orc, scene)
    print(rep, flush=True)
    assert rep["covered_pixels"] > 1_000_000
    assert rep["pass0"]["meshlets"] > 100_000


@pytest.fixture(scope="module")
def batch64():
    from lighthugger_amd.lib import BatchRenderer

    scene = HostScene.preset(abi.PRESET_LARGE)
    cams = batch_cameras(4, 64, 256, 2048)
    print("case: 64 cameras, sides", sorted(int(c.width) for c in cams), flush=True)
    br = BatchRenderer(scene, cams, big_capacity=1 << 20)
    yield scene, cams, br
    br.close()


def _outputs(br):
    out = []
    for v in br.views:
        t = v.handle[0]
        for name in ("visbuffer", "shadow", "radiance", "ldr"):
            o, n = v.offsets[name], v.sizes[name]
            out.append(t[o: o + n].clone())
    return out


def test_batch_64_cameras_equals_loop_full_size(batch64):
    import torch

    from lighthugger_amd.lib import BATCH_BATCHED, BATCH_LOOP

    scene, cams, br = batch64
    br.run(BATCH_LOOP)
    torch.cuda.synchronize()
    ref = _outputs(br)
    assert not any(int(v.counters()[abi.CNT_OVERFLOW]) for v in br.views)
    for name, graph, reps in (("batched", False, 1), ("batched graph", True, 3)):
        for k in range(reps):
            # scribble over the outputs so a replay that writes nothing fails
            for v in br.views:
                o, n = v.offsets["visbuffer"], v.sizes["visbuffer"]
                v.handle[0][o: o + n].fill_(0x5A)
            br.run(BATCH_BATCHED, graph)
            torch.cuda.synchronize()
            got = _outputs(br)
            bad = [i for i, (a, b) in enumerate(zip(ref, got)) if not torch.equal(a, b)]
            assert not bad, f"{name} replay {k}: outputs {bad[:8]} differ from the per-view loop"
    covered = sum(int((v.visbuffer() != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum()) for v in br.views)
    print("covered pixels over 64 cameras:", covered, flush=True)
    assert covered > 1_000_000


def test_batch_64_cameras_extremes_match_oracle(batch64):
    import torch

    from lighthugger_amd.lib import BATCH_BATCHED

    scene, cams, br = batch64
    br.run(BATCH_BATCHED)
    torch.cuda.synchronize()
    sides = [int(c.width) for c in cams]
    for i in (int(np.argmin(sides)), int(np.argmax(sides))):
        vd = br.views[i]
        print(f"camera {i}: {sides[i]}^2", flush=True)
        orc = oracle_frame(scene, vd.cfg, make_uniforms(cams[i], scene.num_instances), threads=THREADS)
        print(compare_frames(snapshot(vd), orc, scene), flush=True)


def test_backward_full_size_matches_fp64_adjoint():
    import torch

    from lighthugger_amd.diff import DifferentiableResolve
    from lighthugger_amd.lib import Renderer
    from oracle import adjoint as A
    from oracle import lib as oracle_lib
    from oracle.compare import oracle_view

    from .test_gpu_backward import _groups, _rel

    side = 1024
    scene = HostScene.preset(abi.PRESET_LARGE, tri_scale_pct=18)
    base = preset_camera(abi.PRESET_LARGE)
    cam = make_camera(base.position.tolist(), base.target.tolist(), side, side)
    cfg = ViewConfig(side, side, side, big_capacity=1 << 20)
    print(f"case: large scene at 18% ({scene.total_triangles} triangles), {side}^2", flush=True)
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    oracle_lib().orc_frame_threads(sd.address, vd.address, THREADS)
    gpu, orc = snapshot(r.view), snapshot(vd)
    assert np.array_equal(gpu["visbuffer"], orc["visbuffer"]) and np.array_equal(gpu["shadow"], orc["shadow"])
    d = DifferentiableResolve(r)
    p = d.parameters()
    g = np.random.default_rng(11).standard_normal((side, side, 4)).astype(np.float32)
    d.forward_raw(p)
    out = d.raw_backward(p, torch.from_numpy(g).cuda(), recorded=True)
    torch.cuda.synchronize()
    assert not d.unsupported()
    g64 = A.grad64(sd, vd, A.scene_params(sd, vd), A.frozen(sd, vd), g)
    got = {k: v.cpu().numpy().astype(np.float64) for k, v in out.items()}
    pos64, nrm64 = _groups(scene, sd, g64["geometry"])
    pos, nrm = _groups(scene, sd, got["geometry"])
    errs = {"positions": _rel(pos, pos64), "normals": _rel(nrm, nrm64)}
    for k in ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity"):
        errs[k] = _rel(got[k], g64[k])
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert np.linalg.norm(pos64) > 0 and np.linalg.norm(nrm64) > 0
    for k, e in errs.items():
        assert e <= REL_L2, (k, e)
