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


def test_large_frame_parity_full_size():
    """The benchmark workload at full size: 2,000 instances, 1920x1080, 4 x 2048^2."""
    scene = HostScene.preset(abi.PRESET_LARGE)
    cam = preset_camera(abi.PRESET_LARGE)
    rep = _run(scene, cam, ViewConfig(cam.width, cam.height, 2048, big_capacity=1 << 21))
    assert rep["pass0"]["meshlets"] > 10000


@pytest.mark.parametrize("preset", ["small", "large"])
def test_quantized_vertex_frame_parity(preset):
    """The reference's vertex formats (u16 positions on a per-mesh grid, i8
    normals, u16 uvs, u16 remap indices) through every kernel, at the same
    bars; the large case is the benchmark workload at full size."""
    p = {"small": abi.PRESET_SMALL, "large": abi.PRESET_LARGE}[preset]
    scene = HostScene.preset(p, vertex_format=abi.VERTEX_FORMAT_QUANTIZED)
    assert not (scene.mesh_infos["flags"] & abi.MESH_INFO_FLAGS_FP32_VERTICES).any()
    cam = preset_camera(p)
    shadow = 1024 if preset == "small" else 2048
    rep = _run(scene, cam, ViewConfig(cam.width, cam.height, shadow, big_capacity=1 << 21))
    assert rep["covered_pixels"] > 10000


def test_alpha_clip_frame_parity():
    """Alpha-clip geometry (double-sided, base-colour alpha < 0.5 discarded)
    through the alpha instantiations of the meshlet and large-triangle rasters
    and the clip path, main view and cascades: keys and shadow depths
    bit-exact against the oracle, covered samples per pass equal."""
    from .fixtures import alpha_scene

    scene, cam, cfg = alpha_scene()
    rep = _run(scene, cam, cfg)
    assert rep["covered_pixels"] > 1000
    stats = ViewConfig(cfg.width, cfg.height, cfg.shadow_size, big_capacity=cfg.big_capacity,
                       clip_capacity=cfg.clip_capacity, statistics=True)
    gpu = _gpu_frame(scene, cam, stats)
    orc = oracle_frame(scene, stats, make_uniforms(cam, scene.num_instances))
    compare_frames(gpu, orc, scene)
    for p in range(abi.NUM_PASSES):
        assert int(gpu["counters"][abi.CNT_FRAGMENTS + p]) == int(orc["counters"][abi.CNT_FRAGMENTS + p]), p
    assert gpu["counters"][abi.CNT_ALPHA_DRAWS] > 0
    assert gpu["counters"][abi.CNT_BIG] > 0 and gpu["counters"][abi.CNT_CLIP] > 0


@pytest.mark.parametrize("mode", [abi.UNIFORMS_DEBUG_CASCADES, abi.UNIFORMS_DEBUG_TRIANGLE_INDEX,
                                  abi.UNIFORMS_DEBUG_INSTANCE_INDEX, abi.UNIFORMS_DEBUG_NORMALS])
def test_debug_views_match_oracle(mode):
    """Uniforms.debug views of the resolve and (with the normals view) the
    shadow-map overlay of the display transform, against the oracle."""
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    cfg = ViewConfig(cam.width, cam.height, 512)
    u = make_uniforms(cam, scene.num_instances)
    u["debug"] = mode
    u["debug_shadowmaps"] = 1 if mode == abi.UNIFORMS_DEBUG_NORMALS else 0
    r = Renderer(scene, cfg)
    r.view.set_uniforms(u, r.scene)
    r.frame()
    torch.cuda.synchronize()
    gpu = snapshot(r.view)
    orc = oracle_frame(scene, cfg, u)
    print(mode, compare_frames(gpu, orc, scene), flush=True)


def test_tonemap_lut_matches_oracle():
    """A supplied display LUT (lh_view.tonemap_lut, random contents) through
    the fused resolve + display transform, against the oracle."""
    import torch

    from lighthugger_amd.lib import Renderer
    from oracle import lib as oracle_lib
    from oracle.compare import oracle_view

    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    cfg = ViewConfig(cam.width, cam.height, 512)
    n = abi.TONEMAP_LUT_DIMS
    lut = np.random.default_rng(3).random((n, n, n, 4), dtype=np.float32)
    dev = torch.from_numpy(lut).cuda()
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.view.c.tonemap_lut = dev.data_ptr()
    r.frame()
    torch.cuda.synchronize()
    gpu = snapshot(r.view)
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    vd.c.tonemap_lut = lut.ctypes.data
    oracle_lib().orc_frame(sd.address, vd.address)
    orc = snapshot(vd)
    print(compare_frames(gpu, orc, scene), flush=True)
    assert (gpu["ldr"] != orc["ldr"]).mean() < 1e-3


def test_skybox_and_shader_clock_view():
    """A supplied skybox cubemap (lh_view.skybox, random faces): every pixel
    matches the oracle at the radiance bar.  The shader-clock debug view
    (timed on the device clock, so not comparable bit for bit) colours every
    shaded pixel with a point of the heat-map curve temperature(t)."""
    import torch

    from lighthugger_amd.lib import Renderer
    from oracle import lib as oracle_lib
    from oracle.compare import oracle_view

    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    cfg = ViewConfig(cam.width, cam.height, 512)
    faces = np.random.default_rng(4).random((6, 96, 96, 4), dtype=np.float32)
    dev = torch.from_numpy(faces).cuda()
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.view.c.skybox, r.view.c.skybox_size = dev.data_ptr(), 96
    r.frame()
    torch.cuda.synchronize()
    gpu = snapshot(r.view)
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    vd.c.skybox, vd.c.skybox_size = faces.ctypes.data, 96
    oracle_lib().orc_frame(sd.address, vd.address)
    orc = snapshot(vd)
    rep = compare_frames(gpu, orc, scene)
    sky = gpu["visbuffer"] == np.uint64(abi.LH_VISBUFFER_CLEAR)
    print(rep, "sky pixels", int(sky.sum()), flush=True)
    assert sky.sum() > 1000 and np.unique(gpu["radiance"][sky][:, 0]).size > 100
    # shader-clock view
    u = make_uniforms(cam, scene.num_instances)
    u["debug"] = abi.UNIFORMS_DEBUG_SHADER_CLOCK
    r.view.set_uniforms(u, r.scene)
    r.frame()
    torch.cuda.synchronize()
    rad = r.view.radiance()[..., :3]
    cov = ~sky
    curve = np.zeros((8001, 3), np.float32)
    for k, t in enumerate(np.linspace(0.0, 1.0, 8001)):
        oracle_lib().orc_temperature(float(t), curve[k].ctypes.data)
    got = np.unique(rad[cov], axis=0)
    d = np.concatenate([np.min(np.linalg.norm(got[i: i + 1024, None, :] - curve[None, :, :], axis=2), axis=1)
                        for i in range(0, len(got), 1024)])
    print("clock view:", len(got), "distinct colours, max distance to the heat-map curve", float(d.max()), flush=True)
    assert d.max() < 2e-3
    assert np.allclose(rad[sky], gpu["radiance"][sky][:, :3])  # the sky is unchanged


@pytest.mark.parametrize("texture_format", [abi.TEXTURE_BC1, abi.TEXTURE_BC1 | abi.TEXTURE_SRGB,
                                            abi.TEXTURE_BC7 | abi.TEXTURE_SRGB, abi.TEXTURE_RGBA8 | abi.TEXTURE_SRGB],
                         ids=["bc1", "bc1_srgb", "bc7_srgb", "rgba8_srgb"])
def test_block_texture_frame_parity(texture_format):
    """The benchmark workload at full size with every texture stored in the
    given format (BC1 / BC7 blocks decoded on fetch, sRGB decoded per texel
    before filtering): keys bit-exact, radiance at the usual bar."""
    scene = HostScene.preset(abi.PRESET_LARGE, texture_format=texture_format)
    cam = preset_camera(abi.PRESET_LARGE)
    rep = _run(scene, cam, ViewConfig(cam.width, cam.height, 2048, big_capacity=1 << 21))
    assert rep["covered_pixels"] > 100000


def test_fixtures_match_oracle_and_golden():
    import hashlib
    import json
    from pathlib import Path

    from .fixtures import fixtures

    golden = json.loads((Path(__file__).resolve().parent / "golden" / "fixtures.json").read_text())
    for name, (scene, cam, cfg) in fixtures().items():
        print("fixture", name, flush=True)
        gpu = _gpu_frame(scene, cam, cfg)
        orc = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))
        compare_frames(gpu, orc, scene)
        assert hashlib.sha256(np.ascontiguousarray(gpu["visbuffer"]).tobytes()).hexdigest() == \
            golden[name]["visbuffer_sha256"], name
        assert hashlib.sha256(np.ascontiguousarray(gpu["shadow"]).tobytes()).hexdigest() == \
            golden[name]["shadow_sha256"], name


def test_per_pass_calls_equal_lh_frame():
    """Calling the passes one by one through the C ABI (the reference's
    render() order) gives the same keys as lh_frame; lh_frame_profiled too."""
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    r = Renderer(scene, ViewConfig(cam.width, cam.height, 512))
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    a = snapshot(r.view)
    r.call("lh_view_reset")
    r.call("lh_cull_instances")
    r.call("lh_cull_meshlets", 0)
    r.call("lh_raster_view")
    r.call("lh_depth_bounds")
    r.call("lh_fit_cascades")
    r.call("lh_cull_instances_shadow")
    for c in range(4):
        r.call("lh_cull_meshlets", 1 + c)
        r.call("lh_raster_shadow", c)
    r.call("lh_resolve")
    torch.cuda.synchronize()
    b = snapshot(r.view)
    for k in ("visbuffer", "shadow", "radiance", "ldr"):
        assert np.array_equal(a[k], b[k]), k
    times = r.frame_profiled()
    names = [n for n, _ in times]
    # the target clears ride on the meshlet culls
    assert names[:3] == ["view_reset", "cull_instances", "cull_meshlets_view"] and names[-1] == "resolve"
    assert "clear_targets" not in names
    assert all(ms >= 0 for _, ms in times)
    c = snapshot(r.view)
    assert np.array_equal(a["visbuffer"], c["visbuffer"])


def test_output_independent_of_tuning():
    """The autotuner's parameters change speed only: keys, shadow depths and
    radiance are bit-identical across extreme settings."""
    import torch

    from lighthugger_amd.lib import Renderer, Tuning, get_tuning_classes, set_tuning, set_tuning_classes

    scene = HostScene.preset(abi.PRESET_LARGE, texture_size=64)
    cam = preset_camera(abi.PRESET_LARGE)
    r = Renderer(scene, ViewConfig(960, 540, 1024, big_capacity=1 << 20))
    from lighthugger_amd.scene import make_camera

    r.set_camera(make_camera(cam.position.tolist(), cam.target.tolist(), 960, 540))
    saved = get_tuning_classes()
    ref = None
    try:
        # the last two give the cascades extents of their own (every shadow
        # triangle in-lane up to 32 px; none in-lane and none wave-cooperative)
        for t in (Tuning(1, 8192, 2048, 256, 1024, 1, 8, 256, 1), Tuning(2, 4096, 512, 64, 512, 16, 512, 2048, 1),
                  Tuning(8, 16384, 2048, 256, 2048, 64, 32, 64, 0), Tuning(32, 1024, 64, 16, 128, 64, 0, 0, 1),
                  Tuning(4, 8192, 2048, 256, 1024, 8, 64, 1024, 1, 32, 64),
                  Tuning(16, 8192, 2048, 256, 1024, 64, 64, 1024, 1, 1, 1)):
            set_tuning(t)
            r.frame()
            torch.cuda.synchronize()
            s = snapshot(r.view)
            assert r.status()[abi.CNT_OVERFLOW] == 0
            if ref is None:
                ref = s
            for k in ("visbuffer", "shadow", "radiance", "ldr"):
                assert np.array_equal(ref[k], s[k]), (t.as_dict(), k)
    finally:
        set_tuning_classes(saved)


@pytest.mark.parametrize("hiz", [False, True])
def test_tiled_raster_equals_per_meshlet_raster(hiz, tiled_raster):
    """The main view through the LDS-binned tile raster (every surviving
    meshlet binned), partly binned (a bin capacity of 2,000 draws: the rest
    take the per-meshlet raster) and not at all (bin_capacity 0): identical
    keys, shadows and radiance, and the oracle's."""
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_LARGE)
    cam = preset_camera(abi.PRESET_LARGE)
    snaps = []
    for bins in (-1, 2000, 0):
        cfg = ViewConfig(cam.width, cam.height, 2048, big_capacity=1 << 21, bin_capacity=bins, hiz=hiz)
        r = Renderer(scene, cfg)
        r.set_camera(cam)
        r.frame()
        r.frame()  # with Hi-Z the second frame uses the first one's visibility history
        torch.cuda.synchronize()
        st = r.status()
        print(f"bins {bins}: tile items {int(st[abi.CNT_TILE_ITEMS])}, direct draws {int(st[abi.CNT_TILE_DIRECT])}",
              flush=True)
        assert st[abi.CNT_OVERFLOW] == 0
        if bins == 2000:
            assert st[abi.CNT_TILE_DIRECT] > 0
        snaps.append(snapshot(r.view))
        del r
    for s in snaps[1:]:
        for k in ("visbuffer", "shadow", "radiance", "ldr"):
            assert np.array_equal(snaps[0][k], s[k]), k
    if not hiz:
        orc = oracle_frame(scene, ViewConfig(cam.width, cam.height, 2048), make_uniforms(cam, scene.num_instances),
                           threads=8)
        print(compare_frames(snaps[0], orc, scene), flush=True)


def test_capacity_overflow_reported_on_device():
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    r = Renderer(scene, ViewConfig(cam.width, cam.height, 256, instance_capacity=8, draw_capacity=100,
                                   big_capacity=4, clip_capacity=4))
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    st = r.status()
    assert st[abi.CNT_OVERFLOW] & 1 and st[abi.CNT_OVERFLOW] & 2
    count, _, vals = r.view.prefix(0)
    assert count > 8 and len(vals) == 8


def test_empty_view_on_device():
    import torch

    from lighthugger_amd.lib import Renderer
    from lighthugger_amd.scene import make_camera

    scene = HostScene.preset(abi.PRESET_SMALL, num_instances=4)
    cam = make_camera((0, 500, 0), (0, 1000, 1), 64, 48)
    cfg = ViewConfig(64, 48, 64)
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    gpu = snapshot(r.view)
    assert gpu["counters"][abi.CNT_FLAGS] & 1
    orc = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))
    compare_frames(gpu, orc, scene)


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


def test_raster_statistics_match_oracle():
    """LH_VIEW_STATISTICS: covered samples per pass equal the oracle's count
    exactly; issued depth atomics are at most the covered samples; the keys are
    those of a frame without statistics."""
    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    base = ViewConfig(cam.width, cam.height, 512)
    cfg = ViewConfig(cam.width, cam.height, 512, statistics=True)
    plain = _gpu_frame(scene, cam, base)
    gpu = _gpu_frame(scene, cam, cfg)
    orc = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))
    compare_frames(gpu, orc, scene)
    for k in ("visbuffer", "shadow", "radiance"):
        assert np.array_equal(plain[k], gpu[k]), k
    g, o = gpu["counters"], orc["counters"]
    for p in range(abi.NUM_PASSES):
        frags = int(g[abi.CNT_FRAGMENTS + p])
        print(f"pass {p}: fragments {frags} atomics {int(g[abi.CNT_ATOMICS + p])}", flush=True)
        assert frags == int(o[abi.CNT_FRAGMENTS + p]), p
        assert int(g[abi.CNT_ATOMICS + p]) <= frags and (frags == 0) == (g[abi.CNT_ATOMICS + p] == 0), p
    assert g[abi.CNT_FRAGMENTS] > 10000
    assert not plain["counters"][abi.CNT_FRAGMENTS:abi.CNT_FRAGMENTS + abi.NUM_PASSES].any()


def test_raster_statistics_leave_the_large_frame_unchanged():
    """The benchmark frame (its large-triangle queue holds ~33k slivers just
    over the medium extent) with LH_VIEW_STATISTICS: two frames in a row are
    bit-identical to the frame without statistics, and every pass counts at
    least one covered sample per covered pixel / texel."""
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_LARGE)
    cam = preset_camera(abi.PRESET_LARGE)
    plain = _gpu_frame(scene, cam, ViewConfig(cam.width, cam.height, 2048, big_capacity=1 << 21))
    r = Renderer(scene, ViewConfig(cam.width, cam.height, 2048, big_capacity=1 << 21, statistics=True))
    r.set_camera(cam)
    for k in range(2):
        r.frame()
        torch.cuda.synchronize()
        got = snapshot(r.view)
        for key in ("visbuffer", "shadow", "radiance", "ldr"):
            assert np.array_equal(plain[key], got[key]), (k, key)
    c = got["counters"]
    assert c[abi.CNT_BIG] > 1000  # the large-triangle raster ran
    assert int(c[abi.CNT_FRAGMENTS]) >= int((got["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum())
    for q in range(abi.NUM_CASCADES):
        covered = int((got["shadow"][q] != np.uint32(abi.LH_SHADOW_CLEAR)).sum())
        print(f"cascade {q}: fragments {int(c[abi.CNT_FRAGMENTS + 1 + q])}, covered texels {covered}", flush=True)
        assert int(c[abi.CNT_FRAGMENTS + 1 + q]) >= covered


def _hiz_pair(scene, cam, cfg_kw):
    """The same frame with and without LH_VIEW_HIZ."""
    import torch

    from lighthugger_amd.lib import Renderer

    out = []
    for hiz in (False, True):
        r = Renderer(scene, ViewConfig(cam.width, cam.height, hiz=hiz, **cfg_kw))
        r.set_camera(cam)
        r.frame()
        torch.cuda.synchronize()
        out.append((snapshot(r.view), r.status()))
    # a second occlusion-culled frame: phase 1 starts from the first frame's visibility
    r.frame()
    torch.cuda.synchronize()
    out.append((snapshot(r.view), r.status()))
    return out


@pytest.mark.parametrize("preset", ["small", "large"])
def test_hiz_keys_identical(preset):
    """Two-phase occlusion culling is conservative: keys, shadow depths and
    radiance are bit-identical with it on and off, and it does cull."""
    p = {"small": abi.PRESET_SMALL, "large": abi.PRESET_LARGE}[preset]
    scene = HostScene.preset(p)
    cam = preset_camera(p)
    shadow = 1024 if preset == "small" else 2048
    (off, st_off), (on, st_on), (warm, st_warm) = _hiz_pair(scene, cam, dict(shadow_size=shadow, big_capacity=1 << 21))
    assert st_on[abi.CNT_OVERFLOW] == 0 and st_off[abi.CNT_OVERFLOW] == 0 and st_warm[abi.CNT_OVERFLOW] == 0
    for k in ("visbuffer", "shadow", "radiance", "ldr", "misc"):
        assert np.array_equal(off[k], on[k]), k
        assert np.array_equal(off[k], warm[k]), k
    print(f"{preset} second frame: phase 1 {int(st_warm[abi.CNT_HIZ_PHASE1])}, candidates "
          f"{int(st_warm[abi.CNT_HIZ_CANDIDATES])}, drawn {int(st_warm[abi.CNT_DRAWS])}", flush=True)
    cand, phase1, drawn = int(st_on[abi.CNT_HIZ_CANDIDATES]), int(st_on[abi.CNT_HIZ_PHASE1]), int(st_on[abi.CNT_DRAWS])
    total = int(st_off[abi.CNT_DRAWS])
    print(f"{preset}: {total} meshlets; phase 1 {phase1}, candidates {cand}, drawn {drawn} "
          f"({total - drawn} culled by Hi-Z)", flush=True)
    assert phase1 + cand == total and phase1 <= drawn <= total
    # what was drawn is a subset of the frustum/cone survivors
    a = {tuple(x) for x in on["draws0"][["instance_index", "meshlet_index"]].tolist()}
    b = {tuple(x) for x in off["draws0"][["instance_index", "meshlet_index"]].tolist()}
    assert a <= b and len(a) == drawn
    if preset == "large":
        assert drawn < total


def test_hiz_fixtures_identical():
    from .fixtures import fixtures

    for name, (scene, cam, cfg) in fixtures().items():
        kw = dict(shadow_size=cfg.shadow_size, big_capacity=cfg.big_capacity, clip_capacity=cfg.clip_capacity)
        (off, _), (on, _), (warm, _) = _hiz_pair(scene, cam, kw)
        for k in ("visbuffer", "shadow", "radiance"):
            assert np.array_equal(off[k], on[k]), (name, k)
            assert np.array_equal(off[k], warm[k]), (name, k)
