"""Whole-frame oracle behaviour: golden fixtures, cross-pass invariants and
host-forced failures (capacity overflow, empty view, non-finite geometry).

Invariants tie the passes together independently of their formulas: every
meshlet id in the visibility buffer belongs to the visible meshlet set; the
depth bounds are the min (non-zero) / max of the buffer's depth bits
(src/shaders/compute/read_depth.glsl:3-18); prefix sectors are consistent
inclusive scans (src/shaders/common/prefix_sum.glsl:17-30).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig, draw_keys
from lighthugger_amd.scene import HostScene, make_camera, make_uniforms, preset_camera
from oracle.compare import check_prefix_invariants, instance_meshlet_counts, oracle_frame, oracle_view, snapshot

from .fixtures import fixtures
from .golden.make_golden import summarize

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "fixtures.json").read_text())


@pytest.fixture(scope="module")
def fx():
    return fixtures()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_fixture_matches_golden(fx, name):
    scene, cam, cfg = fx[name]
    got = summarize(oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances)))
    want = GOLDEN[name]
    for k in ("covered_pixels", "visbuffer_sha256", "shadow_sha256", "draws", "min_depth", "max_depth"):
        assert got[k] == want[k], (name, k)
    np.testing.assert_allclose(got["radiance_mean"], want["radiance_mean"], rtol=1e-6)


@pytest.fixture(scope="module")
def small_frame():
    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    cfg = ViewConfig(cam.width, cam.height, 1024)
    return scene, oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))


def test_small_frame_cross_pass_invariants(small_frame):
    scene, snap = small_frame
    nm = instance_meshlet_counts(scene)
    for p in range(abi.NUM_PASSES):
        check_prefix_invariants(*snap[f"prefix{p}"], nm)
    vis = snap["visbuffer"]
    cov = vis != np.uint64(abi.LH_VISBUFFER_CLEAR)
    assert cov.sum() > 10000
    mids = np.unique((vis[cov] >> np.uint64(7)) & np.uint64((1 << 25) - 1)).astype(np.int64)
    t = scene.meshlet_table[mids]
    drawn = set(draw_keys(snap["draws0"]).tolist())
    keys = (t["instance_index"].astype(np.uint64) << np.uint64(16)) | t["meshlet_index"].astype(np.uint64)
    assert set(keys.tolist()) <= drawn, "a visible key names a meshlet that was culled"
    tri = (vis[cov] & np.uint64(127)).astype(np.int64)
    mi = scene.instances["mesh_info_address"].astype(np.int64)
    counts = np.concatenate([scene.meshlets(int(mi[i]))["triangle_count"][None] for i in range(1)])
    assert tri.max() < counts.max()
    depth = (~(vis >> np.uint64(32))).astype(np.uint32)
    assert int(snap["misc"]["min_depth"]) == int(depth[depth != 0].min())
    assert int(snap["misc"]["max_depth"]) == int(depth.max())
    assert snap["counters"][abi.CNT_FLAGS] == 0


def test_radiance_is_finite_and_sky_where_empty(small_frame):
    _, snap = small_frame
    rad = snap["radiance"]
    assert np.isfinite(rad).all()
    sky = snap["visbuffer"] == np.uint64(abi.LH_VISBUFFER_CLEAR)
    np.testing.assert_array_equal(rad[sky][:, :3], np.tile(np.array([2.8, 4.0, 6.4], np.float32), (sky.sum(), 1)))
    assert (rad[..., 3] == 1).all()


def _run(scene, cam, cfg):
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    st = oracle.lib().orc_frame(sd.address, vd.address)
    return st, snapshot(vd)


def test_instance_capacity_overflow_is_reported_not_written_past():
    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    st, snap = _run(scene, cam, ViewConfig(cam.width, cam.height, 256, instance_capacity=8))
    assert st == abi.LH_ERR_CAPACITY
    assert snap["counters"][abi.CNT_OVERFLOW] & 1
    count, _, vals = snap["prefix0"]
    assert count > 8 and len(vals) == 8


def test_draw_capacity_overflow():
    scene = HostScene.preset(abi.PRESET_SMALL)
    cam = preset_camera(abi.PRESET_SMALL)
    st, snap = _run(scene, cam, ViewConfig(cam.width, cam.height, 256, draw_capacity=100))
    assert st == abi.LH_ERR_CAPACITY
    assert snap["counters"][abi.CNT_OVERFLOW] & 2
    assert snap["counters"][0] > 100 and len(snap["draws0"]) == 100


def test_empty_view_sets_flag_and_renders_sky():
    scene = HostScene.preset(abi.PRESET_SMALL, num_instances=4)
    cam = make_camera((0, 500, 0), (0, 1000, 1), 64, 48)  # looking up, away from everything
    st, snap = _run(scene, cam, ViewConfig(64, 48, 64))
    assert st == abi.LH_OK
    assert snap["counters"][abi.CNT_FLAGS] & 1
    assert (snap["visbuffer"] == np.uint64(abi.LH_VISBUFFER_CLEAR)).all()
    assert np.isfinite(snap["radiance"]).all()


def test_non_finite_instance_is_skipped():
    from .util import quad_mesh, translate

    mesh = quad_mesh(1.0)
    bad = translate(0, 0, 0)
    bad[0, 0] = np.nan
    scene = HostScene.build([mesh], [oracle_material()], [(0, 0, translate(-1.5, 0, 0)), (0, 0, bad)])
    cam = make_camera((0, 0, 5), (0, 0, 0), 64, 48)
    st, snap = _run(scene, cam, ViewConfig(64, 48, 64))
    assert st == abi.LH_OK
    count, _, vals = snap["prefix0"]
    assert vals["index"].tolist() == [0]


def oracle_material():
    from lighthugger_amd.scene import Material

    return Material(base_color=(0.5, 0.5, 0.5), roughness=0.5, metallic=0.0)


def test_threaded_oracle_frame_is_identical():
    """orc_frame_threads (the CPU baseline) produces the scalar frame bit for bit."""
    import numpy as np

    from lighthugger_amd import abi
    from lighthugger_amd.frame import ViewConfig
    from lighthugger_amd.scene import HostScene, make_camera, make_uniforms, preset_camera
    from oracle import lib
    from oracle.compare import oracle_view, snapshot

    scene = HostScene.preset(abi.PRESET_LARGE, num_instances=300, texture_size=32)
    base = preset_camera(abi.PRESET_LARGE)
    cam = make_camera(base.position.tolist(), base.target.tolist(), 320, 200)
    cfg = ViewConfig(320, 200, 256)
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    lib().orc_frame(sd.address, vd.address)
    a = snapshot(vd)
    lib().orc_frame_threads(sd.address, vd.address, 4)
    b = snapshot(vd)
    for k in ("visbuffer", "shadow", "radiance", "ldr", "misc", "counters"):
        assert np.array_equal(a[k], b[k]), k


def test_quantized_frame_close_to_fp32_frame():
    """The oracle reads the reference's quantized formats (u16 positions, i8
    normals, u16 uvs, u16 remap indices): the quantized scene renders the
    same triangles as the FP32 scene at nearly every pixel (positions moved
    by at most half a grid step, ~1e-5 of the mesh size), and the shading
    (through i8 normals) stays close.  A closeness check of the format, not a
    parity bar: the GPU parity tests compare both formats with the oracle
    bit for bit."""
    cam = preset_camera(abi.PRESET_SMALL)
    cfg = ViewConfig(cam.width // 2, cam.height // 2, 512)
    cam = make_camera(cam.position.tolist(), cam.target.tolist(), cfg.width, cfg.height)
    frames = []
    for vf in (abi.VERTEX_FORMAT_FP32, abi.VERTEX_FORMAT_QUANTIZED):
        scene = HostScene.preset(abi.PRESET_SMALL, vertex_format=vf)
        frames.append(oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances)))
    a, b = frames
    clear = np.uint64(abi.LH_VISBUFFER_CLEAR)
    ids_a, ids_b = a["visbuffer"] & np.uint64(0xFFFFFFFF), b["visbuffer"] & np.uint64(0xFFFFFFFF)
    covered = a["visbuffer"] != clear
    assert covered.sum() > 4000
    same = (ids_a == ids_b) & covered
    assert same.sum() >= 0.99 * covered.sum(), same.sum() / covered.sum()
    ra, rb = a["radiance"][..., :3][same], b["radiance"][..., :3][same]
    rel = np.abs(ra - rb) / np.maximum(np.abs(ra), 1e-3)
    # i8 normals move the shading by ~1% (more in highlights); shadow edges flip a few pixels
    assert np.median(rel) < 1e-2 and np.quantile(rel, 0.9) < 0.1
