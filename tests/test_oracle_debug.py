"""Debug views of the resolve (Uniforms.debug, render_geometry.comp:268-308)
and the shadow-map overlay of the display transform (Uniforms.debug_shadowmaps,
display_transform.comp:51-75), in the oracle.

Known answers: the triangle and meshlet views colour each pixel with
DEBUG_COLOURS[index % 10] of the index its key names; the normals view of a
quad facing the camera is (0.5, 0.5, 1); the cascade view is the base colour
times the colour of a cascade in 0..4; the overlay shows the four cascades in
the window's top-left quarter (cleared depth 1.0 -> white) and leaves the rest
as tone-mapped.
"""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.scene import HostScene, Material, make_uniforms
from oracle.compare import oracle_frame

from .util import fixture_camera, grid_patch, quad_mesh

# src/shaders/common/debug.glsl:1-11
DEBUG = np.array([[0, 2, 91], [0, 108, 251], [0, 221, 221], [51, 221, 0], [255, 252, 0], [255, 180, 0],
                  [255, 104, 0], [226, 22, 0], [191, 0, 83], [145, 0, 65]], np.float32) / np.float32(255.0)


def _scene():
    mat = Material(base_color=(0.8, 0.6, 0.4), roughness=0.5, metallic=0.0)
    # a 30 x 30 grid (1,800 triangles, several meshlets) facing +z
    return HostScene.build([grid_patch(30, 30)], [mat], [(0, 0, np.eye(4, dtype=np.float32))])


def _frame(debug=0, shadowmaps=0):
    scene = _scene()
    cam, cfg = fixture_camera(96, 64, position=(0, 0, 2.2))
    u = make_uniforms(cam, scene.num_instances)
    u["debug"] = debug
    u["debug_shadowmaps"] = shadowmaps
    return scene, oracle_frame(scene, cfg, u)


def _covered(f):
    return f["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)


def test_triangle_and_meshlet_views():
    scene, f = _frame(abi.UNIFORMS_DEBUG_TRIANGLE_INDEX)
    cov = _covered(f)
    assert cov.sum() > 2000
    keys = f["visbuffer"][cov]
    tri = (keys & np.uint64(127)).astype(np.int64)
    assert np.array_equal(f["radiance"][cov][:, :3], DEBUG[tri % 10])
    assert len(np.unique(tri % 10)) == 10
    scene, f = _frame(abi.UNIFORMS_DEBUG_INSTANCE_INDEX)
    keys = f["visbuffer"][cov]
    mid = ((keys >> np.uint64(7)) & np.uint64(0xFFFFFF)).astype(np.int64)
    ml = scene.meshlet_table["meshlet_index"][mid].astype(np.int64)
    assert np.array_equal(f["radiance"][cov][:, :3], DEBUG[ml % 10])
    assert len(np.unique(ml)) > 3


def test_normals_and_cascade_views():
    _, f = _frame(abi.UNIFORMS_DEBUG_NORMALS)
    cov = _covered(f)
    np.testing.assert_allclose(f["radiance"][cov][:, :3], np.broadcast_to([0.5, 0.5, 1.0], (int(cov.sum()), 3)), atol=1e-6)
    _, f = _frame(abi.UNIFORMS_DEBUG_CASCADES)
    ratio = f["radiance"][cov][:, :3] / np.array([0.8, 0.6, 0.4], np.float32)
    d = np.abs(ratio[:, None, :] - DEBUG[None, :5, :]).max(axis=2)  # cascades 0..3, or 4 outside every cascade
    assert (d.min(axis=1) < 1e-5).all()


@pytest.mark.parametrize("mode", [abi.UNIFORMS_DEBUG_OFF, abi.UNIFORMS_DEBUG_SHADER_CLOCK])
def test_off_and_shader_clock_shade_normally(mode):
    _, a = _frame(abi.UNIFORMS_DEBUG_OFF)
    _, b = _frame(mode)
    assert np.array_equal(a["radiance"], b["radiance"])


def test_shadow_map_overlay():
    _, a = _frame()
    _, b = _frame(shadowmaps=1)
    H, W = a["ldr"].shape
    ys, xs = np.mgrid[0:H, 0:W]
    quarter = ((xs + 0.5) / W < 0.5) & ((ys + 0.5) / H < 0.5)
    assert np.array_equal(a["ldr"][~quarter], b["ldr"][~quarter])
    assert np.array_equal(a["radiance"], b["radiance"])  # the overlay is display-only
    q = b["ldr"][quarter]
    grey = (q & 0xFF) == ((q >> 8) & 0xFF)
    assert grey.all() and ((q >> 16) & 0xFF == q & 0xFF).all()
    assert (q == 0xFFFFFFFF).any()  # cleared shadow texels: depth 1.0 -> white


def _lut(fn):
    """LH_TONEMAP_LUT_DIMS^3 RGBA float LUT, red fastest: entry (r, g, b) = fn(r, g, b) / 47."""
    n = abi.TONEMAP_LUT_DIMS
    b, g, r = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rgb = np.stack(fn(r, g, b), -1).astype(np.float32) / np.float32(n - 1)
    return np.ascontiguousarray(np.concatenate([rgb, np.ones(rgb.shape[:-1] + (1,), np.float32)], -1))


def _lut_frame(lut):
    from oracle import lib as oracle_lib
    from oracle.compare import oracle_view, snapshot

    scene = _scene()
    cam, cfg = fixture_camera(96, 64, position=(0, 0, 2.2))
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    if lut is not None:
        vd.c.tonemap_lut = lut.ctypes.data
    oracle_lib().orc_frame(sd.address, vd.address)
    return snapshot(vd)


def _channels(ldr):
    return np.stack([(ldr >> 16) & 0xFF, (ldr >> 8) & 0xFF, ldr & 0xFF], -1).astype(np.int64)


def test_tonemap_lut_known_answers():
    """tony_mc_mapface with a supplied LUT (display_transform.comp:5-17): a
    constant LUT of 0.5 gives sRGB(0.5) everywhere; the identity LUT gives
    the LUT-free curve x / (x + 1) to one step; a LUT that swaps red and blue
    swaps the channels (the axis order: red fastest)."""
    plain = _channels(_lut_frame(None)["ldr"])
    const = _channels(_lut_frame(_lut(lambda r, g, b: (r * 0 + 23.5, g * 0 + 23.5, b * 0 + 23.5)))["ldr"])
    x = 0.5
    srgb = 1.055 * x ** (1 / 2.4) - 0.055
    assert (const == int(srgb * 255 + 0.5)).all()
    ident = _channels(_lut_frame(_lut(lambda r, g, b: (r, g, b)))["ldr"])
    assert np.abs(ident - plain).max() <= 1
    swap = _channels(_lut_frame(_lut(lambda r, g, b: (b, g, r)))["ldr"])
    assert np.abs(swap - ident[..., ::-1]).max() <= 1
    assert (np.abs(ident[..., 0] - ident[..., 2]) > 3).any()  # the image is not grey
