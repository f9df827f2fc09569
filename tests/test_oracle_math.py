"""Pins the oracle's cascade fitting, barycentrics, BRDF and display transform
by invariants and closed forms (not by retyping the formulas).

generate_shadow_matrices (src/shaders/compute/generate_shadow_matrices.glsl:15-110):
every corner of a cascade's depth slice lands inside its ortho box; slice
boundaries sit at the split distances; radius is a 1/16 multiple covering the
slice; the shadow view looks along -sun_dir.
CalcFullBary (src/shaders/common/vbuffer.glsl:11-63): unit vectors at the
vertices, world position equal to the FP64 ray/triangle intersection, ddx and
ddy equal to the barycentric difference one pixel right / one pixel up.
BRDF (src/shaders/common/pbr.glsl:44-96): normal-incidence closed forms,
grazing light, NaN guard.  sRGB OETF (display_transform.comp:19-25) closed form.
"""
import numpy as np
import pytest

import oracle
from lighthugger_amd import abi
from lighthugger_amd.frame import HostMemory, SceneData, ViewConfig, ViewData
from lighthugger_amd.scene import Material, make_camera, make_uniforms

from .util import IDENTITY, quad_mesh, single_scene


def _fit(cam, min_depth, max_depth):
    scene = single_scene(quad_mesh(), [IDENTITY])
    mem = HostMemory()
    sd = SceneData(scene, mem)
    vd = ViewData(ViewConfig(cam.width, cam.height, 64), sd, mem)
    vd.set_uniforms(make_uniforms(cam, 1), sd)
    misc = np.zeros(1, abi.MISC_STORAGE)
    misc["min_depth"] = np.float32(min_depth).view(np.uint32)
    misc["max_depth"] = np.float32(max_depth).view(np.uint32)
    vd.write_misc(misc[0])
    assert oracle.lib().orc_fit_cascades(sd.address, vd.address) == 0
    return vd.misc(), make_uniforms(cam, 1)[0]


def _mat(m):  # [column][row] -> row-major 4x4 float64
    return np.asarray(m, np.float64).T


@pytest.mark.parametrize("split_pow", [3.0, 1.0, 2.0])
def test_cascade_slices_fit_their_ortho_boxes(split_pow):
    cam = make_camera((3, 8, 20), (0, 0, 0), 640, 360, near=0.1, sun_dir=(0.35, 0.85, 0.4),
                      cascade_split_pow=split_pow)
    near = 0.1
    dmin, dmax = 8.0, 120.0  # view distances of nearest / farthest geometry
    misc, u = _fit(cam, near / dmax, near / dmin)
    inv = _mat(u["inv_perspective_view"])
    view = _mat(u["view"])
    a, b = 0.5 ** split_pow, 0.75 ** split_pow
    dist_bounds = [(dmin, dmin), (dmin, dmin + (dmax - dmin) * a), (dmin + (dmax - dmin) * a, dmin + (dmax - dmin) * b),
                   (dmin + (dmax - dmin) * b, dmax)]
    sun = np.array([0.35, 0.85, 0.4]) / np.linalg.norm([0.35, 0.85, 0.4])
    for c in range(4):
        M = _mat(misc["shadow_matrices"][c])
        r = float(misc["shadow_sphere_radii"][c])
        assert r * 16 == np.round(r * 16), "radius must be a multiple of 1/16"
        d0, d1 = dist_bounds[c]
        for d in (d0, d1):
            for sx in (-1, 1):
                for sy in (-1, 1):
                    p = inv @ np.array([sx, sy, near / d, 1.0])
                    p = p / p[3]
                    vz = -(view @ p)[2]
                    assert abs(vz - d) < 2e-3 * d, "slice corner is not at the split distance"
                    q = M @ p
                    assert abs(q[0]) <= 1 + 1e-5 and abs(q[1]) <= 1 + 1e-5, "slice corner outside the ortho box"
                    assert -1e-6 <= q[2] <= 1 + 1e-6
        V = _mat(misc["shadow_view_matrices"][c])
        fwd = -V[2, :3]  # camera looks along -z of its view space
        assert np.allclose(fwd, -sun, atol=1e-5)
        UV = _mat(misc["uv_space_shadow_matrices"][c])
        bias = np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0, 0.5], [0, 0, 1, 0], [0, 0, 0, 1]])
        np.testing.assert_allclose(UV, bias @ M, rtol=1e-6, atol=1e-6)


def test_empty_view_gives_nan_matrices_and_flag():
    cam = make_camera((0, 0, 5), (0, 0, 0), 64, 48)
    misc, _ = _fit(cam, np.float32(np.uint32(0xFFFFFFFF).view(np.float32)), 0.0)
    assert np.isnan(misc["shadow_matrices"]).any()


def _bary(p0, p1, p2, ndc, win):
    out = np.zeros(9, np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (p0, p1, p2, ndc, win)]
    oracle.lib().orc_calc_full_bary(*(x.ctypes.data for x in a), out.ctypes.data)
    return out[:3], out[3:6], out[6:9]


def test_calc_full_bary_vertices_intersection_and_derivatives():
    W, H = 640, 480
    cam = make_camera((0.3, 0.2, 6), (0, 0, 0), W, H)
    u = make_uniforms(cam, 1)[0]
    PV = _mat(u["combined_perspective_view"])
    inv = _mat(u["inv_perspective_view"])
    tri = np.array([[-1.5, -1.0, 0.3], [1.7, -0.8, -1.2], [0.2, 1.6, 0.5]])
    clip = [(PV @ np.append(p, 1.0)).astype(np.float32) for p in tri]
    win = np.array([W, H], np.float32)
    for k in range(3):
        ndc = (clip[k][:2] / clip[k][3]).astype(np.float32)
        lam, _, _ = _bary(*clip, ndc, win)
        np.testing.assert_allclose(lam, np.eye(3)[k], atol=2e-5)
    rng = np.random.default_rng(5)
    for _ in range(20):
        w = rng.dirichlet([1, 1, 1])
        target = w @ tri
        q = PV @ np.append(target, 1.0)
        ndc = (q[:2] / q[3]).astype(np.float32)
        lam, ddx, ddy = _bary(*clip, ndc, win)
        # FP64 ray through ndc intersected with the triangle plane
        o = inv @ np.array([ndc[0], ndc[1], 1.0, 1.0])
        f = inv @ np.array([ndc[0], ndc[1], 1e-3, 1.0])
        o, f = o[:3] / o[3], f[:3] / f[3]
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        t = np.dot(tri[0] - o, n) / np.dot(f - o, n)
        hit = o + t * (f - o)
        np.testing.assert_allclose(lam @ tri, hit, atol=2e-4)
        lx, _, _ = _bary(*clip, ndc + np.array([2.0 / W, 0], np.float32), win)
        ly, _, _ = _bary(*clip, ndc - np.array([0, 2.0 / H], np.float32), win)
        np.testing.assert_allclose(ddx, lx - lam, atol=2e-5)
        np.testing.assert_allclose(ddy, ly - lam, atol=2e-5)


def _brdf(v, l, n, rough, metal, base, thin=0):
    out = np.zeros(3, np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (v, l, n, base)]
    oracle.lib().orc_brdf(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, rough, metal, a[3].ctypes.data, thin,
                          out.ctypes.data)
    return out


def test_brdf_closed_forms():
    z = [0, 0, 1]
    base = np.array([0.2, 0.5, 0.8])
    # normal incidence, roughness 1: D = 1/pi, V = 1/4, F = f0, Burley = 1/pi
    np.testing.assert_allclose(_brdf(z, z, z, 1.0, 0.0, base), (base + 0.01) / np.pi, rtol=1e-6)
    np.testing.assert_allclose(_brdf(z, z, z, 1.0, 1.0, base), base * 0.25 / np.pi, rtol=1e-6)
    # light in the tangent plane: no irradiance
    assert np.all(_brdf(z, [1, 0, 0], z, 0.5, 0.0, base) == 0)
    # view opposite to light: h = normalize(0) is NaN, guarded to 0
    assert np.all(_brdf([0, 0, 1], [0, 0, -1], z, 0.5, 0.3, base) == 0)
    # thin-walled light from behind reaches the diffuse term
    l_back = np.array([0.3, 0.0, -0.95]) / np.linalg.norm([0.3, 0.0, -0.95])
    assert np.all(_brdf(z, l_back, z, 1.0, 0.0, base, thin=0) == 0)
    assert np.all(_brdf(z, l_back, z, 1.0, 0.0, base, thin=1) > 0)


def test_srgb_oetf():
    f = oracle.lib().orc_srgb_oetf
    assert abs(f(0.5) - (1.055 * 0.5 ** (1 / 2.4) - 0.055)) < 1e-6
    assert abs(f(1.0) - 1.0) < 1e-6
    assert f(0.001) == pytest.approx(0.01292, rel=1e-6)
    assert abs(f(0.003131) - f(0.003130)) < 1e-4


def test_resolve_centre_pixel_closed_form():
    """A camera-facing quad lit along the view axis: at the centre pixel
    v = l = n, so radiance = pi * (b + 0.01) / pi + ambient 0.025 b."""
    W, H = 64, 48
    b = (0.3, 0.6, 0.9)
    scene = single_scene(quad_mesh(3.0), [IDENTITY], Material(base_color=b, roughness=1.0, metallic=0.0))
    cam = make_camera((0, 0, 5), (0, 0, 0), W, H, sun_dir=(0, 0, 1))
    mem = HostMemory()
    sd = SceneData(scene, mem)
    vd = ViewData(ViewConfig(W, H, 64), sd, mem)
    vd.set_uniforms(make_uniforms(cam, 1), sd)
    assert oracle.lib().orc_frame(sd.address, vd.address) == 0
    rad = vd.radiance()
    c = rad[H // 2, W // 2, :3]
    np.testing.assert_allclose(c, 1.025 * np.array(b) + 0.01, rtol=2e-6)
    # the corner pixel sees the sky
    np.testing.assert_array_equal(rad[0, 0, :3], np.array([0.35, 0.5, 0.8], np.float32) * np.float32(8))
