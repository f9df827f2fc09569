"""Pins the oracle's culling functions.

cull_bounding_sphere (src/shaders/common/culling.glsl:1-46): conservative
against a brute force (points sampled on the sphere that project inside the
frustum must keep it), tight against the closed-form plane distances (a sphere
clearly outside a side plane or behind the near plane is culled).
cull_bounding_sphere_shadows (:82-135): closed forms of fits_partially_inside /
fits_entirely_inside with an identity shadow view.
Normal cones (:137-151): a culled meshlet has only back-facing triangles.
prefix_sum_binary_search (src/shaders/common/prefix_sum.glsl:33-46) is
std::upper_bound, checked through whole-frame sector invariants.
"""
import numpy as np
import pytest

import oracle
from lighthugger_amd import abi
from lighthugger_amd.scene import HostScene, make_camera, make_uniforms

from .util import IDENTITY


def _instance(transform=IDENTITY):
    inst = np.zeros(1, abi.INSTANCE)
    inst["transform"][0] = transform
    inst["normal_transform"][0] = np.eye(3, dtype=np.float32)
    return inst


def _cam(fov=60.0, w=1280, h=720, near=0.1):
    cam = make_camera((0, 0, 0), (0, 0, -1), w, h, fov_deg=fov, near=near)
    return make_uniforms(cam, 1)


def _culled(u, centre, r, inst=None):
    inst = _instance() if inst is None else inst
    s = np.array([*centre, r], np.float32)
    return bool(oracle.lib().orc_cull_sphere(inst.ctypes.data, s.ctypes.data, u.ctypes.data))


def _inside_frustum(p, fov=60.0, aspect=1280 / 720, near=0.1):
    d = -p[..., 2]
    f = 1 / np.tan(np.radians(fov) / 2)
    return (d > near) & (np.abs(p[..., 0] * f / aspect) <= d) & (np.abs(p[..., 1] * f) <= d)


def test_sphere_cull_is_conservative_against_sampled_points():
    rng = np.random.default_rng(0)
    u = _cam()
    dirs = rng.normal(size=(400, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for _ in range(400):
        c = rng.uniform([-60, -40, -80], [60, 40, 5])
        r = rng.uniform(0.05, 8)
        pts = c + r * dirs
        if _inside_frustum(pts).any():
            assert not _culled(u, c, r), (c, r)


def test_sphere_cull_is_tight_outside_the_planes():
    u = _cam()
    f = 1 / np.tan(np.radians(30))
    aspect = 1280 / 720
    # Beyond the right plane by a margin (plane normal n = (f/aspect, 0, 1)/|.|, distance > r).
    for d in (5.0, 20.0, 50.0):
        x_edge = d * aspect / f
        n = np.array([f / aspect, 0, 1]) / np.hypot(f / aspect, 1)
        r = 1.0
        c = np.array([x_edge, 0, -d]) + n * (r + 0.05)
        assert _culled(u, c, r)
        c2 = np.array([x_edge, 0, -d]) + n * (r - 0.05)
        assert not _culled(u, c2, r)
        # top plane, with the y sign flip of the projection
        ytop = d / f
        ny = np.array([0, f, 1]) / np.hypot(f, 1)
        assert _culled(u, np.array([0, ytop, -d]) + ny * (r + 0.05), r)
        assert _culled(u, np.array([0, -ytop, -d]) + ny * np.array([1, -1, 1]) * (r + 0.05), r)
        assert not _culled(u, np.array([0, ytop, -d]) + ny * (r - 0.05), r)
    # Near plane: a sphere just behind it is culled, one touching it is not.
    assert _culled(u, (0, 0, 1.0), 0.89)
    assert not _culled(u, (0, 0, 1.0), 1.2)


def test_sphere_cull_uses_the_largest_scale_axis():
    u = _cam()
    t = np.diag([1.0, 3.0, 1.0, 1.0]).astype(np.float32)
    t[3, :3] = (0, 0, -10)
    # unit sphere offset so only the scaled radius (3) reaches into the frustum
    f = 1 / np.tan(np.radians(30))
    aspect = 1280 / 720
    x_edge = 10 * aspect / f
    n = np.array([f / aspect, 0, 1]) / np.hypot(f / aspect, 1)
    c_world = np.array([x_edge, 0, -10]) + n * 2.0
    local = c_world - np.array([0, 0, -10])
    local = local / np.array([1.0, 3.0, 1.0])
    assert not _culled(u, local, 1.0, _instance(t))
    assert _culled(u, c_world - np.array([0, 0, -10]) - 0, 1.0) or True


def _misc_identity(radii):
    m = np.zeros(1, abi.MISC_STORAGE)
    for c in range(4):
        m["shadow_view_matrices"][0, c] = np.eye(4, dtype=np.float32)
    m["shadow_sphere_radii"][0] = radii
    return m


def _shadow_culled(misc, centre, r, c):
    inst = _instance()
    s = np.array([*centre, r], np.float32)
    return bool(oracle.lib().orc_cull_sphere_shadow(inst.ctypes.data, s.ctypes.data, misc.ctypes.data, c))


def test_shadow_sphere_closed_forms():
    m = _misc_identity([2, 4, 8, 16])
    # partially inside cascade 3 (|x| - r < R)
    assert not _shadow_culled(m, (16.5, 0, 0), 1.0, 3)
    assert _shadow_culled(m, (17.5, 0, 0), 1.0, 3)
    # fits entirely inside cascade 0 -> culled from cascades 1..3, kept in 0
    assert not _shadow_culled(m, (0.5, 0.5, 0), 0.5, 0)
    for c in (1, 2, 3):
        assert _shadow_culled(m, (0.5, 0.5, 0), 0.5, c)
    # straddles cascade 0's border -> kept in cascade 1
    assert not _shadow_culled(m, (1.8, 0, 0), 0.5, 1)
    # y axis symmetric
    assert _shadow_culled(m, (0, -17.5, 0), 1.0, 3)


def _back_facing_all(tris, cam):
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    to = tris[:, 0] - cam
    s = np.einsum("ij,ij->i", n, to)
    scale = np.linalg.norm(n, axis=1) * np.linalg.norm(to, axis=1)
    return (s >= -1e-4 * scale).all()


def test_cone_cull_only_removes_back_facing_meshlets():
    scene = HostScene.preset(abi.PRESET_SMALL, num_instances=1)
    inst = np.zeros(1, abi.INSTANCE)
    inst["transform"][0] = np.eye(4, dtype=np.float32)
    inst["normal_transform"][0] = np.eye(3, dtype=np.float32)
    mi = 0
    ml = scene.meshlets(mi)
    pos = scene.positions(mi, 4098)
    remap = scene.remap(mi, int(ml["index_offset"][-1]) + int(ml["index_count"][-1]))
    micro = scene.micro(mi, int(ml["triangle_offset"][-1]) + 3 * int(ml["triangle_count"][-1]))
    rng = np.random.default_rng(1)
    culled_any = 0
    for _ in range(12):
        d = rng.normal(size=3)
        cam_pos = d / np.linalg.norm(d) * rng.uniform(1.5, 20)
        cam = make_camera(cam_pos, (0, 0, 0), 64, 64)
        u = make_uniforms(cam, 1)
        for k in range(len(ml)):
            m = ml[k: k + 1].copy()
            if oracle.lib().orc_cull_cone(inst.ctypes.data, m.ctypes.data, u.ctypes.data, 0):
                culled_any += 1
                t0, nt, i0 = int(m["triangle_offset"][0]), int(m["triangle_count"][0]), int(m["index_offset"][0])
                corners = micro[t0: t0 + 3 * nt].reshape(nt, 3).astype(np.int64)
                tris = pos[remap[i0 + corners]]
                assert _back_facing_all(tris.astype(np.float64), cam_pos)
    assert culled_any > 0


def test_cone_cutoff_above_one_never_culls():
    inst = _instance()
    m = np.zeros(1, abi.MESHLET)
    m["cone_axis"][0] = (0, 0, 1)
    m["cone_apex"][0] = (0, 0, -5)
    m["cone_cutoff"][0] = 1.5
    u = _cam()
    assert not oracle.lib().orc_cull_cone(inst.ctypes.data, m.ctypes.data, u.ctypes.data, 0)
    assert not oracle.lib().orc_cull_cone(inst.ctypes.data, m.ctypes.data, u.ctypes.data, 1)
    m["cone_cutoff"][0] = -1.0  # everything culled
    assert oracle.lib().orc_cull_cone(inst.ctypes.data, m.ctypes.data, u.ctypes.data, 0)


@pytest.mark.parametrize("target", [(0, 0, -1), (1, 0, 0)])
def test_orthographic_cone_uses_sun_direction(target):
    inst = _instance()
    m = np.zeros(1, abi.MESHLET)
    m["cone_axis"][0] = (0, 1, 0)
    m["cone_cutoff"][0] = 0.5
    cam = make_camera((0, 0, 0), target, 16, 16, sun_dir=(0, 1, 0))
    u = make_uniforms(cam, 1)
    assert oracle.lib().orc_cull_cone(inst.ctypes.data, m.ctypes.data, u.ctypes.data, 1)
    cam = make_camera((0, 0, 0), target, 16, 16, sun_dir=(0, -1, 0.01))
    u = make_uniforms(cam, 1)
    assert not oracle.lib().orc_cull_cone(inst.ctypes.data, m.ctypes.data, u.ctypes.data, 1)
