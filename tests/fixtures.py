"""Hand-built fixture scenes (single triangle, shared edge, closed cube,
near-plane straddler, a mesh of exactly 64 vertices / 98 triangles, and a
124-triangle meshlet), each with a camera and view configuration."""
from __future__ import annotations

import numpy as np

from lighthugger_amd.scene import HostScene, Material

from .util import cube_mesh, fixture_camera, grid_patch, quad_mesh, translate, triangle_mesh


def _rot_y(a):
    c, s = np.cos(a), np.sin(a)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c
    return m


def _strip_124():
    """One meshlet with 124 triangles: a 2 x 62 quad strip has 126 vertices,
    so use a 63 x 1 strip of quads (128 verts) trimmed to 124 triangles."""
    n = 62
    xs = np.linspace(-1, 1, n + 1)
    pos = np.concatenate([np.stack([xs, -0.1 * np.ones_like(xs), np.zeros_like(xs)], 1),
                          np.stack([xs, 0.1 * np.ones_like(xs), np.zeros_like(xs)], 1)]).astype(np.float32)
    tris = []
    for i in range(n):
        a, b, c, d = i, i + 1, n + 1 + i + 1, n + 1 + i
        tris += [(a, b, c), (a, c, d)]
    return pos, None, np.zeros((len(pos), 2), np.float32), np.array(tris[:124], np.uint32)


def fixtures():
    mat = Material(base_color=(0.7, 0.5, 0.3), roughness=0.6, metallic=0.0)
    out = {}
    cam, cfg = fixture_camera(64, 48, position=(0, 0, 3))
    out["one_triangle"] = (HostScene.build([triangle_mesh()], [mat], [(0, 0, np.eye(4, dtype=np.float32))]), cam, cfg)
    out["shared_edge"] = (HostScene.build([quad_mesh(0.8)], [mat], [(0, 0, np.eye(4, dtype=np.float32))]), cam, cfg)
    cam2, cfg2 = fixture_camera(96, 64, position=(2.5, 2.0, 3.5))
    out["closed_cube"] = (HostScene.build([cube_mesh(0.7)], [mat], [(0, 0, _rot_y(0.4))]), cam2, cfg2)
    # A 12 m floor quad (the xy quad rotated to face +y, scaled by 6) at y = -0.5
    # around the camera: its triangles cross the near plane.  Columns of m are
    # the images of the x, y, z axes and the translation.
    m = np.array([[6, 0, 0, 0], [0, 0, -6, 0], [0, 1, 0, 0], [0, -0.5, 0, 1]], np.float32)
    cam3, cfg3 = fixture_camera(80, 60, position=(0, 0, 0.0), target=(0, -0.2, -1), near=0.1)
    floor_quad = quad_mesh(1.0)
    out["near_plane_floor"] = (HostScene.build([floor_quad], [mat], [(0, 0, m)]), cam3, cfg3)
    cam4, cfg4 = fixture_camera(64, 64, position=(0.3, 0.2, 2.2))
    out["grid_64v_98t"] = (HostScene.build([grid_patch(7, 7)], [mat], [(0, 0, np.eye(4, dtype=np.float32))]), cam4,
                           cfg4)
    cam5, cfg5 = fixture_camera(128, 32, position=(0, 0, 1.6))
    out["strip_124t"] = (HostScene.build([_strip_124()], [mat], [(0, 0, np.eye(4, dtype=np.float32))]), cam5, cfg5)
    return out
