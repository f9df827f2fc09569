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


def alpha_texture(n=32, seed=1, cell=4):
    """RGBA8 texture whose alpha is 0 or 255 in cell x cell blocks (noise RGB)."""
    rng = np.random.default_rng(seed)
    cells = rng.random((n // cell, n // cell)) < 0.5
    a = (np.kron(cells, np.ones((cell, cell))) * 255).astype(np.uint8)
    rgb = rng.integers(64, 256, (n, n, 3)).astype(np.uint8)
    return np.dstack([rgb, a])


def _card(scale, yaw, pitch, t):
    """Quad transform: scale, then pitch about x, yaw about y, then translate."""
    cy, sy, cp, sp = np.cos(yaw), np.sin(yaw), np.cos(pitch), np.sin(pitch)
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    m = np.eye(4)
    m[:3, :3] = ry @ rx * scale
    m[:3, 3] = t
    return m.T.astype(np.float32)  # column-major


def alpha_scene(width=96, height=64, shadow=128):
    """Alpha-clipped cards (double-sided, base-colour alpha < 0.5 discarded)
    over an opaque ground: cards facing towards and away from the camera, one
    close enough for the large-triangle path, one through the near plane, one
    alpha-clipped card without a texture (alpha 1), all casting shadows."""
    ground = Material(base_color=(0.6, 0.6, 0.6), roughness=0.8, metallic=0.0)
    card = Material(base_color=(0.9, 0.8, 0.7), roughness=0.5, metallic=0.0, base_color_texture=0, alpha_clip=True)
    plain = Material(base_color=(0.3, 0.7, 0.4), roughness=0.4, metallic=0.0, alpha_clip=True)
    rng = np.random.default_rng(7)
    inst = [(0, 0, _card(6.0, 0.0, -np.pi / 2, (0, -1.0, 0)))]
    for _ in range(10):
        t = (rng.uniform(-2.5, 2.5), rng.uniform(-0.6, 1.2), rng.uniform(-2.5, 1.0))
        inst.append((1, 1, _card(rng.uniform(0.4, 1.0), rng.uniform(0, 2 * np.pi), rng.uniform(-0.5, 0.5), t)))
    inst.append((1, 1, _card(1.2, 0.3, 0.0, (0.4, 0.3, 2.6))))           # near the camera: large triangles
    inst.append((1, 1, _card(1.5, np.pi / 2, 0.0, (0.3, 0.0, 3.5))))     # along the view axis, through the near plane
    inst.append((1, 2, _card(0.5, 2.5, 0.2, (-1.2, 0.2, 0.5))))          # no texture: alpha 1
    scene = HostScene.build([quad_mesh(1.0), grid_patch(6, 6)], [ground, card, plain], inst,
                            textures=[alpha_texture()])
    cam, cfg = fixture_camera(width, height, position=(0.2, 0.4, 4.0), target=(0, 0, 0), near=0.1, shadow=shadow)
    return scene, cam, cfg
