"""Shared helpers for the test suite (hand-built fixture scenes, host views)."""
from __future__ import annotations

import numpy as np

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, Material, make_camera, make_uniforms

IDENTITY = np.eye(4, dtype=np.float32)


def translate(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)  # column-major: column 3 = translation
    return m


def quad_mesh(size=1.0):
    """Two triangles sharing an edge: a square in the xy plane facing +z (CCW)."""
    s = size
    pos = np.array([[-s, -s, 0], [s, -s, 0], [s, s, 0], [-s, s, 0]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return pos, nrm, uv, idx


def triangle_mesh():
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
    return pos, nrm, np.zeros((3, 2), np.float32), np.array([[0, 1, 2]], np.uint32)


def cube_mesh(half=1.0):
    """Closed cube, 8 shared vertices, 12 CCW-outward triangles."""
    h = half
    pos = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)], np.float32)
    # vertex id = 4*ix + 2*iy + iz
    faces = [
        (0, 1, 3, 2),  # -x
        (4, 6, 7, 5),  # +x
        (0, 4, 5, 1),  # -y
        (2, 3, 7, 6),  # +y
        (0, 2, 6, 4),  # -z
        (1, 5, 7, 3),  # +z
    ]
    tris = []
    for a, b, c, d in faces:
        tris += [(a, b, c), (a, c, d)]
    idx = np.array(tris, np.uint32)
    # fix winding to outward
    out = []
    for t in idx:
        p0, p1, p2 = pos[t]
        n = np.cross(p1 - p0, p2 - p0)
        if np.dot(n, (p0 + p1 + p2) / 3) < 0:
            t = t[[0, 2, 1]]
        out.append(t)
    return pos, None, np.zeros((8, 2), np.float32), np.array(out, np.uint32)


def grid_patch(nx, ny):
    """(nx+1) x (ny+1) vertex grid in the xy plane facing +z."""
    xs, ys = np.meshgrid(np.linspace(-1, 1, nx + 1), np.linspace(-1, 1, ny + 1))
    pos = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], 1).astype(np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (len(pos), 1))
    uv = np.stack([xs.ravel() * 0.5 + 0.5, ys.ravel() * 0.5 + 0.5], 1).astype(np.float32)
    tris = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = j * (nx + 1) + i, j * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i + 1, (j + 1) * (nx + 1) + i
            tris += [(a, b, c), (a, c, d)]
    return pos, nrm, uv, np.array(tris, np.uint32)


def single_scene(mesh, transforms, material=None, textures=()):
    material = material or Material(base_color=(0.8, 0.6, 0.4), roughness=0.5, metallic=0.0)
    return HostScene.build([mesh], [material], [(0, 0, t) for t in transforms], textures=textures)


def fixture_camera(width=64, height=48, position=(0, 0, 5), target=(0, 0, 0), near=0.1, shadow=64):
    cam = make_camera(position, target, width, height, near=near)
    return cam, ViewConfig(width, height, shadow_size=shadow, big_capacity=1 << 16, clip_capacity=1 << 12)


def uniforms_for(cam, scene):
    return make_uniforms(cam, scene.num_instances)


def screen_to_clip(pts, W, H, z=0.5, w=1.0):
    """Framebuffer pixel coordinates -> clip coordinates with the given w."""
    out = []
    for x, y in pts:
        xn, yn = 2.0 * x / W - 1.0, 2.0 * y / H - 1.0
        out.append([xn * w, yn * w, z * w, w])
    return np.array(out, np.float32)


def covered(buf, clear=abi.LH_VISBUFFER_CLEAR):
    return buf != np.uint64(clear) if buf.dtype == np.uint64 else buf != np.uint32(abi.LH_SHADOW_CLEAR)
