"""Pins the oracle's software-raster rules (no reference code exists for them:
the reference uses Vulkan's fixed-function rasterizer, src/pipelines.cpp:10-77).

Known answers from the Vulkan rules: pixel-centre sampling, top-left fill rule,
CCW front faces in y-down framebuffer coordinates, 8-bit subpixel snapping,
clipping to 0 <= z <= w, screen-linear depth; invariants: shared edges covered
exactly once, watertight fans, order independence of atomicMin.
"""
import numpy as np
import pytest

import oracle
from lighthugger_amd import abi

from .util import screen_to_clip

W, H = 32, 24
CLEAR = np.uint64(abi.LH_VISBUFFER_CLEAR)


def raster(clip, kind=0, tid=1, W=W, H=H, vis=None, depth=None):
    clip = np.ascontiguousarray(clip, np.float32).reshape(3, 4)
    if vis is None:
        vis = np.full(W * H, CLEAR, np.uint64)
    if depth is None:
        depth = np.full(W * H, abi.LH_SHADOW_CLEAR, np.uint32)
    n = oracle.lib().orc_raster_clip_triangle(clip.ctypes.data, W, H, kind, tid, vis.ctypes.data, depth.ctypes.data)
    return int(n), vis.reshape(H, W), depth.reshape(H, W)


def cover_mask(vis):
    return vis != CLEAR


def test_right_triangle_known_pixels():
    # Right angle at the top-left; CCW on screen (y down) = front facing.
    n, vis, _ = raster(screen_to_clip([(0, 0), (0, 8), (8, 0)], W, H))
    m = cover_mask(vis)
    ys, xs = np.nonzero(m)
    # Centres strictly inside satisfy x + y < 7; the hypotenuse (x + y == 7) is a
    # bottom-right edge and is excluded.
    expect = {(x, y) for x in range(8) for y in range(8) if x + y <= 6}
    assert set(zip(xs.tolist(), ys.tolist())) == expect
    assert n == len(expect) == 28


def test_complementary_triangle_takes_the_shared_edge():
    a = cover_mask(raster(screen_to_clip([(0, 0), (0, 8), (8, 0)], W, H))[1])
    b = cover_mask(raster(screen_to_clip([(8, 8), (8, 0), (0, 8)], W, H))[1])
    assert not (a & b).any(), "a pixel on the shared edge was covered twice"
    square = np.zeros((H, W), bool)
    square[:8, :8] = True
    assert np.array_equal(a | b, square), "the two halves must tile the square exactly"
    assert b.sum() == 36


def test_back_faces_culled_in_main_view_and_front_faces_in_shadow():
    ccw = screen_to_clip([(0, 0), (0, 8), (8, 0)], W, H)
    cw = ccw[[0, 2, 1]]
    assert raster(cw, kind=0)[0] == 0
    assert raster(ccw, kind=1)[0] == 0
    assert raster(cw, kind=1)[0] == 28


def test_left_edge_included_right_edge_excluded():
    # Rectangle x in [4.5, 9.5) covering centres of columns 4..8: left edge on the
    # centres of column 4 is included, right edge on column 9 excluded.
    quad = [(4.5, 0.5), (4.5, 6.5), (9.5, 6.5), (9.5, 0.5)]
    m1 = cover_mask(raster(screen_to_clip([quad[0], quad[1], quad[2]], W, H))[1])
    m2 = cover_mask(raster(screen_to_clip([quad[0], quad[2], quad[3]], W, H))[1])
    m = m1 | m2
    assert not (m1 & m2).any()
    cols = np.nonzero(m.any(0))[0]
    rows = np.nonzero(m.any(1))[0]
    assert cols.tolist() == [4, 5, 6, 7, 8]
    # top edge y = 0.5 (centres of row 0) included, bottom edge y = 6.5 excluded
    assert rows.tolist() == [0, 1, 2, 3, 4, 5]


def test_subpixel_snapping_rounds_to_nearest_256th():
    # A vertex 0.3/256 px away from a pixel centre snaps onto it.
    eps = 0.3 / 256
    a = cover_mask(raster(screen_to_clip([(2.5 + eps, 2.5), (2.5, 8.5), (8.5, 2.5)], W, H))[1])
    b = cover_mask(raster(screen_to_clip([(2.5, 2.5), (2.5, 8.5), (8.5, 2.5)], W, H))[1])
    assert np.array_equal(a, b)


def test_fan_is_watertight_and_exclusive():
    rng = np.random.default_rng(7)
    ang = np.sort(rng.uniform(0, 2 * np.pi, 11))
    r = rng.uniform(6, 10, 11)
    c = np.array([15.37, 11.61])
    ring = np.stack([c[0] + r * np.cos(ang), c[1] + r * np.sin(ang)], 1)
    count = np.zeros((H, W), int)
    for i in range(len(ring)):
        a, b = ring[i], ring[(i + 1) % len(ring)]
        tri = [tuple(c), tuple(b), tuple(a)]  # orientation chosen below
        clip = screen_to_clip(tri, W, H)
        n, vis, _ = raster(clip)
        if n == 0:  # wrong winding for this ordering: flip
            n, vis, _ = raster(clip[[0, 2, 1]])
        count += cover_mask(vis)
    assert count.max() == 1, "fan pieces overlap"
    # Every pixel centre strictly inside the star polygon is covered.
    ys, xs = np.mgrid[0:H, 0:W]
    px, py = xs + 0.5, ys + 0.5
    inside = np.ones((H, W), bool)
    for i in range(len(ring)):
        a, b = ring[i], ring[(i + 1) % len(ring)]
        cr = (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0])
        inside &= cr > 1e-3  # ring is CCW in math orientation (y down)
    assert (count[inside] == 1).all(), "hole inside a closed fan"


def test_order_independence_of_nearest_key():
    rng = np.random.default_rng(3)
    tris = []
    for k in range(12):
        p = rng.uniform(0, W, (3, 2))
        z = rng.uniform(0.1, 0.9)
        clip = screen_to_clip(p, W, H, z=z)
        if raster(clip)[0] == 0:
            clip = clip[[0, 2, 1]]
        tris.append(clip)
    v1 = np.full(W * H, CLEAR, np.uint64)
    v2 = np.full(W * H, CLEAR, np.uint64)
    for k, t in enumerate(tris):
        raster(t, tid=k + 1, vis=v1)
    for k in rng.permutation(len(tris)):
        raster(tris[k], tid=int(k) + 1, vis=v2)
    assert np.array_equal(v1, v2)


def test_constant_depth_is_exact_and_nearest_wins():
    z = np.float32(0.37)
    n, vis, _ = raster(screen_to_clip([(0, 0), (0, 8), (8, 0)], W, H, z=float(z)), tid=5)
    keys = vis[cover_mask(vis)]
    depth_bits = (~(keys >> np.uint64(32))).astype(np.uint32)
    assert (depth_bits == z.view(np.uint32)).all()
    assert ((keys & np.uint64(0xFFFFFFFF)) == 5).all()
    # A nearer triangle (reverse-Z: larger z) overwrites, a farther one does not.
    vis2 = vis.ravel().copy()
    raster(screen_to_clip([(0, 0), (0, 8), (8, 0)], W, H, z=0.2), tid=6, vis=vis2)
    assert ((vis2.reshape(H, W)[cover_mask(vis)] & np.uint64(0xFFFFFFFF)) == 5).all()
    raster(screen_to_clip([(0, 0), (0, 8), (8, 0)], W, H, z=0.6), tid=7, vis=vis2)
    assert ((vis2.reshape(H, W)[cover_mask(vis)] & np.uint64(0xFFFFFFFF)) == 7).all()


def test_depth_interpolates_screen_linearly():
    # z varies linearly in x: z = 0.1 + 0.02 * x_f at the vertices.
    pts = [(0, 0), (0, 16), (16, 0)]
    clip = screen_to_clip(pts, W, H)
    for i, (x, _) in enumerate(pts):
        clip[i, 2] = 0.1 + 0.02 * x
    _, vis, _ = raster(clip)
    m = cover_mask(vis)
    ys, xs = np.nonzero(m)
    z = (~(vis[m] >> np.uint64(32))).astype(np.uint32).view(np.float32)
    np.testing.assert_allclose(z, 0.1 + 0.02 * (xs + 0.5), rtol=0, atol=2e-6)


def test_near_plane_clipping_keeps_the_visible_part():
    # Reverse-Z infinite perspective clip coords: z_c = near, w_c = distance.
    near = 0.1
    # A wall triangle from distance 0.05 (behind the near plane) to 4.
    def clip_of(x, y, d):
        f = 1.0  # focal length (90 degree fov)
        return [x * f, -y * f, near, d]
    clip = np.array([clip_of(-1, -1, 0.05), clip_of(3, -1, 4.0), clip_of(-1, 3, 4.0)], np.float32)
    n1, vis, _ = raster(clip, W=W, H=H)
    if n1 == 0:
        n1, vis, _ = raster(clip[[0, 2, 1]])
    assert n1 > 0
    z = (~(vis[cover_mask(vis)] >> np.uint64(32))).astype(np.uint32).view(np.float32)
    assert (z <= 1.0).all() and (z > 0).all()
    # Entirely behind the near plane: nothing.
    behind = np.array([clip_of(-1, -1, 0.05), clip_of(1, -1, 0.05), clip_of(0, 1, 0.06)], np.float32)
    assert raster(behind)[0] == 0 and raster(behind[[0, 2, 1]])[0] == 0


def test_guard_band_full_screen_quad():
    # Two huge triangles (far outside the guard band) cover every pixel exactly once.
    big = [(-500, -500), (-500, 500), (500, 500), (500, -500)]
    a = screen_to_clip([big[0], big[1], big[2]], W, H)
    b = screen_to_clip([big[0], big[2], big[3]], W, H)
    cnt = np.zeros((H, W), int)
    for t in (a, b):
        n, vis, _ = raster(t)
        if n == 0:
            n, vis, _ = raster(t[[0, 2, 1]])
        cnt += cover_mask(vis)
    assert (cnt == 1).all()


def test_shadow_depth_and_nonfinite_rejection():
    clip = screen_to_clip([(0, 0), (8, 0), (0, 8)], W, H, z=0.25)  # CW on screen: back face, kept by shadows
    n, _, d = raster(clip, kind=1)
    assert n == 28
    assert (d[d != abi.LH_SHADOW_CLEAR] == np.float32(0.25).view(np.uint32)).all()
    bad = clip.copy()
    bad[1, 0] = np.nan
    assert raster(bad, kind=1)[0] == 0
    bad[1, 0] = np.inf
    assert raster(bad, kind=1)[0] == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_triangles_match_brute_force_with_float64(seed):
    """Coverage equals a float64 brute force with the same snapping and fill rule
    (independent restatement in numpy)."""
    rng = np.random.default_rng(seed)
    for _ in range(20):
        p = rng.uniform(-4, W + 4, (3, 2))
        clip = screen_to_clip(p, W, H)
        n, vis, _ = raster(clip)
        if n == 0:
            clip = clip[[0, 2, 1]]
            p = p[[0, 2, 1]]
            n, vis, _ = raster(clip)
        # brute force in float64 on the snapped vertices
        q = np.rint(((clip[:, :2] / clip[:, 3:4]) * 0.5 + 0.5) * np.array([W, H]) * 256).astype(np.int64)
        area = (q[1, 0] - q[0, 0]) * (q[2, 1] - q[0, 1]) - (q[1, 1] - q[0, 1]) * (q[2, 0] - q[0, 0])
        if area >= 0:
            assert n == 0
            continue
        q = q[[0, 2, 1]]
        ys, xs = np.mgrid[0:H, 0:W]
        P = np.stack([xs * 256 + 128, ys * 256 + 128], -1).astype(np.int64)
        ok = np.ones((H, W), bool)
        for a, b in ((1, 2), (2, 0), (0, 1)):
            A, B = q[a], q[b]
            e = (B[0] - A[0]) * (P[..., 1] - A[1]) - (B[1] - A[1]) * (P[..., 0] - A[0])
            dx, dy = B[0] - A[0], B[1] - A[1]
            tl = dy < 0 or (dy == 0 and dx > 0)
            ok &= (e > 0) | ((e == 0) & tl)
        assert np.array_equal(ok, cover_mask(vis))
