"""Scene producer invariants (C++ host, include/lh/lh_scene.h).

Meshlets: <= 64 vertices, <= 124 triangles, u8 micro indices in range,
triangle blocks 4-byte aligned (src/resources/meshlets.cpp:62-64), every source
triangle exactly once.  Bounds: spheres contain their vertices; the stable
meshlet table inverts the instance-order scan; the normal matrix is
transpose(inverse(M)); uniforms follow src/projection.h:3-20.
meshoptimizer's exact meshlet grouping is not reproduced (parity unpinned:
the library is absent); the kernels and the oracle consume the same meshlets.
"""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.scene import HostScene, Material, make_camera, make_uniforms, preset_camera

from .util import cube_mesh, grid_patch, single_scene, translate


@pytest.fixture(scope="module")
def small():
    return HostScene.preset(abi.PRESET_SMALL)


def _mesh_arrays(scene, mi, nverts):
    ml = scene.meshlets(mi)
    last = ml[-1]
    remap = scene.remap(mi, int(last["index_offset"]) + int(last["index_count"]))
    micro = scene.micro(mi, int(last["triangle_offset"]) + 3 * int(last["triangle_count"]))
    pos = scene.positions(mi, nverts)
    return ml, remap, micro, pos


def test_small_preset_shape(small):
    assert small.num_instances == 64
    assert small.num_meshes == 1
    assert small.unique_triangles == 8192
    assert small.total_triangles == 64 * 8192
    assert len(small.mesh_infos) == 4  # one sphere x 4 materials


def _check_meshlets(scene, mi, nverts, src=None):
    ml, remap, micro, pos = _mesh_arrays(scene, mi, nverts)
    assert (ml["index_count"] <= 64).all() and (ml["triangle_count"] <= 124).all()
    assert (ml["triangle_count"] > 0).all()
    assert (ml["triangle_offset"] % 4 == 0).all()
    tris = []
    for m in ml:
        t0, nt, i0, nv = int(m["triangle_offset"]), int(m["triangle_count"]), int(m["index_offset"]), int(m["index_count"])
        c = micro[t0: t0 + 3 * nt].reshape(nt, 3)
        assert c.max() < nv
        v = remap[i0 + c.astype(np.int64)]
        tris.append(v)
        p = pos[remap[i0: i0 + nv]]
        cs, r = m["bounding_sphere"][:3], m["bounding_sphere"][3]
        d = np.linalg.norm(p[np.unique(c)] - cs, axis=1)
        assert (d <= r * (1 + 1e-5) + 1e-6).all(), "meshlet sphere must contain its vertices"
    tris = np.concatenate(tris)
    if src is not None:
        # compared as position triples: the producer renumbers vertices
        # (meshlet first-use order) but keeps every triangle exactly once
        src_pos, src_idx = np.asarray(src[0], np.float32), np.asarray(src[1])

        def key(P, idx):
            t = P[idx.astype(np.int64)].reshape(len(idx), 9)
            t = t.reshape(-1, 3, 3)
            t = np.sort(t.view([("x", "f4"), ("y", "f4"), ("z", "f4")]).reshape(-1, 3), axis=1)
            t = t.view(np.float32).reshape(-1, 9)
            return t[np.lexsort(t.T[::-1])]

        assert np.array_equal(key(pos, tris), key(src_pos, src_idx)), \
            "meshlets must hold every source triangle exactly once"
    sph = scene.mesh_infos[mi]["bounding_sphere"]
    d = np.linalg.norm(pos - sph[:3], axis=1)
    assert (d <= sph[3] * (1 + 1e-5) + 1e-6).all()
    return ml


def test_small_meshlets(small):
    ml = _check_meshlets(small, 0, 4098)
    assert ml["triangle_count"].mean() > 80


def test_grid_and_cube_meshlets_cover_every_triangle():
    for mesh in (grid_patch(20, 17), cube_mesh()):
        s = single_scene(mesh, [np.eye(4, dtype=np.float32)])
        _check_meshlets(s, 0, len(mesh[0]), (mesh[0], mesh[3]))


def test_mesh_at_exactly_64_vertices_and_124_triangles():
    # A fan of 62 rim vertices + centre... build a strip with 64 vertices and
    # <= 124 triangles: 2 x 32 vertex strip has 62 triangles; two stacked rows
    # give 3 x 21 = 63 vertices and 80 triangles.  Use a 8 x 8 grid: 64 verts, 98 tris.
    mesh = grid_patch(7, 7)
    assert len(mesh[0]) == 64 and len(mesh[3]) == 98
    s = single_scene(mesh, [np.eye(4, dtype=np.float32)])
    ml = _check_meshlets(s, 0, 64, (mesh[0], mesh[3]))
    assert len(ml) == 1 and ml[0]["index_count"] == 64 and ml[0]["triangle_count"] == 98


def test_stable_meshlet_table(small):
    base = small.instance_meshlet_base
    nm = small.mesh_infos["num_meshlets"][small.instances["mesh_info_address"].astype(np.int64)].astype(np.int64)
    assert np.array_equal(base.astype(np.int64), np.concatenate([[0], np.cumsum(nm)[:-1]]))
    t = small.meshlet_table
    assert len(t) == nm.sum()
    for i in (0, 17, 63):
        seg = t[base[i]: base[i] + nm[i]]
        assert (seg["instance_index"] == i).all()
        assert np.array_equal(seg["meshlet_index"], np.arange(nm[i]))


def test_normal_matrix_is_inverse_transpose(small):
    for k in range(0, 64, 9):
        m = np.asarray(small.instances["transform"][k], np.float64).T  # row-major
        n = np.asarray(small.instances["normal_transform"][k], np.float64).T
        np.testing.assert_allclose(n, np.linalg.inv(m[:3, :3]).T, rtol=1e-5, atol=1e-6)


def test_uniforms_follow_infinite_reverse_z():
    cam = make_camera((1, 2, 3), (0, 0, 0), 1280, 720, fov_deg=60, near=0.1)
    u = make_uniforms(cam, 7)[0]
    P = np.asarray(u["perspective"], np.float64).T
    f = 1 / np.tan(np.radians(30))
    np.testing.assert_allclose(P, [[f / (1280 / 720), 0, 0, 0], [0, -f, 0, 0], [0, 0, 0, 0.1], [0, 0, -1, 0]],
                               rtol=1e-6, atol=1e-7)
    V = np.asarray(u["view"], np.float64).T
    np.testing.assert_allclose(V @ [1, 2, 3, 1], [0, 0, 0, 1], atol=1e-5)
    # reverse-Z: a point at distance d has depth near / d
    fwd = -np.array([1, 2, 3]) / np.linalg.norm([1, 2, 3])
    q = np.asarray(u["combined_perspective_view"], np.float64).T @ np.append(np.array([1, 2, 3]) + 10 * fwd, 1)
    assert abs(q[2] / q[3] - 0.01) < 1e-6
    assert u["num_instances"] == 7 and tuple(u["window_size"]) == (1280, 720)
    inv = np.asarray(u["inv_perspective_view"], np.float64).T
    np.testing.assert_allclose(inv @ np.asarray(u["combined_perspective_view"], np.float64).T, np.eye(4), atol=1e-4)


def test_large_preset_statistics():
    s = HostScene.preset(abi.PRESET_LARGE, texture_size=16, tri_scale_pct=5)
    assert s.num_instances == 2000 and s.num_meshes == 50
    assert len(s.textures) == 64
    assert (s.textures["levels"] == 5).all()
    tri = s.instance_triangles()
    assert tri.sum() == s.total_triangles


def test_preset_cameras():
    for p, (w, h) in ((abi.PRESET_SMALL, (1280, 720)), (abi.PRESET_LARGE, (1920, 1080)), (abi.PRESET_STRESS, (3840, 2160))):
        c = preset_camera(p)
        assert (c.width, c.height) == (w, h)
        assert abs(c.near_plane - 0.1) < 1e-7


def test_invalid_inputs_rejected():
    pos = np.zeros((3, 3), np.float32)
    with pytest.raises(ValueError):
        HostScene.build([(pos, None, None, np.array([[0, 1, 5]], np.uint32))], [], [])


def _world(scene, inst, local):
    m = np.asarray(scene.instances["transform"][inst], np.float64).T  # row-major
    return local.astype(np.float64) @ m[:3, :3].T + m[:3, 3]


@pytest.mark.parametrize("preset", [abi.PRESET_SMALL, abi.PRESET_LARGE])
def test_quantized_vertices_dequantize_to_the_fp32_scene(preset):
    """LH_VERTEX_FORMAT_QUANTIZED (the reference's u16 / i8 / u16 formats):
    the meshlets and vertex order come from connectivity alone, so they equal
    the FP32 scene's; world positions through the folded instance transform
    are within half a grid step of the FP32 ones, normals within half an i8
    step, texture coordinates (through the folded texture transform) within
    half a u16 step; meshes up to 65,536 vertices use u16 remap indices."""
    kw = dict(tri_scale_pct=5, texture_size=16) if preset == abi.PRESET_LARGE else {}
    a = HostScene.preset(preset, **kw)
    b = HostScene.preset(preset, vertex_format=abi.VERTEX_FORMAT_QUANTIZED, **kw)
    assert b.geometry.nbytes < a.geometry.nbytes
    assert not (b.mesh_infos["flags"] & abi.MESH_INFO_FLAGS_FP32_VERTICES).any()
    inst_of = {int(k): i for i, k in reversed(list(enumerate(a.instances["mesh_info_address"])))}
    for mi in range(len(a.mesh_infos)):
        ma, mb = a.meshlets(mi), b.meshlets(mi)
        for f in ("index_offset", "index_count", "triangle_offset", "triangle_count"):
            assert np.array_equal(ma[f], mb[f]), f
        n_idx = int(ma["index_offset"][-1]) + int(ma["index_count"][-1])
        ra, rb = a.remap(mi, n_idx), b.remap(mi, n_idx)
        assert np.array_equal(ra, rb)
        nv = int(ra.max()) + 1
        assert bool(b.mesh_infos["flags"][mi] & abi.MESH_INFO_FLAGS_32_BIT_INDICES) == (nv > 65536)
        q = b.positions(mi, nv)
        assert q.min() >= 0 and q.max() <= 65535 and np.array_equal(q, np.round(q))
        # the grid: origin at the mesh's minimum corner, one step for all axes
        pa = a.positions(mi, nv).astype(np.float64)
        lo = pa.min(axis=0)
        grid = (pa.max(axis=0) - lo).max() / 65535.0
        assert np.abs(q - np.round((pa - lo) / grid)).max() <= 1.0
        i = inst_of[mi]
        ma_, mb_ = (np.asarray(x.instances["transform"][i], np.float64) for x in (a, b))
        step = np.linalg.norm(mb_[0, :3]) / np.linalg.norm(ma_[0, :3])  # column 0 = M * step
        assert abs(step - grid) <= 1e-5 * grid
        # world positions: half a step per axis, plus the FP32 rounding of both translations
        err = np.abs(_world(a, i, a.positions(mi, nv)) - _world(b, i, q)).max()
        fp = 2.0 ** -23 * (np.abs(ma_[3, :3]).max() + np.abs(mb_[3, :3]).max() + 1.0)
        assert err <= 0.5 * step * np.linalg.norm(ma_[0, :3]) * np.sqrt(3) * (1 + 1e-3) + 2 * fp, (mi, err, step)
        na = a.normals(mi, nv)
        assert np.abs(b.normals(mi, nv) / 127.0 - na).max() <= 0.5 / 127 + 1e-6
        # uv * scale + offset, as the resolve applies it
        ua = a.uvs(mi, nv) * a.mesh_infos["texture_scale"][mi] + a.mesh_infos["texture_offset"][mi]
        ub = b.uvs(mi, nv).astype(np.float64) * b.mesh_infos["texture_scale"][mi] + b.mesh_infos["texture_offset"][mi]
        half = 0.5 * np.asarray(b.mesh_infos["texture_scale"][mi], np.float64)
        assert (np.abs(ua - ub) <= half * (1 + 1e-3) + 1e-5 * (1 + np.abs(ua))).all(), mi
        # normal matrix of the folded transform
        m = np.asarray(b.instances["transform"][i], np.float64).T
        n = np.asarray(b.instances["normal_transform"][i], np.float64).T
        np.testing.assert_allclose(n, np.linalg.inv(m[:3, :3]).T, rtol=1e-4)


def test_vertex_format_rejected_after_first_mesh():
    from lighthugger_amd.scene import scene_lib

    lib = scene_lib()
    h = lib.lh_scene_create()
    try:
        assert lib.lh_scene_set_vertex_format(h, 2) != 0
        assert lib.lh_scene_set_vertex_format(h, abi.VERTEX_FORMAT_QUANTIZED) == 0
        pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
        idx = np.array([[0, 1, 2]], np.uint32)
        assert lib.lh_scene_add_mesh(h, pos.ctypes.data, None, None, 3, idx.ctypes.data, 1) == 0
        assert lib.lh_scene_set_vertex_format(h, abi.VERTEX_FORMAT_FP32) != 0
    finally:
        lib.lh_scene_destroy(h)


@pytest.mark.parametrize("vertex_format", [abi.VERTEX_FORMAT_FP32, abi.VERTEX_FORMAT_QUANTIZED])
def test_built_mesh_cache(tmp_path, vertex_format):
    """The on-disk cache of built meshes (the reference's FsCache): a second
    build reads every mesh back and gives the same scene; a damaged file is
    ignored and rebuilt; the vertex format is part of the key."""
    kw = dict(tri_scale_pct=5, texture_size=16, vertex_format=vertex_format)
    plain = HostScene.preset(abi.PRESET_LARGE, **kw)
    first = HostScene.preset(abi.PRESET_LARGE, cache_dir=tmp_path, **kw)
    files = sorted(tmp_path.glob("lhmesh-*.bin"))
    # content-addressed: meshes with identical sources (some boxes) share a file
    assert 0.8 * plain.num_meshes <= len(files) <= plain.num_meshes
    second = HostScene.preset(abi.PRESET_LARGE, cache_dir=tmp_path, **kw)
    for s in (first, second):
        for k in ("geometry", "mesh_infos", "instances", "meshlet_table", "texels"):
            assert np.array_equal(getattr(s, k), getattr(plain, k)), k
    # damage two files: truncated, and a remap entry out of range
    files[0].write_bytes(files[0].read_bytes()[:100])
    raw = bytearray(files[1].read_bytes())
    raw[-30:-20] = b"\xff" * 10
    files[1].write_bytes(bytes(raw))
    third = HostScene.preset(abi.PRESET_LARGE, cache_dir=tmp_path, **kw)
    assert np.array_equal(third.geometry, plain.geometry) and np.array_equal(third.mesh_infos, plain.mesh_infos)
    other = abi.VERTEX_FORMAT_QUANTIZED if vertex_format == abi.VERTEX_FORMAT_FP32 else abi.VERTEX_FORMAT_FP32
    HostScene.preset(abi.PRESET_LARGE, cache_dir=tmp_path, **{**kw, "vertex_format": other})
    assert len(list(tmp_path.glob("lhmesh-*.bin"))) == 2 * len(files)


def _fnv1a64(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("damage", ["micro_index", "triangle_count", "vertex_count"])
def test_built_mesh_cache_rejects_meshlets_beyond_the_raster_limits(tmp_path, damage):
    """A cache file that passes its checksum but names a micro index at or
    beyond its meshlet's vertex count, or a meshlet over 124 triangles / 64
    vertices, is ignored and the mesh rebuilt (the raster assumes those
    limits, the resolve indexes the remap table with the micro indices)."""
    import struct

    mesh = [grid_patch(12, 12)]
    mat = [Material(base_color=(0.5, 0.5, 0.5))]
    inst = [(0, 0, np.eye(4, dtype=np.float32))]
    plain = HostScene.build(mesh, mat, inst)
    HostScene.build(mesh, mat, inst, cache_dir=tmp_path)
    (path,) = tmp_path.glob("lhmesh-*.bin")
    raw = bytearray(path.read_bytes())
    off = 8  # magic, version
    sizes = {}
    for name, item in (("pos", 4), ("nrm", 4), ("uv", 4), ("idx", 4), ("meshlets", 56), ("remap", 4), ("micro", 1)):
        (n,) = struct.unpack_from("<Q", raw, off)
        sizes[name] = (off + 8, n)
        off += 8 + n * item
    ml0 = sizes["meshlets"][0]
    tri_off, _idx_off, ntri, nvert = struct.unpack_from("<IIBB", raw, ml0 + 44)
    if damage == "micro_index":
        raw[sizes["micro"][0] + tri_off + 1] = nvert  # one past the meshlet's vertices
    elif damage == "triangle_count":
        raw[ml0 + 52] = 125
    else:
        raw[ml0 + 53] = 65
    raw[-8:] = struct.pack("<Q", _fnv1a64(bytes(raw[:-8])))
    path.write_bytes(bytes(raw))
    again = HostScene.build(mesh, mat, inst, cache_dir=tmp_path)
    for k in ("geometry", "mesh_infos", "meshlet_table"):
        assert np.array_equal(getattr(again, k), getattr(plain, k)), k
