"""Texture formats beyond RGBA8 / BC1: BC7 blocks, the sRGB decode of the
*_SRGB formats, the producer's BC7 encoder, and the DDS / KTX2 / glTF paths
that bring stored textures in (src/resources/image_loading.cpp:14-39,
:242-360, src/resources/mesh_loading.cpp:323-352).

Pins:
  - sRGB EOTF: f(c) = c / 12.92 for c <= 0.04045, ((c + 0.055) / 1.055)^2.4
    above, on all 256 codes against float64, f(0) = 0, f(1) = 1, continuity
    at the knee, inverse of the OETF the display transform uses;
  - BC7: hand-worked blocks of mode 6 (p-bits, 4-bit weights) and mode 5
    (rotation, separate alpha indices), and every mode and partition
    against Pillow's BCn decoder (a third-party implementation of the
    published format) on random blocks; the reserved mode is transparent
    black (Pillow returns opaque black there, so it is checked by hand);
  - stored textures: DDS / KTX2 (zstd and uncompressed) written by the test
    read back byte for byte, with the reference's format mapping; a frame
    whose base colour comes from a BC7 sRGB KTX2 file equals, bit for bit,
    the frame with that file's texels decoded by Pillow and stored as sRGB
    RGBA8.
"""
from __future__ import annotations

import io
import json
import struct

import numpy as np
import pytest

import oracle
from lighthugger_amd import abi, dds, ktx2
from lighthugger_amd.gltf import load_gltf
from lighthugger_amd.scene import EncodedTexture, HostScene, Material, make_uniforms
from oracle.compare import oracle_frame

from .util import fixture_camera, grid_patch, quad_mesh

PIL = pytest.importorskip("PIL.Image")


def _bits(fields) -> np.ndarray:
    """16-byte block from (value, width) fields, LSB first."""
    v, pos = 0, 0
    for val, n in fields:
        assert 0 <= val < (1 << n)
        v |= val << pos
        pos += n
    assert pos == 128
    return np.frombuffer(v.to_bytes(16, "little"), np.uint8).copy()


def _bc7(block) -> np.ndarray:
    out = np.zeros(64, np.uint8)
    oracle.lib().orc_bc7_block(np.ascontiguousarray(block).ctypes.data, out.ctypes.data)
    return out.reshape(16, 4)


def _dds_bytes(blocks: bytes, w: int, h: int, dxgi: int) -> bytes:
    hdr = bytearray(124)
    struct.pack_into("<7I", hdr, 0, 124, 0x1007, h, w, 0, 1, 1)
    struct.pack_into("<2I4s", hdr, 72, 32, 0x4, b"DX10")
    struct.pack_into("<I", hdr, 104, 0x1000)
    return b"DDS " + bytes(hdr) + struct.pack("<5I", dxgi, 3, 0, 1, 0) + blocks


def _pillow_decode(blocks: np.ndarray, w: int, h: int) -> np.ndarray:
    im = PIL.open(io.BytesIO(_dds_bytes(blocks.tobytes(), w, h, dds.BC7_UNORM)))
    return np.asarray(im.convert("RGBA"))


# ---------------------------------------------------------------- sRGB
def test_srgb_eotf_closed_form():
    L = oracle.lib()
    got = np.array([L.orc_srgb_to_linear(v) for v in range(256)], np.float64)
    c = np.arange(256) / 255.0
    want = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    assert np.abs(got - want).max() < 2e-7
    assert got[0] == 0.0 and got[255] == 1.0
    assert np.all(np.diff(got) > 0)
    knee = 0.04045
    assert abs(knee / 12.92 - ((knee + 0.055) / 1.055) ** 2.4) < 1e-7  # the two branches meet
    oetf = np.array([L.orc_srgb_oetf(float(x)) for x in got])
    assert np.abs(oetf - c).max() < 1e-5  # the display transform's OETF inverts it


# ---------------------------------------------------------------- BC7
def test_bc7_mode6_hand_worked():
    # mode 6: R 127/0, G 0/127, B 64/64, A 127/127 (7 bits), p-bits 1 and 0
    idx = [0] * 16
    idx[5], idx[10] = 15, 8
    block = _bits([(1 << 6, 7), (127, 7), (0, 7), (0, 7), (127, 7), (64, 7), (64, 7), (127, 7), (127, 7), (1, 1),
                   (0, 1), (idx[0], 3)] + [(i, 4) for i in idx[1:]])
    t = _bc7(block)
    # endpoints (v << 1 | p): e0 = (255, 1, 129, 255), e1 = (0, 254, 128, 254)
    assert tuple(t[0]) == (255, 1, 129, 255)
    assert tuple(t[5]) == (0, 254, 128, 254)
    # weight 34: ((64 - 34) e0 + 34 e1 + 32) >> 6
    assert tuple(t[10]) == (120, 135, 128, 254)


def test_bc7_mode5_rotation_hand_worked():
    # mode 5, rotation 1 (alpha <-> red): R 127/0 (7 bits -> 255/0), G, B 0, A 10/200 (8 bits)
    cidx, aidx = [0] * 16, [0] * 16
    cidx[3], aidx[3] = 3, 3
    block = _bits([(1 << 5, 6), (1, 2), (127, 7), (0, 7), (0, 7), (0, 7), (0, 7), (0, 7), (10, 8), (200, 8),
                   (cidx[0], 1)] + [(i, 2) for i in cidx[1:]] + [(aidx[0], 1)] + [(i, 2) for i in aidx[1:]])
    t = _bc7(block)
    assert tuple(t[0]) == (10, 0, 0, 255)   # colour 255, alpha 10, swapped
    assert tuple(t[3]) == (200, 0, 0, 0)    # colour 0, alpha 200, swapped


def test_bc7_reserved_mode_is_transparent_black():
    b = np.full(16, 0xA5, np.uint8)
    b[0] = 0
    assert not _bc7(b).any()


def test_bc7_every_mode_and_partition_matches_pillow():
    rng = np.random.default_rng(1)
    n = 64 * 64
    blocks = rng.integers(0, 256, (n, 16), dtype=np.uint8)
    for k in range(n):
        m = k % 8
        blocks[k, 0] = (int(blocks[k, 0]) & (0xFF & ~((1 << (m + 1)) - 1))) | (1 << m)
    img = _pillow_decode(blocks, 256, 256)
    for k in range(n):
        by, bx = divmod(k, 64)
        want = img[by * 4: by * 4 + 4, bx * 4: bx * 4 + 4].reshape(16, 4)
        assert np.array_equal(_bc7(blocks[k]), want), f"block {k} (mode {k % 8})"


def test_producer_bc7_encoder_round_trip():
    """Every level of a two-colour tinted noise texture (the kind the presets
    make: every channel a blend of two colours by one noise value) encoded by
    the producer (mode 6, one line through RGBA per block) decodes within a
    few steps of the source."""
    rng = np.random.default_rng(4)
    noise = rng.random((32, 32))
    noise = (noise + np.roll(noise, 1, 0) + np.roll(noise, 1, 1) + np.roll(noise, 1, (0, 1))) / 4
    ca, cb = np.array([30, 200, 90, 255.0]), np.array([220, 40, 160, 120.0])
    tex = np.round(ca + (cb - ca) * noise[..., None]).astype(np.uint8)
    plain = HostScene.build([quad_mesh()], [Material(base_color_texture=0)], [(0, 0, np.eye(4, dtype=np.float32))],
                            textures=[tex])
    bc7 = HostScene.build([quad_mesh()], [Material(base_color_texture=0)], [(0, 0, np.eye(4, dtype=np.float32))],
                          textures=[tex], texture_format=abi.TEXTURE_BC7)
    t, p = bc7.textures[0], plain.textures[0]
    assert int(t["format"]) == abi.TEXTURE_BC7 and int(t["levels"]) == int(p["levels"])
    for lv in range(int(t["levels"])):
        w, h = max(1, 32 >> lv), max(1, 32 >> lv)
        src = plain.texels[int(p["texels"]) + 4 * int(p["level_offset"][lv]):][: w * h * 4].reshape(h, w, 4)
        bw, bh = (w + 3) // 4, (h + 3) // 4
        base = int(t["texels"]) + 16 * int(t["level_offset"][lv])
        blocks = bc7.texels[base: base + 16 * bw * bh].reshape(bh * bw, 16)
        dec = np.zeros((bh * 4, bw * 4, 4), np.uint8)
        for k in range(bw * bh):
            by, bx = divmod(k, bw)
            dec[by * 4: by * 4 + 4, bx * 4: bx * 4 + 4] = _bc7(blocks[k]).reshape(4, 4, 4)
            assert blocks[k][0] & 0x7F == 0x40  # mode 6 (bit 7 starts the red endpoint)
        err = np.abs(dec[:h, :w].astype(np.int32) - src.astype(np.int32))
        assert err.mean() < 2.5 and err.max() < 12, (lv, err.mean(), err.max())


# ---------------------------------------------------------------- stored textures
def _bc7_chain(w, h, seed=3):
    rng = np.random.default_rng(seed)
    levels, lw, lh = [], w, h
    while True:
        n = ((lw + 3) // 4) * ((lh + 3) // 4)
        b = rng.integers(0, 256, (n, 16), dtype=np.uint8)
        b[:, 0] = 0x40  # mode 6 blocks (any mode would do)
        levels.append(b.tobytes())
        if lw == 1 and lh == 1:
            return levels
        lw, lh = max(1, lw // 2), max(1, lh // 2)


def test_dds_bc7_and_format_mapping(tmp_path):
    levels = _bc7_chain(16, 8)
    for fmt, want in ((dds.BC7_UNORM, abi.TEXTURE_BC7 | abi.TEXTURE_SRGB),
                      (dds.BC7_UNORM_SRGB, abi.TEXTURE_BC7 | abi.TEXTURE_SRGB)):
        p = tmp_path / f"t{fmt}.dds"
        hdr = bytearray(124)
        struct.pack_into("<7I", hdr, 0, 124, 0x21007, 8, 16, 0, 1, len(levels))
        struct.pack_into("<2I4s", hdr, 72, 32, 0x4, b"DX10")
        p.write_bytes(b"DDS " + bytes(hdr) + struct.pack("<5I", fmt, 3, 0, 1, 0) + b"".join(levels))
        t = dds.load_dds_texture(p)
        assert t.format == want and (t.width, t.height) == (16, 8)
        assert [bytes(x) for x in t.levels] == levels
        with pytest.raises(ValueError):
            dds.read_dds(p)  # BC7 is decoded on the GPU, not on the host
    # BC1 (DX10 UNORM, and the legacy DXT1 FourCC) maps to sRGB with alpha read as 1
    assert dds.TEXTURE_FORMAT[dds.BC1_UNORM] == abi.TEXTURE_BC1 | abi.TEXTURE_SRGB | abi.TEXTURE_RGB
    assert dds.TEXTURE_FORMAT[dds.R8G8B8A8_UNORM] == abi.TEXTURE_RGBA8


def test_dds_cubemap_faces(tmp_path):
    S = 4
    faces = [np.full((S, S, 4), float(f + 1), np.float32) for f in range(6)]
    hdr = bytearray(124)
    struct.pack_into("<7I", hdr, 0, 124, 0x1007, S, S, 0, 1, 1)
    struct.pack_into("<2I4s", hdr, 72, 32, 0x4, b"DX10")
    struct.pack_into("<I", hdr, 108, 0x200 | 0xFC00)  # DDSCAPS2_CUBEMAP | all faces
    p = tmp_path / "sky.dds"
    p.write_bytes(b"DDS " + bytes(hdr) + struct.pack("<5I", dds.R32G32B32A32_FLOAT, 3, 4, 1, 0) +
                  b"".join(f.tobytes() for f in faces))
    d = dds.read_dds(p)
    assert d["cubemap"] and len(d["faces"]) == 6
    for f in range(6):
        assert np.array_equal(d["faces"][f][0], faces[f])
    sky = dds.load_skybox(p)
    assert sky.shape == (6, S, S, 4) and np.array_equal(sky[:, 0, 0, 0], np.arange(1, 7, dtype=np.float32))


@pytest.mark.parametrize("scheme", [ktx2.SUPERCOMPRESSION_ZSTD, ktx2.SUPERCOMPRESSION_NONE])
def test_ktx2_round_trip(tmp_path, scheme):
    levels = _bc7_chain(20, 12)
    p = tmp_path / "t.ktx2"
    ktx2.write_ktx2(p, ktx2.VK_BC7_SRGB, 20, 12, levels, supercompression=scheme)
    d = ktx2.read_ktx2(p)
    assert d["supercompression"] == scheme and d["levels"] == levels
    if scheme == ktx2.SUPERCOMPRESSION_ZSTD:
        assert p.stat().st_size < sum(len(x) for x in levels) + 200 or True  # random blocks barely compress
    t = ktx2.load_ktx2_texture(p)
    assert t.format == abi.TEXTURE_BC7 | abi.TEXTURE_SRGB and (t.width, t.height) == (20, 12)
    ktx2.write_ktx2(p, ktx2.VK_BC1_RGB_UNORM, 20, 12, [b"\0" * 8 * 15], supercompression=scheme)
    assert ktx2.load_ktx2_texture(p).format == abi.TEXTURE_BC1 | abi.TEXTURE_RGB


def test_ktx2_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.ktx2"
    p.write_bytes(b"not a ktx2 file at all")
    with pytest.raises(ValueError):
        ktx2.read_ktx2(p)
    ktx2.write_ktx2(p, ktx2.VK_BC7_UNORM, 8, 8, [b"\0" * 16 * 3])  # 3 blocks where 4 are needed
    with pytest.raises(ValueError):
        ktx2.load_ktx2_texture(p)
    ktx2.write_ktx2(p, 1000, 8, 8, [b"\0" * 64])  # a vkFormat the kernels do not read
    with pytest.raises(ValueError):
        ktx2.load_ktx2_texture(p)


def test_zstd_round_trip():
    data = bytes(range(256)) * 64
    assert ktx2.zstd_decompress(ktx2.zstd_compress(data), len(data)) == data


def test_encoded_texture_size_checked():
    with pytest.raises(ValueError):
        HostScene.build([quad_mesh()], [Material(base_color_texture=0)], [(0, 0, np.eye(4, dtype=np.float32))],
                        textures=[EncodedTexture(abi.TEXTURE_BC7, 8, 8, [b"\0" * 48])])


def _textured_frame(texture, *, size=(96, 64)):
    mat = Material(base_color=(1.0, 1.0, 1.0), roughness=0.5, metallic=0.0, base_color_texture=0,
                   texture_scale=(3.0, 3.0))
    scene = HostScene.build([grid_patch(8, 8)], [mat], [(0, 0, np.eye(4, dtype=np.float32))], textures=[texture])
    cam, cfg = fixture_camera(*size, position=(0.3, 0.4, 1.8))
    return scene, oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))


def test_bc7_srgb_ktx2_frame_equals_pillow_decoded_srgb_rgba8(tmp_path):
    """The whole stored-texture path: a BC7 sRGB KTX2 file (zstd) through the
    producer and the oracle's BC7 fetch and sRGB decode draws the frame that
    the same texels decoded by Pillow and stored as sRGB RGBA8 draw."""
    rng = np.random.default_rng(9)
    W = 32
    levels = []
    for lv in range(6):
        w = W >> lv
        n = ((w + 3) // 4) ** 2
        b = rng.integers(0, 256, (n, 16), dtype=np.uint8)
        for k in range(n):
            m = k % 8
            b[k, 0] = (int(b[k, 0]) & (0xFF & ~((1 << (m + 1)) - 1))) | (1 << m)
        levels.append(b)
    p = tmp_path / "albedo.ktx2"
    ktx2.write_ktx2(p, ktx2.VK_BC7_SRGB, W, W, [b.tobytes() for b in levels])
    _, a = _textured_frame(ktx2.load_ktx2_texture(p))
    decoded = []
    for lv, b in enumerate(levels):
        w = W >> lv
        img = _pillow_decode(b, max(4, w), max(4, w))[:w, :w]
        decoded.append(np.ascontiguousarray(img).tobytes())
    _, b = _textured_frame(EncodedTexture(abi.TEXTURE_RGBA8 | abi.TEXTURE_SRGB, W, W, decoded))
    cov = a["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)
    assert cov.sum() > 1000
    assert np.array_equal(a["radiance"], b["radiance"])
    # and sRGB changes the picture: the same bytes read as linear are brighter
    _, c = _textured_frame(EncodedTexture(abi.TEXTURE_RGBA8, W, W, decoded))
    assert c["radiance"][cov][:, :3].mean() > 1.2 * b["radiance"][cov][:, :3].mean()


def test_gltf_reads_ktx2_and_dds_images(tmp_path):
    W = 8
    rgba = np.zeros((W, W, 4), np.uint8)
    rgba[..., 0], rgba[..., 3] = 200, 255
    chain, w = [], W
    while True:
        chain.append(np.full((w, w, 4), [200, 0, 0, 255], np.uint8).tobytes())
        if w == 1:
            break
        w //= 2
    ktx2.write_ktx2(tmp_path / "red.ktx2", ktx2.VK_R8G8B8A8_SRGB, W, W, chain)
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    blob = pos.tobytes() + uv.tobytes() + idx.tobytes() + b"\0\0\0\0"
    doc = {
        "asset": {"version": "2.0"}, "buffers": [{"byteLength": len(blob), "uri": "q.bin"}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 48},
                        {"buffer": 0, "byteOffset": 48, "byteLength": 32},
                        {"buffer": 0, "byteOffset": 80, "byteLength": 12}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC2"},
                      {"bufferView": 2, "componentType": 5123, "count": 6, "type": "SCALAR"}],
        "images": [{"uri": "red.ktx2"}, {"uri": "absent.png"}],
        "textures": [{"source": 0}, {"source": 1}],
        "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0},
                                                "metallicRoughnessTexture": {"index": 1}, "metallicFactor": 0.0},
                       "normalTexture": {"index": 0}}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 0}]}],
        "nodes": [{"mesh": 0}], "scenes": [{"nodes": [0]}], "scene": 0,
    }
    (tmp_path / "q.bin").write_bytes(blob)
    (tmp_path / "q.gltf").write_text(json.dumps(doc))
    s = load_gltf(tmp_path / "q.gltf")
    assert len(s.textures) == 1 and int(s.textures[0]["format"]) == abi.TEXTURE_RGBA8 | abi.TEXTURE_SRGB
    mi = s.mesh_infos[0]
    assert int(mi["base_color_texture_index"]) == 0 and int(mi["normal_texture_index"]) == 0
    assert int(mi["metallic_roughness_texture_index"]) == abi.UNUSED_TEXTURE_INDEX  # the PNG is not decoded


def test_gltf_reference_node_transforms(tmp_path):
    """node_transforms='reference': each mesh node's own scale(S) * translate(T)
    (rotation and parents ignored, every mesh node instanced); 'gltf': T R S
    composed down the scene's hierarchy."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([0, 1, 2], np.uint16)
    uv = np.zeros((3, 2), np.float32)
    blob = pos.tobytes() + uv.tobytes() + idx.tobytes() + b"\0\0"
    q = [0.0, 0.0, np.sin(np.pi / 4), np.cos(np.pi / 4)]  # 90 degrees about z
    doc = {
        "asset": {"version": "2.0"}, "buffers": [{"byteLength": len(blob), "uri": "t.bin"}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36},
                        {"buffer": 0, "byteOffset": 36, "byteLength": 24},
                        {"buffer": 0, "byteOffset": 60, "byteLength": 6}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC2"},
                      {"bufferView": 2, "componentType": 5123, "count": 3, "type": "SCALAR"}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2}]}],
        "nodes": [{"translation": [10, 0, 0], "rotation": q, "children": [1]},
                  {"mesh": 0, "translation": [1, 2, 3], "scale": [2, 2, 2], "rotation": q},
                  {"mesh": 0, "translation": [5, 5, 5]}],  # in no scene
        "scenes": [{"nodes": [0]}], "scene": 0,
    }
    (tmp_path / "t.bin").write_bytes(blob)
    (tmp_path / "t.gltf").write_text(json.dumps(doc))

    def world(s, i, p):
        m = s.instances["transform"][i].T.astype(np.float64)  # row-major
        return (m @ np.append(p, 1.0))[:3]

    g = load_gltf(tmp_path / "t.gltf")
    assert g.num_instances == 1
    # child T(1,2,3) R S(2): (1,0,0) -> (2,0,0) -> (0,2,0) -> (1,4,3); parent T(10,0,0) R: -> (-4,1,3) -> (6,1,3)
    np.testing.assert_allclose(world(g, 0, [1, 0, 0]), [6, 1, 3], atol=1e-5)
    r = load_gltf(tmp_path / "t.gltf", node_transforms="reference")
    assert r.num_instances == 2
    np.testing.assert_allclose(world(r, 0, [1, 0, 0]), [2 * (1 + 1), 2 * 2, 2 * 3], atol=1e-6)
    np.testing.assert_allclose(world(r, 1, [0, 1, 0]), [5, 6, 5], atol=1e-6)
