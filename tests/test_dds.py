"""DDS reading (lighthugger_amd/dds.py; the reference's load_dds,
src/resources/image_loading.cpp:67-160): files written here with the
DDS_HEADER + DDS_HEADER_DXT10 layout.  RGB9E5 is pinned by hand-worked words
(m * 2^(e - 24)), BC1 by the oracle's decoder, and the LUT path end to end:
an identity LUT stored as RGB9E5 gives the LUT-free display curve."""
import struct

import numpy as np

import oracle
from lighthugger_amd import abi, dds
from lighthugger_amd.scene import HostScene, Material, make_uniforms

from .util import fixture_camera, grid_patch


def _write_dds(path, width, height, depth, fmt, payload, mips=1):
    hdr = bytearray(124)
    struct.pack_into("<7I", hdr, 0, 124, 0x1007 | (0x800000 if depth > 1 else 0), height, width, 0, depth, mips)
    struct.pack_into("<I", hdr, 72, 32)
    hdr[80:84] = b"DX10"
    dx10 = struct.pack("<5I", fmt, 4 if depth > 1 else 3, 0, 1, 0)
    path.write_bytes(b"DDS " + bytes(hdr) + dx10 + payload)


def _rgb9e5(v):
    """Encoder for the test (values in [0, 65408]): smallest exponent keeping the largest mantissa <= 511."""
    v = np.asarray(v, np.float64)
    m = np.maximum(v.max(-1), 2.0 ** -16)
    e = np.clip(np.floor(np.log2(m)).astype(np.int64) + 16, 0, 31)
    e = np.where(np.round(m / 2.0 ** (e - 24)) > 511, e + 1, e)
    q = np.clip(np.round(v / 2.0 ** (e - 24)[..., None]), 0, 511).astype(np.uint32)
    return (q[..., 0] | (q[..., 1] << 9) | (q[..., 2] << 18) | (e.astype(np.uint32) << 27)).astype(np.uint32)


def test_rgb9e5_known_words():
    words = np.array([256 | (16 << 27), 511 | (511 << 9) | (511 << 18) | (31 << 27), (1 << 9) | (0 << 27), 0],
                     np.uint32)
    v = dds.rgb9e5_to_float(words)
    np.testing.assert_array_equal(v[0], [1.0, 0, 0])
    np.testing.assert_array_equal(v[1], [65408.0] * 3)  # the format's largest value
    np.testing.assert_array_equal(v[2], [0, 2.0 ** -24, 0])
    np.testing.assert_array_equal(v[3], [0, 0, 0])


def test_bc1_dds_matches_oracle_decoder(tmp_path):
    rng = np.random.default_rng(5)
    blocks = rng.integers(0, 256, (3 * 2, 8), dtype=np.uint8)  # 12 x 8 texels
    p = tmp_path / "t.dds"
    _write_dds(p, 12, 8, 1, dds.BC1_UNORM, blocks.tobytes())
    img = dds.read_dds(p)["levels"][0]
    assert img.shape == (8, 12, 4)
    for by in range(2):
        for bx in range(3):
            blk = np.ascontiguousarray(blocks[by * 3 + bx])
            for k in range(16):
                v = oracle.lib().orc_bc1_texel(blk.ctypes.data, k & 3, k >> 2)
                want = (v & 255, (v >> 8) & 255, (v >> 16) & 255, v >> 24)
                assert tuple(img[by * 4 + (k >> 2), bx * 4 + (k & 3)]) == want


def test_tonemap_lut_from_rgb9e5_dds(tmp_path):
    n = abi.TONEMAP_LUT_DIMS
    b, g, r = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ident = np.stack([r, g, b], -1) / (n - 1)
    p = tmp_path / "lut.dds"
    _write_dds(p, n, n, n, dds.R9G9B9E5_SHAREDEXP, _rgb9e5(ident).tobytes())
    lut = dds.load_tonemap_lut(p)
    assert lut.shape == (n, n, n, 4) and lut.dtype == np.float32
    np.testing.assert_allclose(lut[..., :3], ident, atol=1.0 / 512)
    # end to end through the oracle's display transform
    from oracle import lib as oracle_lib
    from oracle.compare import oracle_view, snapshot

    mat = Material(base_color=(0.8, 0.6, 0.4), roughness=0.5, metallic=0.0)
    scene = HostScene.build([grid_patch(4, 4)], [mat], [(0, 0, np.eye(4, dtype=np.float32))])
    cam, cfg = fixture_camera(48, 32, position=(0, 0, 2.2))
    outs = []
    for use in (False, True):
        sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
        if use:
            vd.c.tonemap_lut = lut.ctypes.data
        oracle_lib().orc_frame(sd.address, vd.address)
        outs.append(snapshot(vd)["ldr"])
    ch = [np.stack([(x >> 16) & 255, (x >> 8) & 255, x & 255], -1).astype(np.int64) for x in outs]
    assert np.abs(ch[0] - ch[1]).max() <= 2
