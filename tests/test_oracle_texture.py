"""BC1 (S3TC) textures: the decoder the oracle shares with the kernels'
restatement, the producer's encoder, and frames with BC1 textures.

The decoder is pinned by hand-worked blocks of the published BC1 format
(little-endian RGB565 endpoints c0, c1; 2-bit indices, texel (x, y) at bits
2 (4 y + x); endpoints widened by bit replication; c0 > c1: thirds rounded to
nearest, else the rounded midpoint and transparent black) and by an
independent numpy decoder over every block of an encoded scene.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from lighthugger_amd import abi
from lighthugger_amd.scene import HostScene, make_camera, make_uniforms, preset_camera
from oracle.compare import oracle_frame


def _block(c0, c1, idx):
    return np.array([c0 & 255, c0 >> 8, c1 & 255, c1 >> 8, idx & 255, (idx >> 8) & 255, (idx >> 16) & 255,
                     idx >> 24], np.uint8)


def _decode(block, x, y):
    v = oracle.lib().orc_bc1_texel(block.ctypes.data, x, y)
    return (v & 255, (v >> 8) & 255, (v >> 16) & 255, v >> 24)


def test_bc1_known_blocks():
    idx = sum((k % 4) << (2 * k) for k in range(16))  # texel k uses index k % 4
    four = _block(0xF800, 0x001F, idx)  # red > blue: four-colour mode
    want4 = [(255, 0, 0, 255), (0, 0, 255, 255), (170, 0, 85, 255), (85, 0, 170, 255)]
    three = _block(0x001F, 0xF800, idx)  # blue < red: three colours + transparent black
    want3 = [(0, 0, 255, 255), (255, 0, 0, 255), (128, 0, 128, 255), (0, 0, 0, 0)]
    for k in range(16):
        assert _decode(four, k & 3, k >> 2) == want4[k % 4], k
        assert _decode(three, k & 3, k >> 2) == want3[k % 4], k
    # bit replication: r5 = 16 -> 132, g6 = 32 -> 130, b5 = 16 -> 132
    grey = _block(0x8410, 0x0000, 0)
    assert _decode(grey, 2, 3) == (132, 130, 132, 255)


def _np_decode_level(data, w, h):
    """Independent numpy BC1 decoder of one level (blocks row-major)."""
    bw, bh = (w + 3) // 4, (h + 3) // 4
    blk = data[: bw * bh * 8].reshape(bh, bw, 8).astype(np.uint32)
    c0 = blk[..., 0] | (blk[..., 1] << 8)
    c1 = blk[..., 2] | (blk[..., 3] << 8)
    idx = blk[..., 4] | (blk[..., 5] << 8) | (blk[..., 6] << 16) | (blk[..., 7] << 24)

    def widen(c):
        r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
        return np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], -1).astype(np.int64)

    a, b = widen(c0), widen(c1)
    four = (c0 > c1)[..., None]
    pal = np.stack([a, b, np.where(four, (2 * a + b + 1) // 3, (a + b + 1) // 2),
                    np.where(four, (a + 2 * b + 1) // 3, 0)], -2)  # [bh, bw, 4, 3]
    out = np.zeros((bh * 4, bw * 4, 3), np.int64)
    for k in range(16):
        sel = ((idx >> (2 * k)) & 3).astype(np.int64)
        col = np.take_along_axis(pal, sel[..., None, None].repeat(3, -1), axis=-2)[..., 0, :]
        out[(k >> 2)::4, (k & 3)::4] = col
    return out[:h, :w]


def test_encoder_round_trip_and_decoders_agree():
    kw = dict(tri_scale_pct=5, texture_size=64)
    a = HostScene.preset(abi.PRESET_LARGE, **kw)
    b = HostScene.preset(abi.PRESET_LARGE, texture_format=abi.TEXTURE_BC1, **kw)
    assert (b.textures["format"] == abi.TEXTURE_BC1).all() and (a.textures["format"] == abi.TEXTURE_RGBA8).all()
    assert b.texels.nbytes * 6 < a.texels.nbytes  # 8 bytes per 16 texels vs 64
    for t in range(0, len(a.textures), 7):
        ta, tb = a.textures[t], b.textures[t]
        for level in (0, 2, int(ta["levels"]) - 1):
            w, h = max(1, int(ta["width"]) >> level), max(1, int(ta["height"]) >> level)
            src = a.texels[int(ta["texels"]) + 4 * int(ta["level_offset"][level]):][: 4 * w * h]
            src = src.reshape(h, w, 4)[..., :3].astype(np.int64)
            data = b.texels[int(tb["texels"]) + 8 * int(tb["level_offset"][level]):]
            dec = _np_decode_level(data, w, h)
            err = np.abs(dec - src)
            assert err.mean() < 8.0 and err.max() <= 64, (t, level, err.mean(), err.max())
            if level == 0:  # the oracle's decoder agrees with the numpy one on every texel of a few blocks
                for by in range(0, (h + 3) // 4, 5):
                    for bx in range(0, (w + 3) // 4, 5):
                        blk = np.ascontiguousarray(data[8 * (by * ((w + 3) // 4) + bx):][:8])
                        for k in range(16):
                            x, y = bx * 4 + (k & 3), by * 4 + (k >> 2)
                            assert _decode(blk, k & 3, k >> 2)[:3] == tuple(dec[y, x]), (t, bx, by, k)


def test_bc1_frame_close_to_rgba8_frame():
    kw = dict(tri_scale_pct=5, texture_size=64)
    base = preset_camera(abi.PRESET_LARGE)
    cam = make_camera(base.position.tolist(), base.target.tolist(), 160, 96)
    from lighthugger_amd.frame import ViewConfig

    cfg = ViewConfig(160, 96, 256, big_capacity=1 << 16)
    frames = []
    for fmt in (abi.TEXTURE_RGBA8, abi.TEXTURE_BC1):
        s = HostScene.preset(abi.PRESET_LARGE, texture_format=fmt, **kw)
        frames.append(oracle_frame(s, cfg, make_uniforms(cam, s.num_instances)))
    a, b = frames
    assert np.array_equal(a["visbuffer"], b["visbuffer"])
    cov = a["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)
    rel = np.abs(a["radiance"][cov][:, :3] - b["radiance"][cov][:, :3]) / np.maximum(a["radiance"][cov][:, :3], 1e-3)
    assert np.median(rel) < 0.05
