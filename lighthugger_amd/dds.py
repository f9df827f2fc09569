"""DDS reading for the texture files the reference ships (load_dds,
src/resources/image_loading.cpp:67-160, header layout src/resources/dds.h):
magic "DDS ", the 124-byte DDS_HEADER and the DDS_HEADER_DXT10 extension
(the reference always reads it; files with the legacy "DXT1" FourCC and no
extension are accepted too).

    lut = load_tonemap_lut("tony_mc_mapface.dds")   # 48^3 RGBA float for lh_view.tonemap_lut
    img = read_dds("albedo.dds")                     # levels decoded to RGBA8 (BC1, RGBA8) or float (RGB9E5)

Formats: DXGI_FORMAT_R9G9B9E5_SHAREDEXP (67; shared-exponent floats, the
tony-mc-mapface LUT), DXGI_FORMAT_BC1_UNORM(_SRGB) (71, 72; decoded with the
published BC1 rules the kernels use), DXGI_FORMAT_R8G8B8A8_UNORM(_SRGB) (28,
29), DXGI_FORMAT_R32G32B32A32_FLOAT (2).  BC7 and zstd-supercompressed KTX2
are not read.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

R32G32B32A32_FLOAT, R8G8B8A8_UNORM, R8G8B8A8_UNORM_SRGB = 2, 28, 29
R9G9B9E5_SHAREDEXP, BC1_UNORM, BC1_UNORM_SRGB = 67, 71, 72
_DDSCAPS2_CUBEMAP = 0x200


def rgb9e5_to_float(words: np.ndarray) -> np.ndarray:
    """Shared-exponent RGB9E5 words -> float32 [..., 3]: channel = m * 2^(e - 15 - 9)
    with 9-bit mantissas m (bits 0-8, 9-17, 18-26) and the exponent e (bits 27-31)."""
    w = np.asarray(words, np.uint32)
    scale = np.ldexp(np.float32(1.0), (w >> 27).astype(np.int32) - 24).astype(np.float32)
    return np.stack([(w & 511).astype(np.float32) * scale, ((w >> 9) & 511).astype(np.float32) * scale,
                     ((w >> 18) & 511).astype(np.float32) * scale], -1)


def bc1_to_rgba8(blocks: np.ndarray, width: int, height: int) -> np.ndarray:
    """BC1 blocks (8 bytes per 4x4, row-major) -> RGBA8 [height, width, 4]."""
    bw, bh = (width + 3) // 4, (height + 3) // 4
    b = np.asarray(blocks, np.uint8)[: bw * bh * 8].reshape(bh, bw, 8).astype(np.uint32)
    c0, c1 = b[..., 0] | (b[..., 1] << 8), b[..., 2] | (b[..., 3] << 8)
    idx = b[..., 4] | (b[..., 5] << 8) | (b[..., 6] << 16) | (b[..., 7] << 24)

    def widen(c):
        r, g, bl = (c >> 11) & 31, (c >> 5) & 63, c & 31
        return np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (bl << 3) | (bl >> 2)], -1).astype(np.int64)

    a, z = widen(c0), widen(c1)
    four = (c0 > c1)[..., None]
    pal = np.stack([a, z, np.where(four, (2 * a + z + 1) // 3, (a + z + 1) // 2),
                    np.where(four, (a + 2 * z + 1) // 3, 0)], -2)
    alpha = np.stack([np.full(c0.shape, 255), np.full(c0.shape, 255), np.full(c0.shape, 255),
                      np.where(c0 > c1, 255, 0)], -1)
    out = np.zeros((bh * 4, bw * 4, 4), np.uint8)
    for k in range(16):
        sel = ((idx >> (2 * k)) & 3).astype(np.int64)
        rgb = np.take_along_axis(pal, sel[..., None, None].repeat(3, -1), axis=-2)[..., 0, :]
        out[(k >> 2)::4, (k & 3)::4, :3] = rgb
        out[(k >> 2)::4, (k & 3)::4, 3] = np.take_along_axis(alpha, sel[..., None], -1)[..., 0]
    return out[:height, :width]


def read_dds(path) -> dict:
    """{"width", "height", "depth", "format" (DXGI), "cubemap", "levels": [arrays]}:
    each mip level decoded, RGBA8 uint8 [d?, h, w, 4] or float32 [d?, h, w, 3|4]
    (the depth axis only for volume textures)."""
    data = Path(path).read_bytes()
    if data[:4] != b"DDS ":
        raise ValueError(f"{path}: not a DDS file")
    (size, _flags, height, width, _pitch, depth, mips) = struct.unpack_from("<7I", data, 4)
    if size != 124:
        raise ValueError(f"{path}: DDS_HEADER size {size}")
    fourcc = data[4 + 80: 4 + 84]
    caps2 = struct.unpack_from("<I", data, 4 + 108)[0]
    offset = 4 + 124
    if fourcc == b"DX10":
        fmt = struct.unpack_from("<I", data, offset)[0]
        offset += 20
    elif fourcc == b"DXT1":
        fmt = BC1_UNORM
    else:
        raise ValueError(f"{path}: unsupported pixel format {fourcc!r}")
    depth, mips = max(depth, 1), max(mips, 1)
    levels, off = [], offset
    for level in range(mips):
        w, h, d = max(1, width >> level), max(1, height >> level), max(1, depth >> level)
        if fmt in (BC1_UNORM, BC1_UNORM_SRGB):
            n = ((w + 3) // 4) * ((h + 3) // 4) * 8
            img = np.stack([bc1_to_rgba8(np.frombuffer(data, np.uint8, n, off + z * n), w, h) for z in range(d)])
            off += n * d
        elif fmt in (R8G8B8A8_UNORM, R8G8B8A8_UNORM_SRGB):
            n = w * h * d * 4
            img = np.frombuffer(data, np.uint8, n, off).reshape(d, h, w, 4).copy()
            off += n
        elif fmt == R9G9B9E5_SHAREDEXP:
            n = w * h * d
            img = rgb9e5_to_float(np.frombuffer(data, np.uint32, n, off).reshape(d, h, w))
            off += 4 * n
        elif fmt == R32G32B32A32_FLOAT:
            n = w * h * d * 4
            img = np.frombuffer(data, np.float32, n, off).reshape(d, h, w, 4).copy()
            off += 4 * n
        else:
            raise ValueError(f"{path}: DXGI format {fmt} is not read")
        levels.append(img if depth > 1 else img[0])
    return {"width": width, "height": height, "depth": depth, "format": fmt,
            "cubemap": bool(caps2 & _DDSCAPS2_CUBEMAP), "levels": levels}


def load_tonemap_lut(path) -> np.ndarray:
    """The display LUT of a 3D DDS (the reference's tony_mc_mapface.dds,
    48^3 RGB9E5) as the contiguous RGBA float32 [48, 48, 48, 4] array
    lh_view.tonemap_lut expects (red fastest, alpha 1)."""
    from . import abi

    d = read_dds(path)
    n = abi.TONEMAP_LUT_DIMS
    lut = np.asarray(d["levels"][0], np.float32)
    if lut.shape[:3] != (n, n, n):
        raise ValueError(f"{path}: LUT of shape {lut.shape[:3]}, expected {n}^3")
    if lut.shape[3] == 3:
        lut = np.concatenate([lut, np.ones(lut.shape[:3] + (1,), np.float32)], -1)
    return np.ascontiguousarray(lut)
