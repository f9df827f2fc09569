"""DDS reading for the texture files the reference ships (load_dds,
src/resources/image_loading.cpp:67-240, header layout src/resources/dds.h):
magic "DDS ", the 124-byte DDS_HEADER and the DDS_HEADER_DXT10 extension
(the reference always reads it; files with the legacy "DXT1" FourCC and no
extension are accepted too).

    lut = load_tonemap_lut("tony_mc_mapface.dds")   # 48^3 RGBA float for lh_view.tonemap_lut
    tex = load_dds_texture("albedo.dds")             # EncodedTexture for HostScene.build(textures=...)
    img = read_dds("albedo.dds")                     # levels decoded (BC1, RGBA8, RGB9E5, RGBA32F)
    raw = read_dds("albedo.dds", raw=True)           # levels as stored (every format, BC7 included)

Formats: DXGI_FORMAT_R9G9B9E5_SHAREDEXP (67; shared-exponent floats, the
tony-mc-mapface LUT and the skybox), DXGI_FORMAT_BC1_UNORM(_SRGB) (71, 72),
DXGI_FORMAT_BC7_UNORM(_SRGB) (98, 99), DXGI_FORMAT_R8G8B8A8_UNORM(_SRGB) (28,
29), DXGI_FORMAT_R32G32B32A32_FLOAT (2).  As the reference uploads them
(translate_format, src/resources/image_loading.cpp:14-39), BC1_UNORM and both
BC7 formats become sRGB textures (Vulkan's BC1_RGB_SRGB / BC7_SRGB): the
sampler decodes them to linear before filtering, and BC1 reads alpha 1.

Cube maps: the six faces are returned in the DDS order (+X, -X, +Y, -Y, +Z,
-Z), each face with its own mip chain, face after face as DDS files store
them.  The reference's copy regions read every level's six faces
consecutively (src/resources/image_loading.cpp:216-240), which is the same
order for single-level cube maps such as its skybox.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

from . import abi

R32G32B32A32_FLOAT, R8G8B8A8_UNORM, R8G8B8A8_UNORM_SRGB = 2, 28, 29
R9G9B9E5_SHAREDEXP, BC1_UNORM, BC1_UNORM_SRGB = 67, 71, 72
BC7_UNORM, BC7_UNORM_SRGB = 98, 99
_DDSCAPS2_CUBEMAP = 0x200

# DXGI format -> LhTexture.format, as the reference maps them to Vulkan formats
TEXTURE_FORMAT = {
    BC1_UNORM: abi.TEXTURE_BC1 | abi.TEXTURE_SRGB | abi.TEXTURE_RGB,
    BC1_UNORM_SRGB: abi.TEXTURE_BC1 | abi.TEXTURE_SRGB | abi.TEXTURE_RGB,
    BC7_UNORM: abi.TEXTURE_BC7 | abi.TEXTURE_SRGB,
    BC7_UNORM_SRGB: abi.TEXTURE_BC7 | abi.TEXTURE_SRGB,
    R8G8B8A8_UNORM: abi.TEXTURE_RGBA8,
    R8G8B8A8_UNORM_SRGB: abi.TEXTURE_RGBA8 | abi.TEXTURE_SRGB,
}


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


def level_bytes(fmt: int, w: int, h: int, d: int = 1) -> int:
    """Stored size of one level (one face) of a DXGI format."""
    if fmt in (BC1_UNORM, BC1_UNORM_SRGB):
        return ((w + 3) // 4) * ((h + 3) // 4) * 8 * d
    if fmt in (BC7_UNORM, BC7_UNORM_SRGB):
        return ((w + 3) // 4) * ((h + 3) // 4) * 16 * d
    if fmt in (R8G8B8A8_UNORM, R8G8B8A8_UNORM_SRGB, R9G9B9E5_SHAREDEXP):
        return w * h * d * 4
    if fmt == R32G32B32A32_FLOAT:
        return w * h * d * 16
    raise ValueError(f"DXGI format {fmt} is not read")


def _decode(fmt: int, buf: bytes, w: int, h: int, d: int):
    if fmt in (BC1_UNORM, BC1_UNORM_SRGB):
        n = level_bytes(fmt, w, h)
        return np.stack([bc1_to_rgba8(np.frombuffer(buf, np.uint8, n, z * n), w, h) for z in range(d)])
    if fmt in (R8G8B8A8_UNORM, R8G8B8A8_UNORM_SRGB):
        return np.frombuffer(buf, np.uint8).reshape(d, h, w, 4).copy()
    if fmt == R9G9B9E5_SHAREDEXP:
        return rgb9e5_to_float(np.frombuffer(buf, np.uint32).reshape(d, h, w))
    if fmt == R32G32B32A32_FLOAT:
        return np.frombuffer(buf, np.float32).reshape(d, h, w, 4).copy()
    raise ValueError(f"DXGI format {fmt} is not decoded on the host (read it with raw=True)")


def read_dds(path, raw: bool = False) -> dict:
    """{"width", "height", "depth", "format" (DXGI), "cubemap", "levels", "faces"}.

    levels: each mip level of the first face - decoded (RGBA8 uint8 [d?, h, w, 4]
    or float32 [d?, h, w, 3|4], the depth axis only for volume textures), or
    with raw=True the stored bytes of each level (every format, BC7 included).
    faces: for a cube map, the six faces' level lists (else [levels])."""
    data = Path(path).read_bytes()
    if data[:4] != b"DDS ":
        raise ValueError(f"{path}: not a DDS file")
    (size, _flags, height, width, _pitch, depth, mips) = struct.unpack_from("<7I", data, 4)
    if size != 124:
        raise ValueError(f"{path}: DDS_HEADER size {size}")
    fourcc = data[4 + 80: 4 + 84]
    caps2 = struct.unpack_from("<I", data, 4 + 108)[0]
    offset = 4 + 124
    array_size = 1
    if fourcc == b"DX10":
        fmt, _dim, _misc, array_size = struct.unpack_from("<4I", data, offset)
        offset += 20
    elif fourcc == b"DXT1":
        fmt = BC1_UNORM
    else:
        raise ValueError(f"{path}: unsupported pixel format {fourcc!r}")
    cubemap = bool(caps2 & _DDSCAPS2_CUBEMAP)
    depth, mips = max(depth, 1), max(mips, 1)
    nfaces = 6 if cubemap else max(1, array_size)
    faces, off = [], offset
    for _ in range(nfaces):
        levels = []
        for level in range(mips):
            w, h, d = max(1, width >> level), max(1, height >> level), max(1, depth >> level)
            n = level_bytes(fmt, w, h, d)
            if off + n > len(data):
                raise ValueError(f"{path}: truncated at level {level}")
            buf = data[off: off + n]
            off += n
            if raw:
                levels.append(np.frombuffer(buf, np.uint8).copy())
            else:
                img = _decode(fmt, buf, w, h, d)
                levels.append(img if depth > 1 else img[0])
        faces.append(levels)
    return {"width": width, "height": height, "depth": depth, "format": fmt, "cubemap": cubemap,
            "levels": faces[0], "faces": faces}


def load_dds_texture(path):
    """A DDS file as the EncodedTexture the scene producer takes, with the
    reference's format mapping (BC1 and BC7 sampled as sRGB)."""
    from .scene import EncodedTexture

    d = read_dds(path, raw=True)
    if d["format"] not in TEXTURE_FORMAT:
        raise ValueError(f"{path}: DXGI format {d['format']} is not a texture format of the kernels")
    if d["cubemap"] or d["depth"] > 1:
        raise ValueError(f"{path}: a 2D texture was expected")
    return EncodedTexture(TEXTURE_FORMAT[d["format"]], d["width"], d["height"], d["levels"])


def load_tonemap_lut(path) -> np.ndarray:
    """The display LUT of a 3D DDS (the reference's tony_mc_mapface.dds,
    48^3 RGB9E5) as the contiguous RGBA float32 [48, 48, 48, 4] array
    lh_view.tonemap_lut expects (red fastest, alpha 1)."""
    d = read_dds(path)
    n = abi.TONEMAP_LUT_DIMS
    lut = np.asarray(d["levels"][0], np.float32)
    if lut.shape[:3] != (n, n, n):
        raise ValueError(f"{path}: LUT of shape {lut.shape[:3]}, expected {n}^3")
    if lut.shape[3] == 3:
        lut = np.concatenate([lut, np.ones(lut.shape[:3] + (1,), np.float32)], -1)
    return np.ascontiguousarray(lut)


def load_skybox(path) -> np.ndarray:
    """A cube-map DDS (the reference's hdr-cubemap-1024x1024.dds, RGB9E5) as
    the contiguous RGBA float32 [6, S, S, 4] array of level 0 that
    lh_view.skybox expects (faces +X, -X, +Y, -Y, +Z, -Z; alpha 1)."""
    d = read_dds(path)
    if not d["cubemap"] or d["width"] != d["height"]:
        raise ValueError(f"{path}: a square cube map was expected")
    faces = [np.asarray(f[0], np.float32) for f in d["faces"]]
    if faces[0].ndim != 3:
        raise ValueError(f"{path}: float faces were expected (RGB9E5 or RGBA32F)")
    faces = [f if f.shape[-1] == 4 else np.concatenate([f, np.ones(f.shape[:2] + (1,), np.float32)], -1)
             for f in faces]
    return np.ascontiguousarray(np.stack(faces))
