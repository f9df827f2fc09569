"""KTX2 reading for the texture files the reference's glTF assets use
(load_ktx2_image, src/resources/image_loading.cpp:242-360; header structs
src/resources/ktx2.h): the 12-byte identifier, the header (vkFormat,
typeSize, pixel size, layer / face / level counts, supercompression scheme),
the index of the data-format descriptor, key/value and supercompression
global data, and one (byteOffset, byteLength, uncompressedByteLength) entry
per mip level, level 0 first.  Levels supercompressed with Zstandard
(scheme 2) are decompressed with the system libzstd (ZSTD_decompress, as the
reference does); scheme 0 is read as stored.

    tex = load_ktx2_texture("albedo.ktx2")   # EncodedTexture for HostScene.build(textures=...)

The file's vkFormat decides the texture format: R8G8B8A8_UNORM / _SRGB
(37, 43), BC1_RGB(A)_UNORM / _SRGB (131-134) and BC7_UNORM / _SRGB (145, 146);
the reference creates its image with that format as it is.

write_ktx2 writes the same layout (used by the tests and by callers who
convert textures offline).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import struct
from pathlib import Path

import numpy as np

from . import abi

IDENTIFIER = bytes([0xAB, 0x4B, 0x54, 0x58, 0x20, 0x32, 0x30, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A])
SUPERCOMPRESSION_NONE, SUPERCOMPRESSION_ZSTD = 0, 2

VK_R8G8B8A8_UNORM, VK_R8G8B8A8_SRGB = 37, 43
VK_BC1_RGB_UNORM, VK_BC1_RGB_SRGB, VK_BC1_RGBA_UNORM, VK_BC1_RGBA_SRGB = 131, 132, 133, 134
VK_BC7_UNORM, VK_BC7_SRGB = 145, 146

# vkFormat -> LhTexture.format
TEXTURE_FORMAT = {
    VK_R8G8B8A8_UNORM: abi.TEXTURE_RGBA8,
    VK_R8G8B8A8_SRGB: abi.TEXTURE_RGBA8 | abi.TEXTURE_SRGB,
    VK_BC1_RGB_UNORM: abi.TEXTURE_BC1 | abi.TEXTURE_RGB,
    VK_BC1_RGB_SRGB: abi.TEXTURE_BC1 | abi.TEXTURE_SRGB | abi.TEXTURE_RGB,
    VK_BC1_RGBA_UNORM: abi.TEXTURE_BC1,
    VK_BC1_RGBA_SRGB: abi.TEXTURE_BC1 | abi.TEXTURE_SRGB,
    VK_BC7_UNORM: abi.TEXTURE_BC7,
    VK_BC7_SRGB: abi.TEXTURE_BC7 | abi.TEXTURE_SRGB,
}
_BLOCK_BYTES = {abi.TEXTURE_RGBA8: 0, abi.TEXTURE_BC1: 8, abi.TEXTURE_BC7: 16}

_HEADER = struct.Struct("<9I")       # vkFormat .. supercompressionScheme
_INDEX = struct.Struct("<4I2Q")      # dfd offset/length, kvd offset/length, sgd offset/length
_LEVEL = struct.Struct("<3Q")        # byteOffset, byteLength, uncompressedByteLength

_zstd = None


def zstd() -> C.CDLL:
    """The system Zstandard library (libzstd.so.1)."""
    global _zstd
    if _zstd is None:
        name = ctypes.util.find_library("zstd") or "libzstd.so.1"
        L = C.CDLL(name)
        L.ZSTD_decompress.restype = C.c_size_t
        L.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        L.ZSTD_compress.restype = C.c_size_t
        L.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
        L.ZSTD_compressBound.restype = C.c_size_t
        L.ZSTD_compressBound.argtypes = [C.c_size_t]
        L.ZSTD_isError.restype = C.c_uint
        L.ZSTD_isError.argtypes = [C.c_size_t]
        _zstd = L
    return _zstd


def zstd_decompress(data: bytes, size: int) -> bytes:
    L = zstd()
    out = C.create_string_buffer(max(size, 1))
    n = L.ZSTD_decompress(out, size, data, len(data))
    if L.ZSTD_isError(n) or n != size:
        raise ValueError(f"zstd: level did not decompress to {size} bytes")
    return out.raw[:size]


def zstd_compress(data: bytes, level: int = 3) -> bytes:
    L = zstd()
    cap = L.ZSTD_compressBound(len(data))
    out = C.create_string_buffer(cap)
    n = L.ZSTD_compress(out, cap, data, len(data), level)
    if L.ZSTD_isError(n):
        raise ValueError("zstd: compression failed")
    return out.raw[:n]


def level_bytes(texture_format: int, w: int, h: int) -> int:
    unit = _BLOCK_BYTES[texture_format & abi.TEXTURE_FORMAT_MASK]
    return w * h * 4 if unit == 0 else ((w + 3) // 4) * ((h + 3) // 4) * unit


def read_ktx2(path) -> dict:
    """{"vk_format", "width", "height", "depth", "layers", "faces", "supercompression",
    "levels": [bytes of each level, level 0 first, decompressed]}."""
    data = Path(path).read_bytes()
    if data[:12] != IDENTIFIER:
        raise ValueError(f"{path}: not a KTX2 file")
    (vk_format, _type_size, width, height, depth, layers, faces, nlevels,
     scheme) = _HEADER.unpack_from(data, 12)
    _INDEX.unpack_from(data, 12 + _HEADER.size)
    if scheme not in (SUPERCOMPRESSION_NONE, SUPERCOMPRESSION_ZSTD):
        raise ValueError(f"{path}: supercompression scheme {scheme} is not read")
    levels = []
    base = 12 + _HEADER.size + _INDEX.size
    for i in range(max(1, nlevels)):
        off, length, raw_len = _LEVEL.unpack_from(data, base + i * _LEVEL.size)
        if off + length > len(data):
            raise ValueError(f"{path}: level {i} lies beyond the end of the file")
        buf = data[off: off + length]
        if scheme == SUPERCOMPRESSION_ZSTD:
            buf = zstd_decompress(buf, raw_len)
        elif length != raw_len:
            raise ValueError(f"{path}: level {i} length {length} != uncompressed length {raw_len}")
        levels.append(buf)
    return {"vk_format": vk_format, "width": width, "height": height, "depth": depth, "layers": layers,
            "faces": faces, "supercompression": scheme, "levels": levels}


def load_ktx2_texture(path):
    """A 2D KTX2 file as the EncodedTexture the scene producer takes."""
    from .scene import EncodedTexture

    d = read_ktx2(path)
    if d["vk_format"] not in TEXTURE_FORMAT:
        raise ValueError(f"{path}: vkFormat {d['vk_format']} is not a texture format of the kernels")
    if d["faces"] > 1 or d["depth"] > 1 or d["layers"] > 1:
        raise ValueError(f"{path}: a 2D texture was expected")
    fmt = TEXTURE_FORMAT[d["vk_format"]]
    for i, lv in enumerate(d["levels"]):
        n = level_bytes(fmt, max(1, d["width"] >> i), max(1, d["height"] >> i))
        if len(lv) != n:
            raise ValueError(f"{path}: level {i} holds {len(lv)} bytes, expected {n}")
    return EncodedTexture(fmt, d["width"], d["height"], [np.frombuffer(lv, np.uint8) for lv in d["levels"]])


def write_ktx2(path, vk_format: int, width: int, height: int, levels, supercompression=SUPERCOMPRESSION_ZSTD):
    """Writes a 2D KTX2 file: `levels` are the stored bytes of each mip level,
    level 0 first; level data is laid out smallest level first, each level
    8-byte aligned, as the KTX2 specification recommends."""
    levels = [bytes(lv) for lv in levels]
    payload = [zstd_compress(lv) if supercompression == SUPERCOMPRESSION_ZSTD else lv for lv in levels]
    n = len(levels)
    header = IDENTIFIER + _HEADER.pack(vk_format, 1, width, height, 0, 0, 1, n, supercompression)
    index_end = len(header) + _INDEX.size + n * _LEVEL.size
    dfd = struct.pack("<I", 4)  # an empty data-format descriptor block (total size only)
    pos = index_end + len(dfd)
    offsets = [0] * n
    body = b""
    for i in reversed(range(n)):
        pad = (-pos) % 8
        body += b"\0" * pad
        pos += pad
        offsets[i] = pos
        body += payload[i]
        pos += len(payload[i])
    index = _INDEX.pack(index_end, len(dfd), 0, 0, 0, 0)
    level_index = b"".join(_LEVEL.pack(offsets[i], len(payload[i]), len(levels[i])) for i in range(n))
    Path(path).write_bytes(header + index + level_index + dfd + body)
