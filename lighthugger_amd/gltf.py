"""glTF 2.0 ingest for the scene producer (the reference's load_gltf,
src/resources/mesh_loading.cpp:241-738, which reads its assets through
fastgltf with KHR_mesh_quantization and KHR_texture_transform).

    scene = load_gltf("model.gltf")                 # .gltf (+ .bin / data: URIs) or .glb
    scene = load_gltf(path, vertex_format=abi.VERTEX_FORMAT_QUANTIZED, images={0: rgba})

What is read:
  * triangle primitives (mode 4) of every mesh a node references, with
    POSITION, NORMAL, TEXCOORD_0 and indices accessors of any component type
    KHR_mesh_quantization allows (float, (unsigned) byte / short, normalized
    or not; indices u8 / u16 / u32), dequantized by the glTF rules
    (normalized: c / 255, c / 65535, max(c / 127, -1), max(c / 32767, -1));
  * node transforms, node_transforms="gltf" (default): the glTF rules - the
    node's matrix, or translation * rotation * scale, composed down the
    hierarchy of the chosen scene, each (node, primitive) pair one instance
    with the node's world matrix (KHR_mesh_quantization puts the position
    dequantization there); node_transforms="reference": what the reference's
    NodeTree computes (src/resources/mesh_loading.cpp:16-47, 356-368) - every
    node with a mesh, in node order, whether or not a scene lists it, with its
    own scale(S) * translate(T) (rotation, matrices and parents ignored);
  * materials: base-colour factor, roughness and metallic factors,
    alphaMode MASK (cutoff 0.5, as the reference requires) -> alpha clip; the
    base-colour, metallic-roughness and normal textures, each with its
    KHR_texture_transform scale / offset (the last one present wins, as in
    the reference, src/resources/mesh_loading.cpp:495-550);
  * images: .ktx2 and .dds URIs are loaded as stored (ktx2.py, dds.py: BC1,
    BC7, RGBA8, sRGB as the file / the reference says), decoded on fetch by
    the kernels; other images (PNG / JPEG) only when the caller passes them
    decoded (`images`: glTF image index -> RGBA8 array H x W x 4, or an
    EncodedTexture), which also overrides a file.  A texture whose image is
    neither is left unused.

The producer then builds meshlets and bounds, and writes the vertices in
the requested vertex format (FP32, or the reference's quantized formats).
"""
from __future__ import annotations

import base64
import json
import struct
from pathlib import Path

import numpy as np

from . import abi
from .dds import load_dds_texture
from .ktx2 import load_ktx2_texture
from .scene import EncodedTexture, HostScene, Material

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_COMPONENTS = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def _read_container(path: Path) -> tuple[dict, list[bytes]]:
    data = path.read_bytes()
    if data[:4] == b"glTF":  # binary container: header, JSON chunk, optional BIN chunk
        _, _, length = struct.unpack_from("<4sII", data, 0)
        off, doc, bin_chunk = 12, None, None
        while off < length:
            n, kind = struct.unpack_from("<II", data, off)
            body = data[off + 8: off + 8 + n]
            if kind == 0x4E4F534A:
                doc = json.loads(body.decode("utf-8"))
            elif kind == 0x004E4942:
                bin_chunk = body
            off += 8 + n
        if doc is None:
            raise ValueError(f"{path}: no JSON chunk")
    else:
        doc, bin_chunk = json.loads(data.decode("utf-8")), None
    buffers = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise ValueError(f"{path}: buffer {i} has no data")
            buffers.append(bin_chunk)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            buffers.append((path.parent / uri).read_bytes())
    return doc, buffers


def _accessor(doc: dict, buffers: list[bytes], index: int) -> np.ndarray:
    """Accessor `index` as float32 [count, components] (or uint32 for
    integer scalars), dequantized by the glTF normalization rules."""
    a = doc["accessors"][index]
    dtype = np.dtype(_COMPONENT[a["componentType"]])
    ncomp, count = _COMPONENTS[a["type"]], a["count"]
    view = doc["bufferViews"][a["bufferView"]]
    buf = buffers[view["buffer"]]
    start = view.get("byteOffset", 0) + a.get("byteOffset", 0)
    elem = dtype.itemsize * ncomp
    stride = view.get("byteStride", elem)
    raw = np.frombuffer(buf, np.uint8, count=(count - 1) * stride + elem if count else 0, offset=start)
    rows = np.lib.stride_tricks.as_strided(raw, (count, elem), (stride, 1)) if count else raw.reshape(0, elem)
    vals = np.ascontiguousarray(rows).view(dtype).reshape(count, ncomp)
    if a.get("normalized", False):
        if dtype == np.uint8:
            return vals.astype(np.float32) / np.float32(255)
        if dtype == np.uint16:
            return vals.astype(np.float32) / np.float32(65535)
        if dtype == np.int8:
            return np.maximum(vals.astype(np.float32) / np.float32(127), np.float32(-1))
        if dtype == np.int16:
            return np.maximum(vals.astype(np.float32) / np.float32(32767), np.float32(-1))
    if ncomp == 1 and dtype.kind in "iu":
        return vals[:, 0].astype(np.uint32)
    return vals.astype(np.float32)


def _local_matrix(node: dict) -> np.ndarray:
    """Row-major 4x4 of a node: its matrix (column-major in the file) or T * R * S."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T
    t = np.asarray(node.get("translation", [0, 0, 0]), np.float64)
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * s[None, :]
    m[:3, 3] = t
    return m


def _material(doc: dict, index, texture_of_image: dict) -> Material:
    if index is None:
        return Material(base_color=(1.0, 1.0, 1.0), roughness=1.0, metallic=1.0)
    m = doc["materials"][index]
    pbr = m.get("pbrMetallicRoughness", {})
    mode = m.get("alphaMode", "OPAQUE")
    if mode == "MASK" and abs(m.get("alphaCutoff", 0.5) - 0.5) > 1e-6:
        raise ValueError(f"material {index}: alphaCutoff {m['alphaCutoff']} (only 0.5 is supported, as in the reference)")
    mat = Material(base_color=tuple(pbr.get("baseColorFactor", [1, 1, 1, 1])[:3]),
                   roughness=float(pbr.get("roughnessFactor", 1.0)), metallic=float(pbr.get("metallicFactor", 1.0)),
                   alpha_clip=mode == "MASK")
    for field, tex in (("base_color_texture", pbr.get("baseColorTexture")),
                       ("metallic_roughness_texture", pbr.get("metallicRoughnessTexture")),
                       ("normal_texture", m.get("normalTexture"))):
        if tex is None:
            continue
        image = doc["textures"][tex["index"]].get("source")
        if image in texture_of_image:
            setattr(mat, field, texture_of_image[image])
        tt = tex.get("extensions", {}).get("KHR_texture_transform")
        if tt is not None:
            mat.texture_scale = tuple(tt.get("scale", [1.0, 1.0]))
            mat.texture_offset = tuple(tt.get("offset", [0.0, 0.0]))
    return mat


def _reference_node_matrix(node: dict) -> np.ndarray:
    """Row-major scale(S) * translate(T) of a TRS node (identity for a matrix
    node), as the reference's NodeTree forms it (mesh_loading.cpp:16-35)."""
    if "matrix" in node:
        return np.eye(4)
    t = np.asarray(node.get("translation", [0, 0, 0]), np.float64)
    s = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    m = np.eye(4)
    m[:3, :3] = np.diag(s)
    m[:3, 3] = s * t
    return m


def _image_texture(path: Path, image: dict):
    uri = image.get("uri")
    if uri is None or uri.startswith("data:"):
        return None
    suffix = Path(uri).suffix.lower()
    if suffix == ".ktx2":
        return load_ktx2_texture(path.parent / uri)
    if suffix == ".dds":
        return load_dds_texture(path.parent / uri)
    return None


def load_gltf(path, *, vertex_format: int = abi.VERTEX_FORMAT_FP32, images: dict | None = None,
              scene: int | None = None, node_transforms: str = "gltf") -> HostScene:
    """A HostScene of the glTF file's default (or given) scene; see the module
    docstring for what is read."""
    if node_transforms not in ("gltf", "reference"):
        raise ValueError(f"node_transforms {node_transforms!r}: 'gltf' or 'reference'")
    path = Path(path)
    doc, buffers = _read_container(path)
    supplied = images or {}
    textures, texture_of_image = [], {}
    for i, img in enumerate(doc.get("images", [])):
        tex = supplied.get(i)
        if tex is None:
            tex = _image_texture(path, img)
        elif not isinstance(tex, EncodedTexture):
            tex = np.ascontiguousarray(tex, np.uint8)
        if tex is not None:
            texture_of_image[i] = len(textures)
            textures.append(tex)
    meshes, mesh_of, materials, material_of, instances = [], {}, [], {}, []

    def primitive_mesh(mi: int, pi: int) -> int:
        key = (mi, pi)
        if key not in mesh_of:
            prim = doc["meshes"][mi]["primitives"][pi]
            attrs = prim["attributes"]
            pos = _accessor(doc, buffers, attrs["POSITION"])[:, :3]
            nrm = _accessor(doc, buffers, attrs["NORMAL"])[:, :3] if "NORMAL" in attrs else None
            uv = _accessor(doc, buffers, attrs["TEXCOORD_0"])[:, :2] if "TEXCOORD_0" in attrs else None
            if "indices" in prim:
                idx = _accessor(doc, buffers, prim["indices"]).astype(np.uint32)
            else:
                idx = np.arange(len(pos), dtype=np.uint32)
            mesh_of[key] = len(meshes)
            meshes.append((pos, nrm, uv, idx.reshape(-1, 3)))
        return mesh_of[key]

    def walk(node_index: int, parent: np.ndarray) -> None:
        world = parent @ _local_matrix(doc["nodes"][node_index])
        walk_one(node_index, world)
        for child in doc["nodes"][node_index].get("children", []):
            walk(child, world)

    def walk_one(node_index: int, world: np.ndarray) -> None:
        node = doc["nodes"][node_index]
        if "mesh" in node:
            for pi, prim in enumerate(doc["meshes"][node["mesh"]]["primitives"]):
                if prim.get("mode", 4) != 4:
                    continue  # points and lines are not drawn
                mat = prim.get("material")
                if mat not in material_of:
                    material_of[mat] = len(materials)
                    materials.append(_material(doc, mat, texture_of_image))
                instances.append((primitive_mesh(node["mesh"], pi), material_of[mat],
                                  np.ascontiguousarray(world.T, np.float32)))  # column-major

    scenes = doc.get("scenes", [])
    s = scene if scene is not None else doc.get("scene", 0)
    if node_transforms == "reference":
        for ni, node in enumerate(doc.get("nodes", [])):
            if "mesh" in node:
                walk_one(ni, _reference_node_matrix(node))
    else:
        roots = scenes[s]["nodes"] if scenes else list(range(len(doc.get("nodes", []))))
        for r in roots:
            walk(r, np.eye(4))
    if not instances:
        raise ValueError(f"{path}: no triangle primitives in scene {s}")
    return HostScene.build(meshes, materials, instances, textures=textures, vertex_format=vertex_format)
