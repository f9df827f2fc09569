"""glTF 2.0 ingest (lighthugger_amd/gltf.py; the reference's load_gltf,
src/resources/mesh_loading.cpp:241-738): a file written here with
KHR_mesh_quantization accessors (u16 positions dequantized by the node
transform, normalized i8 normals, normalized u16 uvs, u16 indices), a node
hierarchy, an alpha-mask material with KHR_texture_transform and a float
mesh, read back as .gltf (external .bin) and as .glb."""
import json
import struct

import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.gltf import load_gltf
from lighthugger_amd.scene import make_uniforms
from oracle.compare import oracle_frame

from .util import fixture_camera


def _doc(tmp_path):
    # quad on a u16 grid: x, y in {0, 65535} -> [-1, 1] through the child node's transform
    qpos = np.array([[0, 0, 0], [65535, 0, 0], [65535, 65535, 0], [0, 65535, 0]], np.uint16)
    qnrm = np.array([[0, 0, 127, 0]] * 4, np.int8)  # padded to 4 bytes per element
    quv = np.array([[0, 0], [65535, 0], [65535, 65535], [0, 65535]], np.uint16)
    qidx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    tri = np.array([[-0.5, -0.5, 0.3], [0.5, -0.5, 0.3], [0.0, 0.6, 0.3]], np.float32)
    parts, views, off = [], [], 0
    for arr, stride in ((qpos, 8), (qnrm, 4), (quv, 4), (qidx, None), (tri, None)):
        data = arr.tobytes() if stride != 8 else np.concatenate([arr, np.zeros((4, 1), np.uint16)], 1).tobytes()
        v = {"buffer": 0, "byteOffset": off, "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        views.append(v)
        parts.append(data + b"\0" * (-len(data) % 4))
        off += len(parts[-1])
    blob = b"".join(parts)
    doc = {
        "asset": {"version": "2.0"},
        "extensionsUsed": ["KHR_mesh_quantization", "KHR_texture_transform"],
        "buffers": [{"byteLength": len(blob)}],
        "bufferViews": views,
        "accessors": [
            {"bufferView": 0, "componentType": 5123, "count": 4, "type": "VEC3"},
            {"bufferView": 1, "componentType": 5120, "normalized": True, "count": 4, "type": "VEC3"},
            {"bufferView": 2, "componentType": 5123, "normalized": True, "count": 4, "type": "VEC2"},
            {"bufferView": 3, "componentType": 5123, "count": 6, "type": "SCALAR"},
            {"bufferView": 4, "componentType": 5126, "count": 3, "type": "VEC3"},
        ],
        "materials": [
            {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.7, 0.6, 1], "roughnessFactor": 0.4,
                                      "metallicFactor": 0.0,
                                      "baseColorTexture": {"index": 0, "extensions": {
                                          "KHR_texture_transform": {"scale": [2, 3], "offset": [0.25, 0.5]}}}},
             "alphaMode": "MASK", "doubleSided": True},
            {"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.3, 0.9, 1], "roughnessFactor": 0.7,
                                      "metallicFactor": 1.0}},
        ],
        "textures": [{"source": 0}],
        "images": [{"uri": "absent.png"}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3,
                                    "material": 0}]},
                   {"primitives": [{"attributes": {"POSITION": 4}, "material": 1}]}],
        "nodes": [
            {"translation": [0, 0, -0.5], "children": [1, 2]},
            {"mesh": 0, "translation": [-1, -1, 0], "scale": [2 / 65535] * 3},
            {"mesh": 1, "rotation": [0, 0, 0.25881905, 0.96592583]},  # 30 degrees about z
        ],
        "scenes": [{"nodes": [0]}],
        "scene": 0,
    }
    return doc, blob


def _write(tmp_path, binary: bool):
    doc, blob = _doc(tmp_path)
    if binary:
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        p = tmp_path / "scene.glb"
        p.write_bytes(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)
    else:
        (tmp_path / "scene.bin").write_bytes(blob)
        doc["buffers"][0]["uri"] = "scene.bin"
        p = tmp_path / "scene.gltf"
        p.write_text(json.dumps(doc))
    return p


def _tex():
    t = np.full((16, 16, 4), 255, np.uint8)
    t[:, :8, 3] = 0
    return t


@pytest.mark.parametrize("binary", [False, True])
def test_gltf_ingest(tmp_path, binary):
    s = load_gltf(_write(tmp_path, binary), images={0: _tex()})
    assert s.num_instances == 2 and len(s.mesh_infos) == 2
    mi = s.mesh_infos
    inst = s.instances
    # instance 0: the quad; world = parent (translate z -0.5) @ child (translate -1, -1; scale 2/65535)
    m = np.asarray(inst["transform"][0], np.float64).T
    local = s.positions(int(inst["mesh_info_address"][0]), 4).astype(np.float64)
    world = local @ m[:3, :3].T + m[:3, 3]
    corners = sorted(map(tuple, np.round(world, 5)))
    assert corners == sorted([(-1.0, -1.0, -0.5), (1.0, -1.0, -0.5), (1.0, 1.0, -0.5), (-1.0, 1.0, -0.5)])
    q = int(inst["mesh_info_address"][0])
    assert mi["flags"][q] & abi.MESH_INFO_FLAGS_ALPHA_CLIP
    np.testing.assert_allclose(mi["texture_scale"][q], [2, 3])
    np.testing.assert_allclose(mi["texture_offset"][q], [0.25, 0.5])
    np.testing.assert_allclose(s.normals(q, 4), np.tile([0, 0, 1], (4, 1)), atol=1e-6)
    assert sorted(map(tuple, s.uvs(q, 4))) == [(0, 0), (0, 1), (1, 0), (1, 1)]
    np.testing.assert_allclose(mi["base_color_factor"][q], [0.8, 0.7, 0.6], rtol=1e-6)
    assert int(mi["base_color_texture_index"][q]) == 0
    # instance 1: the float triangle rotated 30 degrees about z under the parent
    m1 = np.asarray(inst["transform"][1], np.float64).T
    c, sn = np.cos(np.pi / 6), np.sin(np.pi / 6)
    np.testing.assert_allclose(m1[:3, :3], [[c, -sn, 0], [sn, c, 0], [0, 0, 1]], atol=1e-6)
    np.testing.assert_allclose(m1[:3, 3], [0, 0, -0.5], atol=1e-7)
    o = int(inst["mesh_info_address"][1])
    assert not mi["flags"][o] & abi.MESH_INFO_FLAGS_ALPHA_CLIP
    assert s.factors["metallic_factor"][o] == 1.0
    # renders: the oracle draws both
    cam, cfg = fixture_camera(64, 64, position=(0, 0, 2.5))
    f = oracle_frame(s, cfg, make_uniforms(cam, s.num_instances))
    covered = f["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)
    ids = np.unique(s.meshlet_table["instance_index"][((f["visbuffer"][covered] >> np.uint64(7)) &
                                                        np.uint64(0xFFFFFF)).astype(np.int64)])
    assert set(ids.tolist()) == {0, 1}


def test_gltf_binary_and_text_agree(tmp_path):
    a = load_gltf(_write(tmp_path, False), vertex_format=abi.VERTEX_FORMAT_QUANTIZED)
    b = load_gltf(_write(tmp_path, True), vertex_format=abi.VERTEX_FORMAT_QUANTIZED)
    assert np.array_equal(a.geometry, b.geometry) and np.array_equal(a.instances, b.instances)
    assert np.array_equal(a.mesh_infos, b.mesh_infos)
