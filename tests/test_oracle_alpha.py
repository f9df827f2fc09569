"""Alpha-clip geometry in the oracle (MESH_INFO_FLAGS_ALPHA_CLIP): the
reference's alpha-clip pipelines are double-sided and discard samples whose
base-colour alpha is below 0.5 (src/pipelines.cpp:22-25,
src/shaders/rasterization.glsl:84-97, :130-145).

Known answers: a quad whose texture is transparent for u < 0.5 shows only
its opaque half, seen from the front and, mirrored, from the back; as opaque
geometry it shows whole from the front and not at all from the back.
"""
import numpy as np

from lighthugger_amd import abi
from lighthugger_amd.scene import HostScene, Material, make_uniforms
from oracle.compare import oracle_frame

from .fixtures import alpha_scene
from .util import fixture_camera, quad_mesh


def _half_texture(n=64):
    t = np.full((n, n, 4), 255, np.uint8)
    t[:, : n // 2, 3] = 0  # columns with u < 0.5 transparent
    return t


def _frame(alpha_clip, position):
    mat = Material(base_color=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0, base_color_texture=0,
                   alpha_clip=alpha_clip)
    scene = HostScene.build([quad_mesh(1.0)], [mat], [(0, 0, np.eye(4, dtype=np.float32))],
                            textures=[_half_texture()])
    cam, cfg = fixture_camera(128, 96, position=position, target=(0, 0, 0))
    f = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))
    return f["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)


def test_alpha_clip_discards_transparent_half_front_and_back():
    front_opaque = _frame(False, (0, 0, 3))
    front_alpha = _frame(True, (0, 0, 3))
    back_opaque = _frame(False, (0, 0, -3))
    back_alpha = _frame(True, (0, 0, -3))
    n = int(front_opaque.sum())
    assert n > 2000
    assert not back_opaque.any()  # single-sided: the back face is culled
    assert (front_alpha <= front_opaque).all()
    for covered in (front_alpha, back_alpha):
        assert abs(int(covered.sum()) - n / 2) <= 0.03 * n
    cols = np.nonzero(front_opaque.any(axis=0))[0]
    mid = (cols[0] + cols[-1] + 1) / 2
    # +x (u = 1) is on the right seen from +z and on the left seen from -z
    xs_front = np.nonzero(front_alpha.any(axis=0))[0]
    xs_back = np.nonzero(back_alpha.any(axis=0))[0]
    assert xs_front.min() >= mid - 2 and xs_back.max() <= mid + 2


def test_alpha_scene_keeps_fewer_samples_than_opaque_cards():
    """The fixture scene: alpha-clipped cards cover fewer pixels than the same
    cards drawn opaque and double-sided would; every key names a real draw."""
    scene, cam, cfg = alpha_scene()
    f = oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances))
    covered = f["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)
    assert covered.sum() > 1000
    assert (scene.mesh_infos["flags"] & abi.MESH_INFO_FLAGS_ALPHA_CLIP).any()
    # shadows are written by the cards too
    assert (f["shadow"] != np.uint32(0x3F800000)).any()
