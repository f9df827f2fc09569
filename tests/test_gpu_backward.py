"""GPU tests of the differentiable resolve (lh_resolve_ex, lh_resolve_backward and
the libtorch autograd Function) against the FP64 dual-number adjoint of the
oracle.

Bar: per parameter group (vertex positions, vertex normals, base-colour,
roughness and metallic factors, sun direction, sun intensity) the FP32 GPU
gradient is within 1e-4 of the FP64 adjoint in relative L2 norm,
||g_gpu - g_64|| <= 1e-4 * ||g_64||.
"""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, make_camera, make_uniforms, preset_camera
from oracle import adjoint as A
from oracle import lib as oracle_lib
from oracle.compare import oracle_view, snapshot

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4


def _setup(width=256, height=192, shadow=512, tri_scale_pct=10, vertex_format=abi.VERTEX_FORMAT_FP32,
           texture_format=abi.TEXTURE_RGBA8):
    import torch

    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_LARGE, texture_size=64, tri_scale_pct=tri_scale_pct,
                             vertex_format=vertex_format, texture_format=texture_format)
    base = preset_camera(abi.PRESET_LARGE)
    cam = make_camera(base.position.tolist(), base.target.tolist(), width, height)
    cfg = ViewConfig(width, height, shadow, big_capacity=1 << 20)
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    r.frame()
    torch.cuda.synchronize()
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    oracle_lib().orc_frame(sd.address, vd.address)
    gpu, orc = snapshot(r.view), snapshot(vd)
    assert np.array_equal(gpu["visbuffer"], orc["visbuffer"]) and np.array_equal(gpu["shadow"], orc["shadow"])
    print(f"case: LARGE x{tri_scale_pct}% {width}x{height}, covered "
          f"{int((gpu['visbuffer'] != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum())}", flush=True)
    return scene, r, sd, vd


def _groups(scene, sd, g_geo):
    """Position and normal components of every mesh: the arena words from
    MeshInfo.positions to .normals and from .normals to .uvs."""
    base = sd.geometry_base
    pos, nrm = set(), set()
    for m in sd.mesh_infos:
        p0, n0, u0 = ((int(m[k]) - base) // 4 for k in ("positions", "normals", "uvs"))
        pos.add((p0, n0))
        nrm.add((n0, u0))
    def gather(spans):
        return np.concatenate([g_geo[a:b] for a, b in sorted(spans)])
    return gather(pos), gather(nrm)


def _rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_resolve_ex_with_scene_values_is_lh_resolve():
    import torch

    from lighthugger_amd.diff import DifferentiableResolve

    scene, r, sd, vd = _setup()
    ref = r.view.radiance().copy()
    d = DifferentiableResolve(r)
    d.forward_raw(d.parameters())
    torch.cuda.synchronize()
    assert np.array_equal(ref, r.view.radiance())
    out = d(**d.parameters()).cpu().numpy()
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("texture_format", [abi.TEXTURE_RGBA8, abi.TEXTURE_BC1])
def test_backward_matches_fp64_adjoint(texture_format):
    import torch

    from lighthugger_amd.diff import DifferentiableResolve

    scene, r, sd, vd = _setup(texture_format=texture_format)
    d = DifferentiableResolve(r)
    p = d.parameters()
    rng = np.random.default_rng(5)
    g = rng.standard_normal((r.view.cfg.height, r.view.cfg.width, 4)).astype(np.float32)
    out = d.raw_backward(p, torch.from_numpy(g).cuda())
    torch.cuda.synchronize()
    assert not d.unsupported()
    fz = A.frozen(sd, vd)
    g64 = A.grad64(sd, vd, A.scene_params(sd, vd), fz, g)
    got = {k: v.cpu().numpy().astype(np.float64) for k, v in out.items()}
    pos64, nrm64 = _groups(scene, sd, g64["geometry"])
    pos, nrm = _groups(scene, sd, got["geometry"])
    errs = {"positions": _rel(pos, pos64), "normals": _rel(nrm, nrm64)}
    for k in ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity"):
        errs[k] = _rel(got[k], g64[k])
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert np.linalg.norm(pos64) > 0 and np.linalg.norm(nrm64) > 0
    for k, e in errs.items():
        assert e <= REL_L2, (k, e)
    # the same with the fixed quantities recorded by the forward
    d.forward_raw(p)
    rec = d.raw_backward(p, torch.from_numpy(g).cuda(), recorded=True)
    torch.cuda.synchronize()
    got_r = {k: v.cpu().numpy().astype(np.float64) for k, v in rec.items()}
    for k in ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity", "geometry"):
        assert _rel(got_r[k], got[k]) <= 1e-6, k
    # a record without its valid bits (zero-filled, as a fresh or foreign
    # buffer may be) is not trusted: every pixel recomputes the fixed quantities
    d.frozen.zero_()
    rec0 = d.raw_backward(p, torch.from_numpy(g).cuda(), recorded=True)
    torch.cuda.synchronize()
    got_0 = {k: v.cpu().numpy().astype(np.float64) for k, v in rec0.items()}
    for k in ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity", "geometry"):
        assert _rel(got_0[k], got[k]) <= 1e-6, k


def test_autograd_function_l2_loss():
    """loss = ||radiance - target||^2 through the libtorch autograd Function:
    its parameter gradients equal lh_resolve_backward's with grad 2 (r - t),
    and the FP64 adjoint's within the same bar."""
    import torch

    from lighthugger_amd.diff import DifferentiableResolve

    scene, r, sd, vd = _setup(192, 128)
    d = DifferentiableResolve(r)
    p0 = d.parameters()
    with torch.no_grad():
        target = d(**p0)[..., :3].clone() * 0.9
    p = d.parameters(requires_grad=True)
    geom = d.geometry.detach().clone()
    geo_leaf = d.geometry.detach().requires_grad_(True)  # shares the scene's memory
    try:
        rad = d(**p, geometry=geo_leaf)
        loss = ((rad[..., :3] - target) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        gr = torch.zeros_like(rad)
        gr[..., :3] = 2.0 * (rad.detach()[..., :3] - target)
        ref = d.raw_backward(p0, gr)
        torch.cuda.synchronize()
        for k in ("base_color", "roughness", "metallic", "sun_dir", "sun_intensity"):
            assert _rel(p[k].grad.cpu().numpy(), ref[k].cpu().numpy()) <= 1e-6, k
        assert _rel(geo_leaf.grad.cpu().numpy(), ref["geometry"].cpu().numpy()) <= 1e-6
        g64 = A.grad64(sd, vd, A.scene_params(sd, vd), A.frozen(sd, vd), gr.cpu().numpy())
        assert _rel(p["sun_intensity"].grad.cpu().numpy(), g64["sun_intensity"]) <= REL_L2
        assert _rel(p["base_color"].grad.cpu().numpy(), g64["base_color"]) <= REL_L2
    finally:
        assert torch.equal(geom, d.geometry)


def test_quantized_meshes_are_reported_unsupported():
    """The adjoint covers FP32 vertices only: on a scene in the quantized
    formats it flags LH_FLAG_BACKWARD_UNSUPPORTED and adds no gradient."""
    import torch

    from lighthugger_amd.diff import DifferentiableResolve

    scene, r, sd, vd = _setup(128, 96, 256, tri_scale_pct=5, vertex_format=abi.VERTEX_FORMAT_QUANTIZED)
    d = DifferentiableResolve(r)
    p = d.parameters()
    out = d.raw_backward(p, torch.ones(96, 128, 4, device="cuda"))
    torch.cuda.synchronize()
    assert d.unsupported()
    for k, t in out.items():
        assert not t.any().item(), k
