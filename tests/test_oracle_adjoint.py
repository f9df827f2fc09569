"""The FP64 dual-number adjoint of the resolve (oracle/oracle64.cpp) against
central finite differences of the FP64 resolve, and the FP64 resolve against
the FP32 oracle resolve.

Finite differences with h = 1e-3 (the configured step) agree with the dual
gradients to 1e-6 relative for the parameters the radiance is linear in
(base-colour and metallic factors, sun intensity).  For roughness, sun
direction and vertex positions/normals the radiance is curved at the scale
of h (small and grazing triangles), so the h = 1e-3 difference is compared at
a looser bar and the h -> 0 limit (h = 1e-6) at 1e-5: the adjoint is the
derivative, the discrepancy at h = 1e-3 is the truncation error of the
difference.
"""
import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, make_camera, make_uniforms, preset_camera
from oracle import adjoint as A
from oracle import lib as oracle_lib
from oracle.compare import oracle_view


@pytest.fixture(scope="module")
def case():
    scene = HostScene.preset(abi.PRESET_LARGE, num_instances=300, texture_size=64)
    base = preset_camera(abi.PRESET_LARGE)
    W, H = 96, 64
    cam = make_camera(base.position.tolist(), base.target.tolist(), W, H)
    cfg = ViewConfig(W, H, 256)
    sd, vd = oracle_view(scene, cfg, make_uniforms(cam, scene.num_instances))
    oracle_lib().orc_frame(sd.address, vd.address)
    fz = A.frozen(sd, vd)
    p = A.scene_params(sd, vd)
    g = np.random.default_rng(0).standard_normal((H, W, 4)).astype(np.float32)
    grad = A.grad64(sd, vd, p, fz, g)
    return sd, vd, fz, p, g, grad


def _fd(case, key, idx, h):
    sd, vd, fz, p, g, _ = case
    a = {k: v.copy() for k, v in p.items()}
    b = {k: v.copy() for k, v in p.items()}
    a[key].flat[idx] += h
    b[key].flat[idx] -= h
    la = float((A.resolve64(sd, vd, a, fz) * g[..., :3]).sum())
    lb = float((A.resolve64(sd, vd, b, fz) * g[..., :3]).sum())
    return (la - lb) / (2 * h)


def test_frozen_quantities_cover_the_view(case):
    sd, vd, fz, *_ = case
    kinds = np.bincount(fz["kind"], minlength=3)
    assert kinds[1] > 1000 and kinds[2] == 0
    assert set(np.unique(fz["flip"][fz["kind"] == 1])) <= {-1.0, 1.0}
    assert np.all((fz["shadow"] >= 0) & (fz["shadow"] <= 1))


def test_fp64_resolve_matches_fp32_oracle(case):
    sd, vd, fz, p, _, _ = case
    rad32 = vd.radiance()[..., :3].astype(np.float64)
    rad64 = A.resolve64(sd, vd, p, fz)
    m = fz["kind"].reshape(rad32.shape[:2]) == 1
    err = np.abs(rad64[m] - rad32[m]) / (np.abs(rad64[m]) + 1e-3)
    assert np.median(err) < 1e-6 and err.max() < 1e-2


@pytest.mark.parametrize("key", ["base_color", "metallic", "sun_intensity"])
def test_linear_parameters_match_fd(case, key):
    grad = case[5][key]
    idx = np.flatnonzero(grad)[:6]
    assert len(idx) > 0
    for i in idx:
        fd = _fd(case, key, i, 1e-3)
        assert abs(fd - grad.flat[i]) <= 1e-6 * max(abs(fd), 1e-3), (key, i, fd, grad.flat[i])


@pytest.mark.parametrize("key", ["roughness", "sun_dir", "geometry"])
def test_curved_parameters_match_fd(case, key):
    grad = case[5][key]
    nz = np.flatnonzero(grad)
    rng = np.random.default_rng(1)
    idx = nz if len(nz) <= 8 else rng.choice(nz, 8, replace=False)
    coarse, fine, ref = [], [], []
    for i in idx:
        coarse.append(_fd(case, key, i, 1e-3))
        fine.append(_fd(case, key, i, 1e-6))
        ref.append(grad.flat[i])
    coarse, fine, ref = map(np.asarray, (coarse, fine, ref))
    rel = lambda a: np.linalg.norm(a - ref) / np.linalg.norm(ref)  # noqa: E731
    assert rel(coarse) < 2e-2, (coarse, ref)
    assert rel(fine) < 1e-5, (fine, ref)
