"""lh_frame_noalloc: the frame with its raster scratch in the library's own
device array, against lh_frame with a caller-owned workspace, through the C
ABI.  Two calls in a row, the second with a smaller view: keys, shadow depths,
radiance and BGRA8 bit-identical to lh_frame's, and the first call's larger
contents left in the shared array do not leak into the second."""
import ctypes as C

import numpy as np
import pytest

from lighthugger_amd import abi
from lighthugger_amd.frame import ViewConfig
from lighthugger_amd.scene import HostScene, make_camera
from oracle.compare import snapshot

pytestmark = pytest.mark.gpu

EXACT = ("visbuffer", "shadow", "radiance", "ldr", "misc")


def _frame(r, noalloc):
    import torch

    from lighthugger_amd.lib import check, current_stream, lib

    if noalloc:
        check(lib().lh_frame_noalloc(r.scene.address, r.view.address, current_stream()), "lh_frame_noalloc")
    else:
        r.frame()
    torch.cuda.synchronize()
    return snapshot(r.view)


def _same(a, b, what):
    for k in EXACT:
        if k == "misc":
            assert bytes(np.ascontiguousarray(a[k]).view(np.uint8)) == bytes(np.ascontiguousarray(b[k]).view(np.uint8)), \
                f"{what}: misc storage differs"
        else:
            assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


def test_noalloc_frame_equals_lh_frame_two_calls(tiled_raster):
    from lighthugger_amd.lib import Renderer

    scene = HostScene.preset(abi.PRESET_SMALL, num_instances=48)
    cases = [make_camera((0, 20, 80), (0, 0, 0), 640, 360, near=0.1),
             make_camera((10, 15, 60), (0, 0, 0), 200, 120, near=0.1)]
    sizes = [512, 256]
    for i, (cam, s) in enumerate(zip(cases, sizes)):
        print(f"case {i}: {cam.width}x{cam.height}, shadow {s}", flush=True)
        cfg = ViewConfig(cam.width, cam.height, s, big_capacity=1 << 16, clip_capacity=1 << 12)
        r = Renderer(scene, cfg)
        need = r.view.sizes["scratch"]
        assert need <= abi.STATIC_SCRATCH_BYTES, need
        r.set_camera(cam)
        ref = _frame(r, noalloc=False)
        # scribble over the caller's scratch: the no-alloc frame must not read it
        t = r.view.handle[0]
        o, n = r.view.offsets["scratch"], r.view.sizes["scratch"]
        t[o: o + n].fill_(0xA5)
        got = _frame(r, noalloc=True)
        _same(ref, got, f"call {i}")
        covered = int((got["visbuffer"] != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum())
        print(f"case {i}: covered {covered}", flush=True)
        assert covered > 500


def test_noalloc_rejects_views_over_the_array():
    from lighthugger_amd.lib import Renderer, current_stream, lib

    scene = HostScene.preset(abi.PRESET_SMALL, num_instances=8)
    cam = make_camera((0, 20, 80), (0, 0, 0), 320, 180, near=0.1)
    r = Renderer(scene, ViewConfig(cam.width, cam.height, 256, big_capacity=1 << 21))  # 100 MB big queue
    r.set_camera(cam)
    assert abi.scratch_bytes(cam.width, cam.height, r.view.dcap, 1 << 21, r.view.cfg.clip_capacity,
                             r.view.bcap) > abi.STATIC_SCRATCH_BYTES
    assert lib().lh_frame_noalloc(r.scene.address, r.view.address, current_stream()) == abi.LH_ERR_CAPACITY
    assert lib().lh_frame_noalloc(None, r.view.address, current_stream()) == abi.LH_ERR_INVALID
    assert lib().lh_frame_noalloc(r.scene.address, C.c_void_p(0), current_stream()) == abi.LH_ERR_INVALID
