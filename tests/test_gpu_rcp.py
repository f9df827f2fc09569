"""GPU: the raster's 1 / area (rcp_exact in lh_device.h: the hardware
reciprocal and one Newton step on fused multiply-adds) equals the IEEE
quotient 1.0f / a for every float a in [1, 2^30].

Triangles within 64 px have integer areas below 2^28, so the check covers
every value the in-lane, wave-cooperative and tiled rasters divide by: their
depths, and hence the visibility-buffer keys, are those of the correctly
rounded division the oracle uses.  The check runs the device function itself
(lhp_rcp_check in liblh_probe.so) over all 251,658,241 floats of the range."""
import ctypes
import struct
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def _bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_rcp_exact_equals_ieee_reciprocal_on_every_float_up_to_2_30():
    import torch

    path = ROOT / "lighthugger_amd" / "liblh_probe.so"
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(str(path))
    lib.lhp_rcp_check.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    lib.lhp_rcp_check.restype = ctypes.c_int
    lo, hi = _bits(1.0), _bits(float(1 << 30))
    assert (lo, hi) == (0x3F800000, 0x4E800000)
    print(f"case: rcp_exact over float bits [{lo:#x}, {hi:#x}]", flush=True)
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    assert lib.lhp_rcp_check(lo, hi, ctypes.c_void_p(count.data_ptr()), None) == 0
    torch.cuda.synchronize()
    assert int(count.item()) == 0
    # the argument checks reject an empty pointer and a reversed range
    assert lib.lhp_rcp_check(lo, hi, None, None) == 1
    assert lib.lhp_rcp_check(hi, lo, ctypes.c_void_p(count.data_ptr()), None) == 1
