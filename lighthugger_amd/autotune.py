"""Autotuner for the raster / cull launch tunables (lh_tuning, include/lh/lh.h).

The reference's raster is fixed-function; its software replacement here has
free parameters that change speed but never results: the in-lane and
wave-cooperative triangle-size thresholds and the persistent grid sizes.
This module times whole frames (HIP events on the frame's stream) for a
representative workload of each resolution class and keeps, per class, the
fastest point of a coordinate-descent search.  The result is written to
lighthugger_amd/tuning.json, which lib.lib() applies when it loads the library.

Usage (on a GPU):  python -m lighthugger_amd.autotune [--classes 0,1,2,4]
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

from . import abi

TUNING_FILE = Path(__file__).resolve().parent / "tuning.json"

SPACE = {
    "small_extent_px": [2, 3, 4, 6, 8],
    "medium_extent_px": [8, 12, 16, 32, 64],
    "meshlet_blocks": [4096, 8192, 16384],
    "big_blocks": [512, 2048],
    "clip_blocks": [64, 256],
    "cull_blocks": [512, 1024, 2048],
    "tile_item_entries": [32, 64, 128],
    "tile_blocks": [768, 1024, 2048],
    "tile_raster": [0, 1],
    "shadow_small_extent_px": [0, 4, 6, 8],
    "shadow_medium_extent_px": [0, 8, 16, 32],
}
FIELDS = ("small_extent_px", "meshlet_blocks", "big_blocks", "clip_blocks", "cull_blocks", "medium_extent_px",
          "tile_item_entries", "tile_blocks", "tile_raster", "shadow_small_extent_px", "shadow_medium_extent_px")
# fields absent from older files (0: the cascades take the main view's extents)
DEFAULTS = {"tile_item_entries": 64, "tile_blocks": 1024, "tile_raster": 1, "shadow_small_extent_px": 0,
            "shadow_medium_extent_px": 0}


def _tuning(d: dict):
    from .lib import Tuning

    return Tuning(*[int(d.get(f, DEFAULTS.get(f, 0))) for f in FIELDS])


def apply_file(path: Path = TUNING_FILE) -> bool:
    """Applies the persisted per-class tunings; returns False if there is no file."""
    import ctypes as C

    from .lib import check, lib

    if not path.exists():
        return False
    data = json.loads(path.read_text())
    for cls, d in data.get("classes", {}).items():
        check(lib().lh_set_tuning_class(int(cls), C.byref(_tuning(d["tuning"]))), "lh_set_tuning_class")
    return True


def _set_class(cls: int, d: dict) -> None:
    import ctypes as C

    from .lib import check, lib

    check(lib().lh_set_tuning_class(cls, C.byref(_tuning(d))), "lh_set_tuning_class")


def _get_class(cls: int) -> dict:
    import ctypes as C

    from .lib import Tuning, check, lib

    t = Tuning()
    check(lib().lh_get_tuning_class(cls, C.byref(t)), "lh_get_tuning_class")
    return t.as_dict()


def _timer(run, frames: int) -> float:
    import torch

    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(3):
        e0.record()
        for _ in range(frames):
            run()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) / frames)
    return best


def search(cls: int, run, frames: int, sweeps: int = 2, log=print) -> dict:
    """Coordinate descent over SPACE starting from the class's current tuning."""
    cur = _get_class(cls)
    cur["medium_extent_px"] = max(cur["medium_extent_px"], cur["small_extent_px"])
    _set_class(cls, cur)
    best_ms = _timer(run, frames)
    log(f"class {cls}: start {best_ms:.4f} ms {cur}")
    for _ in range(sweeps):
        improved = False
        for field, values in SPACE.items():
            for val in values:
                if val == cur[field]:
                    continue
                cand = dict(cur, **{field: val})
                if cand["medium_extent_px"] < cand["small_extent_px"]:
                    continue
                shadow_small = cand["shadow_small_extent_px"] or cand["small_extent_px"]
                if cand["shadow_medium_extent_px"] and cand["shadow_medium_extent_px"] < shadow_small:
                    continue
                _set_class(cls, cand)
                ms = _timer(run, frames)
                if ms < best_ms * 0.995:
                    best_ms, cur, improved = ms, cand, True
                    log(f"class {cls}: {field}={val} -> {ms:.4f} ms")
            _set_class(cls, cur)
        if not improved:
            break
    _set_class(cls, cur)
    return {"tuning": cur, "ms": round(best_ms, 5)}


def workloads(classes):
    """(class, description, run, frames, finalizer) for each requested class."""
    from .frame import ViewConfig
    from .lib import BATCH_BATCHED, BatchRenderer, Renderer
    from .scene import HostScene, batch_cameras, make_camera, preset_camera

    scene = HostScene.preset(abi.PRESET_LARGE)
    base = preset_camera(abi.PRESET_LARGE)
    sizes = {0: (640, 640, 640), 1: (1920, 1080, 2048), 2: (3840, 2160, 4096), 3: (4096, 4096, 4096)}
    for cls in classes:
        if cls == 4:
            br = BatchRenderer(scene, batch_cameras(4, 64, 256, 2048), big_capacity=1 << 20)
            yield cls, "64 cameras, sides 256-2048, own 4 cascades", lambda br=br: br.run(BATCH_BATCHED), 2, br.close
            continue
        w, h, s = sizes[cls]
        cam = make_camera(base.position.tolist(), base.target.tolist(), w, h)
        r = Renderer(scene, ViewConfig(w, h, s, big_capacity=1 << 21))
        r.set_camera(cam)
        yield cls, f"{w}x{h} + 4x{s}^2", r.frame, 10, (lambda: None)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default="0,1,2,4")
    ap.add_argument("--out", default=str(TUNING_FILE))
    args = ap.parse_args()
    classes = [int(c) for c in args.classes.split(",") if c]
    out = {"device": None, "written": time.strftime("%Y-%m-%d %H:%M:%S"), "classes": {}}
    import torch

    out["device"] = torch.cuda.get_device_name(0)
    for cls, desc, run, frames, fin in workloads(classes):
        res = search(cls, run, frames)
        res["workload"] = desc
        out["classes"][str(cls)] = res
        fin()
    # a class that was not searched inherits its nearest searched neighbour
    for cls in range(5):
        if str(cls) not in out["classes"]:
            near = min(out["classes"], key=lambda c: abs(int(c) - cls) if int(c) != 4 else 99)
            out["classes"][str(cls)] = dict(out["classes"][near], inherited_from=int(near))
    Path(args.out).write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
