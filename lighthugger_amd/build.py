"""Build recipe for the native libraries (no cmake/ninja needed).

Outputs (in-tree, git-ignored, shipped with the tree):
  lighthugger_amd/liblh_scene.so  host scene producer (g++, C++20)
  lighthugger_amd/liblh.so        HIP kernels + C ABI + frame recorder (hipcc, gfx950)
  lighthugger_amd/_lh_torch.so    libtorch autograd Function for the resolve (hipcc)
  lighthugger_amd/liblh_probe.so  device ceilings for the roofline (atomics, plain copy)
The oracle (test infrastructure) is built by oracle/build.py.

Floating-point flags are part of the contract: both the kernels and the
scalar oracle are compiled without FMA contraction and with correctly rounded
fp32 division and square root, so that every culling decision and every
visibility-buffer key is reproducible bit for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_SOURCES = [
    "cull.hip",
    "raster.hip",
    "tile.hip",
    "hiz.hip",
    "depth.hip",
    "resolve.hip",
    "backward.hip",
    "capi.hip",
    "batch.hip",
    "noalloc.hip",
    "graph.hip",
]

# Device code: gfx950 only; no FMA contraction; IEEE div/sqrt; packed fp32
# arithmetic disabled (some v_pk_*_f32 forms have given wrong results on
# these GPUs without any error).
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd[:4])} ...")


def build_scene_lib() -> Path:
    out = PKG / "liblh_scene.so"
    _run(["g++", *HOST_FLAGS, "-shared", "-pthread", str(CSRC / "host" / "scene.cpp"), "-o", str(out)])
    return out


def build_hip_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    srcs = [s for s in HIP_SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = PKG / "liblh.so"
    tmp = PKG / "liblh.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_probe_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    obj = BUILD / "probe.o"
    _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / "probe.hip"), "-o", str(obj)])
    out = PKG / "liblh_probe.so"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", str(obj), "-o", str(out)])
    return out


def build_all() -> list[Path]:
    outs = [build_scene_lib(), build_hip_lib(), build_probe_lib()]
    torch_src = CSRC / "torch_ext.cpp"
    if torch_src.exists():
        from . import build_torch  # noqa: PLC0415

        outs.append(build_torch.build())
    return outs


if __name__ == "__main__":
    for p in build_all():
        print(p)
