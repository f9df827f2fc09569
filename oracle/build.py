"""Builds the scalar oracle (test infrastructure) into oracle/liblh_oracle.so.

Same floating-point contract as the kernels: no FMA contraction, IEEE
division and square root (x86-64 SSE), so bit-exact results are comparable.

The reference itself (GLSL shaders run through Vulkan) cannot be compiled in
this image: there is no glslc / glslang, no Vulkan loader and no lavapipe ICD,
so there is no oracle/_ref build.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent


def build() -> Path:
    out = HERE / "liblh_oracle.so"
    cmd = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           f"-I{ROOT / 'include'}", f"-I{HERE}", str(HERE / "oracle.cpp"), str(HERE / "oracle64.cpp"), "-o",
           str(out)]
    cmd = [c for c in cmd if not c.endswith("oracle64.cpp") or Path(c).exists()]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("oracle build failed")
    return out


if __name__ == "__main__":
    print(build())
