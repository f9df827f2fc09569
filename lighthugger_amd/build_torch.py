"""Build of the libtorch extension lighthugger_amd/_lh_torch*.so (the
differentiable resolve's torch::autograd::Function, csrc/torch_ext.cpp).

Compiled directly with g++ against torch's headers and libraries (no ninja,
no setuptools); links liblh.so with an $ORIGIN rpath, so the extension loads
the in-tree HIP library.
"""
from __future__ import annotations

import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent


def output_path() -> Path:
    return PKG / ("_lh_torch" + sysconfig.get_config_var("EXT_SUFFIX"))


def build() -> Path:
    import torch
    import torch.utils.cpp_extension as ce

    out = output_path()
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DTORCH_EXTENSION_NAME=_lh_torch", "-DTORCH_API_INCLUDE_EXTENSION_H", "-w",
           *[f"-I{p}" for p in ce.include_paths()], f"-I{sysconfig.get_paths()['include']}", f"-I{ROOT / 'include'}",
           str(PKG / "csrc" / "torch_ext.cpp"), "-o", str(out),
           *[f"-L{p}" for p in ce.library_paths()], "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           f"-L{PKG}", "-llh", "-Wl,-rpath,$ORIGIN"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout[-8000:])
        raise RuntimeError("torch extension build failed")
    return out


if __name__ == "__main__":
    print(build())
