"""Builds libdeepmd_op_pt.so: the torch.ops.deepmd op library over libdpmi.so.

    python -m dpmi.torch_op_build

Loaded with ``torch.ops.load_library`` (see dpmi.torch_ops), as the
reference's deepmd/pt/cxx_op.py loads its own op library.
"""
from __future__ import annotations

import os
import subprocess

import torch.utils.cpp_extension as ce

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(ROOT, "csrc", "torch", "tabulate_op.cpp")
OUT = os.path.join(PKG, "libdeepmd_op_pt.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build(verbose: bool = False) -> str:
    import torch

    inc = []
    for p in ce.include_paths():
        inc += ["-I", p]
    libdirs = ce.library_paths()
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = (["g++", "-std=c++17", "-O2", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
            "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", f"-I{ROCM}/include"] + inc + [SRC, "-o", OUT + ".tmp"]
           + [f"-L{d}" for d in libdirs] + ["-lc10", "-ltorch", "-ltorch_cpu", "-lc10_hip", "-ltorch_hip",
                                            f"-L{PKG}", "-ldpmi", f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{ROCM}/lib",
                                            f"-L{ROCM}/lib", "-lamdhip64"]
           + [f"-Wl,-rpath,{d}" for d in libdirs])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch op build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
