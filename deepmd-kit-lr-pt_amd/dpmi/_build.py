"""Builds libdpmi.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m dpmi._build            # from deepmd-kit-lr-pt_amd/

Objects go to deepmd-kit-lr-pt_amd/_build/, the library to
deepmd-kit-lr-pt_amd/dpmi/libdpmi.so.  Every HIP translation unit is
compiled with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libdpmi.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("DPMI_ARCH", "gfx950")

HIPCC = os.path.join(ROCM, "bin", "hipcc")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{ROCM}/include"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,--no-undefined", "-o", tmp] + objs + [
        f"-L{ROCM}/lib", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


TESTS_CPP = os.path.join(os.path.dirname(ROOT), "tests", "cpp")
HARNESS = os.path.join(TESTS_CPP, "_bin", "dpmi_harness")
MD_DRIVER = os.path.join(TESTS_CPP, "_bin", "dpmi_md")


def _build_caller(src: str, out: str) -> None:
    os.makedirs(os.path.dirname(out), exist_ok=True)
    inc = os.path.join(os.path.dirname(ROOT), "include")
    cmd = [HIPCC, "-std=c++17", "-O2", "-Wall", f"-I{inc}", src, "-o", out + ".tmp", f"-L{PKG}", "-ldpmi",
           f"-Wl,-rpath,{PKG}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(out + ".tmp", out)


def build_harness(verbose: bool = False) -> str:
    """The C++ test callers of the drop-in library, linked against
    libdpmi.so: tests/cpp/harness.cpp (the deepmd:: templates and DeepPot)
    and tests/cpp/md_driver.cpp (NVE dynamics through the LAMMPS pair
    style)."""
    for src, out in ((os.path.join(TESTS_CPP, "harness.cpp"), HARNESS),
                     (os.path.join(TESTS_CPP, "md_driver.cpp"), MD_DRIVER)):
        _build_caller(src, out)
        if verbose:
            print("built", out)
    return HARNESS


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
