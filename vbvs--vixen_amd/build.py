"""Builds libvixen_hip.so (HIP kernels for gfx950 + the C ABI + the host builder).

Run as a script or call build(). Every source is compiled each time: this is the
"does it build" step, and it is cheap (a few seconds per file).

Numerics flags, all deliberate:
  -ffp-contract=off                         no FMA contraction (host and device), so
                                            the traversal matches a strict-FP32 CPU
                                            evaluation bit for bit
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt on the device
  -fno-gpu-flush-denormals-to-zero          keep FP32 denormals, as the host does
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed FP32 (v_pk_*_f32) instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OBJ = PKG / "_build"
LIB = PKG / "libvixen_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HOST_FLAGS = ["-O2", "-fPIC", "-std=c++2b", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
              "-pthread", f"-I{INCLUDE}", f"-I{CSRC}"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++20", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fno-fast-math",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", f"-I{INCLUDE}", f"-I{CSRC}"]

HOST_SOURCES = ["vx_builder.cpp", "vx_scenes.cpp", "vx_normals.cpp", "vx_cache.cpp", "vx_materials.cpp"]
# host code that calls the HIP runtime API (no kernels): the C++ ISVOStructure adapter
HIP_HOST_SOURCES = ["vx_svo.cpp"]
HIP_HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__=1", "-I/opt/rocm/include"]
HIP_SOURCES = ["vx_api.hip", "vx_batch.hip", "vx_albedo.hip", "vx_encode.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    # hipcc's host pass does not know the packed-fp32 feature name; drop that noise
    noise = ("packed-fp32-ops", "is not a recognized feature for this target")
    lines = [ln for ln in r.stdout.splitlines() if ln.strip() and not any(s in ln for s in noise)]
    if lines:
        print("\n".join(lines), file=sys.stderr)


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    objs: list[str] = []
    for src in HOST_SOURCES:
        out = OBJ / (Path(src).stem + ".o")
        cmd = [CXX, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(out))
    for src in HIP_HOST_SOURCES:
        out = OBJ / (Path(src).stem + ".o")
        cmd = [CXX, *HOST_FLAGS, *HIP_HOST_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(out))
    for src in HIP_SOURCES:
        out = OBJ / (Path(src).stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(out))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objs, "-pthread"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


TORCH_EXT = PKG / "vx_torch.so"


def build_torch_ext(verbose: bool = False) -> Path | None:
    """The libtorch autograd binding (vx_torch.so), linked against libvixen_hip.so."""
    try:
        import torch
        import pybind11
    except ImportError:
        return None
    import sysconfig
    tdir = Path(torch.__file__).resolve().parent
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [CXX, "-O2", "-fPIC", "-shared", "-std=c++17", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DTORCH_EXTENSION_NAME=vx_torch", "-DTORCH_API_INCLUDE_EXTENSION_H", "-w",
           f"-I{tdir / 'include'}", f"-I{tdir / 'include' / 'torch' / 'csrc' / 'api' / 'include'}",
           f"-I{sysconfig.get_paths()['include']}", f"-I{pybind11.get_include()}", f"-I{INCLUDE}",
           "-D__HIP_PLATFORM_AMD__=1", "-I/opt/rocm/include",
           str(CSRC / "vx_torch.cpp"), "-o", str(TORCH_EXT.with_suffix(".so.tmp")),
           f"-L{tdir / 'lib'}", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           f"-L{PKG}", "-l:libvixen_hip.so", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tdir / 'lib'}"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(TORCH_EXT.with_suffix(".so.tmp"), TORCH_EXT)
    return TORCH_EXT


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    LIB.unlink(missing_ok=True)
    TORCH_EXT.unlink(missing_ok=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    if "--no-torch" not in sys.argv:
        print(build_torch_ext(verbose="-v" in sys.argv))
