"""Build the gfx950 HIP library ``torchff/libtorchff_hip.so`` from ``csrc/*.hip`` and the operator
library ``torchff/libtorchff_ops.so`` (``csrc/torch_ops.cpp``: ``torch.ops.torchff.*`` registered in C++
over the C ABI, loadable by a libtorch-only process).

Usage: ``python build_lib.py`` (or ``__graft_entry__.build()``).  Compiles the HIP code with
hipcc for gfx950 only, with packed-FP32 instruction selection disabled; the operator library
is host C++ (g++) against the installed torch.
"""

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "torchff", "libtorchff_hip.so")
OPS_OUT = os.path.join(HERE, "torchff", "libtorchff_ops.so")

FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result", "-Wl,-soname,libtorchff_hip.so",
]


def build(verbose=False):
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
           *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtorchff_hip.so")
    os.replace(tmp, OUT)
    build_ops(verbose)
    return OUT


def build_ops(verbose=False):
    """libtorchff_ops.so: host C++ against torch's headers and libraries, linked to libtorchff_hip.so
    (found next to it at run time)."""
    import torch
    tdir = os.path.dirname(torch.__file__)
    tlib = os.path.join(tdir, "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tmp = OPS_OUT + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__", "-DUSE_ROCM",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(tdir, "include"), "-I", os.path.join(tdir, "include", "torch", "csrc", "api", "include"),
           "-I", "/opt/rocm/include", os.path.join(HERE, "csrc", "torch_ops.cpp"),
           "-L", os.path.dirname(OUT), "-l:libtorchff_hip.so", "-L", tlib, "-lc10", "-ltorch", "-ltorch_cpu",
           "-lc10_hip", "-ltorch_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + tlib, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building libtorchff_ops.so")
    os.replace(tmp, OPS_OUT)
    return OPS_OUT


if __name__ == "__main__":
    print(build(verbose=True))
