"""Build recipe for the in-tree native code.

    python -m opencl_dpm_amd.build

1. ``libdpm_hip.so``  — the C ABI (include/dpm.h) and every HIP kernel,
   compiled by hipcc for gfx950 only;
2. ``clDPM<ext>.so``  — the pybind11 module with the reference's Python
   surface, compiled by g++ and linked against libdpm_hip.so (rpath $ORIGIN);
3. the CPU oracle used by the tests (``make -C oracle``), and, where the
   reference checkout exists, its host sources built into oracle/_ref.

Everything lands next to the sources so that it travels with the tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Float semantics the kernels rely on: no contraction (every a*b+c is two
# rounded operations, as in the CPU oracle), IEEE float divide and sqrt.
# Packed FP32 instructions are disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC",
    "-Wall",
]

LIB_NAME = "libdpm_hip.so"
EXT_SUFFIX = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
MODULE_NAME = "clDPM" + EXT_SUFFIX


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_library() -> str:
    out = os.path.join(PKG, LIB_NAME)
    src = os.path.join(CSRC, "dpm_capi.hip")
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}", src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_module() -> str:
    import pybind11

    out = os.path.join(PKG, MODULE_NAME)
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
        f"-I{INCLUDE}", f"-I{CSRC}", f"-I{pybind11.get_include()}", f"-I{py_inc}",
        os.path.join(CSRC, "clDPM.cpp"), os.path.join(CSRC, "dpm_host.cpp"),
        f"-L{PKG}", "-ldpm_hip", "-Wl,-rpath,$ORIGIN",
        "-o", out + ".tmp",
    ]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> None:
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", oracle, "-B", "all"])
    ref = os.environ.get("DPM_REFERENCE_DIR", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src")):
        _run(["make", "-s", "-C", oracle, "-B", "ref", f"REFDIR={ref}"])


def build_all() -> None:
    build_library()
    build_module()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
