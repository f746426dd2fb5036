"""MI355X-native Deformable Particle Model timestep.

`clDPM` (built in-tree as opencl_dpm_amd/clDPM*.so) is the drop-in for the
reference's pybind11 module: same classes Cell2D, Tissue2D, Cell3D,
Tissue3D, attributes and methods.  Reference scripts run unchanged with
this directory on sys.path (``import clDPM``), or use
``from opencl_dpm_amd import clDPM``.

`opencl_dpm_amd.capi` binds the C ABI (include/dpm.h) directly for batched
tissues and timing.  Both load libdpm_hip.so, the HIP implementation; there
is no CPU fallback.

The module is loaded on first access of ``opencl_dpm_amd.clDPM`` so that the
build recipe (``opencl_dpm_amd.build``) can be imported before anything is
built.
"""
from __future__ import annotations

import importlib
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load_module():
    if _HERE not in sys.path:
        sys.path.insert(0, _HERE)
    try:
        return importlib.import_module("clDPM")
    except ImportError as e:
        raise ImportError(
            "clDPM extension is not built in-tree; run `python -m opencl_dpm_amd.build`") from e


def __getattr__(name):
    if name == "clDPM":
        mod = _load_module()
        globals()["clDPM"] = mod
        return mod
    raise AttributeError(f"module 'opencl_dpm_amd' has no attribute {name!r}")


__all__ = ["clDPM"]
