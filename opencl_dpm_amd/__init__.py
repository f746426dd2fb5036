"""MI355X-native Deformable Particle Model timestep.

`clDPM` (built in-tree as opencl_dpm_amd/clDPM*.so) is the drop-in for the
reference's pybind11 module: same classes Cell2D, Tissue2D, Cell3D,
Tissue3D, attributes and methods.  Reference scripts run unchanged with
this directory on sys.path (``import clDPM``), or use
``from opencl_dpm_amd import clDPM``.

`opencl_dpm_amd.capi` binds the C ABI (include/dpm.h) directly for batched
tissues and timing.  Both load libdpm_hip.so, the HIP implementation; there
is no CPU fallback.
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


clDPM = _load_module()

__all__ = ["clDPM"]
