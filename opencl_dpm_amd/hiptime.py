"""HIP event timing on a given stream (ctypes on the HIP runtime)."""
from __future__ import annotations

import ctypes
import os

_rt = None


def runtime() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        for path in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                                    "libamdhip64.so")):
            try:
                _rt = ctypes.CDLL(path)
                break
            except OSError:
                continue
        if _rt is None:
            raise OSError("libamdhip64.so not found")
        _rt.hipEventCreate.argtypes = [ctypes.c_void_p]
        _rt.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _rt.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _rt.hipEventElapsedTime.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _rt.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return _rt


class Event:
    def __init__(self):
        self.ev = ctypes.c_void_p()
        rc = runtime().hipEventCreate(ctypes.byref(self.ev))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")

    def record(self, stream: int) -> None:
        rc = runtime().hipEventRecord(self.ev, ctypes.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def synchronize(self) -> None:
        runtime().hipEventSynchronize(self.ev)

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0.0)
        end.synchronize()
        rc = runtime().hipEventElapsedTime(ctypes.byref(ms), self.ev, end.ev)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return float(ms.value)

    def __del__(self):
        try:
            if self.ev:
                runtime().hipEventDestroy(self.ev)
        except Exception:
            pass
