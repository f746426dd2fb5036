"""ctypes binding of the C ABI in include/dpm.h (libdpm_hip.so).

This is the binding a Python caller uses for the batched forms (many
tissues in one handle) and for timing; the drop-in classes of the reference
(`clDPM.Tissue2D/3D`) reach the same library through the pybind11 module.
The library is the HIP implementation only: if it is missing or no GPU is
visible, calls fail loudly; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdpm_hip.so")

DPM_OK, DPM_EINVAL, DPM_ENONFINITE, DPM_EDEVICE, DPM_ENOMEM, DPM_ERANGE = range(6)

DPM2D_AREA, DPM2D_PERIMETER, DPM2D_BENDING, DPM2D_ATTRACTION, DPM2D_REPULSION = 1, 2, 4, 8, 16
DPM2D_ALL = 31
DPM2D_AREA_GRAD, DPM2D_VREPULSION = 32, 64
DPM2D_CONSERVATIVE = 102
DPM2D_ALL_TERMS = 127
DPM3D_VOLUME, DPM3D_SURFACE, DPM3D_SUBSTRATE, DPM3D_CONTACT = 1, 2, 4, 8
DPM3D_ADHESION, DPM3D_BENDING = 16, 32
DPM3D_REFERENCE, DPM3D_ALL = 15, 63

# Every symbol include/dpm.h declares (checked by the CPU tests).
EXPORTS = [
    "dpm_last_error", "dpm_version", "dpm_device_count",
    "dpm3d_create", "dpm3d_set_positions", "dpm3d_get_positions", "dpm3d_get_forces",
    "dpm3d_set_cell_params", "dpm3d_set_tissue_params", "dpm3d_cell_stats", "dpm3d_probe",
    "dpm3d_probe_adhesion", "dpm2d_probe_attraction", "dpm3d_probe_noalloc", "dpm3d_cell_measures",
    "dpm_disperse", "dpm2d_set_step_kernel", "dpm3d_probe_contact",
    "dpm3d_create_f64", "dpm3d_set_positions_f64", "dpm3d_get_positions_f64",
    "dpm3d_get_forces_f64", "dpm3d_energy", "dpm3d_trajectory", "dpm3d_adjoint", "dpm3d_adjoint_f64",
    "dpm3d_set_compensated", "dpm3d_set_cell_list", "dpm3d_set_bin_width", "dpm3d_bin_width",
    "dpm3d_set_step_kernel", "dpm3d_autotune",
    "dpm2d_create", "dpm2d_set_positions", "dpm2d_get_positions", "dpm2d_get_forces",
    "dpm2d_set_cell_params", "dpm2d_set_tissue_params", "dpm2d_probe",
    "dpm2d_create_f64", "dpm2d_set_positions_f64", "dpm2d_get_positions_f64", "dpm2d_get_forces_f64",
    "dpm2d_get_box", "dpm_fire_defaults", "dpm_jam_defaults", "dpm2d_fire", "dpm2d_energy", "dpm2d_jam",
    "dpm2d_set_relax_threads", "dpm2d_relax_threads", "dpm2d_autotune", "dpm2d_jam_step",
    "dpm_euler", "dpm_euler_async", "dpm_synchronize", "dpm_stream", "dpm_counters",
    "dpm_destroy",
]


class DPMError(RuntimeError):
    pass


_lib: Optional[ctypes.CDLL] = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m opencl_dpm_amd.build` (HIP path required)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i, sz, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_float
    lib.dpm_last_error.restype = ctypes.c_char_p
    lib.dpm_stream.restype = vp
    lib.dpm_stream.argtypes = [vp]
    for name in ["dpm3d_set_positions", "dpm3d_get_positions", "dpm3d_get_forces",
                 "dpm3d_set_positions_f64", "dpm3d_get_positions_f64", "dpm3d_get_forces_f64",
                 "dpm2d_set_positions", "dpm2d_get_positions", "dpm2d_get_forces",
                 "dpm2d_set_positions_f64", "dpm2d_get_positions_f64", "dpm2d_get_forces_f64"]:
        getattr(lib, name).argtypes = [vp, vp, sz]
    lib.dpm_euler.argtypes = [vp, i, fp]
    lib.dpm_euler_async.argtypes = [vp, i, fp]
    lib.dpm_synchronize.argtypes = [vp]
    lib.dpm_counters.argtypes = [vp, vp, i]
    lib.dpm_destroy.argtypes = [vp]
    lib.dpm_destroy.restype = None
    lib.dpm3d_create.argtypes = [vp, vp]
    lib.dpm3d_create_f64.argtypes = [vp, vp]
    lib.dpm3d_energy.argtypes = [vp, vp]
    lib.dpm3d_trajectory.argtypes = [vp, i, fp, i]
    lib.dpm3d_set_compensated.argtypes = [vp, i]
    lib.dpm3d_set_cell_list.argtypes = [vp, i]
    lib.dpm3d_set_bin_width.argtypes = [vp, fp]
    lib.dpm3d_bin_width.argtypes = [vp]
    lib.dpm3d_bin_width.restype = fp
    lib.dpm3d_set_step_kernel.argtypes = [vp, i]
    lib.dpm3d_autotune.argtypes = [vp, i, fp, vp, vp]
    lib.dpm3d_adjoint.argtypes = [vp, vp, vp, vp]
    lib.dpm3d_adjoint_f64.argtypes = [vp, vp, vp, vp]
    lib.dpm2d_create.argtypes = [vp, vp]
    lib.dpm2d_create_f64.argtypes = [vp, vp]
    lib.dpm2d_get_box.argtypes = [vp, vp]
    lib.dpm_fire_defaults.argtypes = [vp]
    lib.dpm_fire_defaults.restype = None
    lib.dpm_jam_defaults.argtypes = [vp]
    lib.dpm_jam_defaults.restype = None
    lib.dpm2d_fire.argtypes = [vp, vp, vp]
    lib.dpm2d_energy.argtypes = [vp, vp]
    lib.dpm2d_jam.argtypes = [vp, vp, vp]
    lib.dpm2d_jam_step.argtypes = [vp, vp, i, i, vp]
    lib.dpm2d_set_relax_threads.argtypes = [vp, i]
    lib.dpm2d_relax_threads.argtypes = [vp]
    lib.dpm2d_autotune.argtypes = [vp, i, vp, vp]
    lib.dpm3d_set_cell_params.argtypes = [vp] + [vp] * 6
    lib.dpm3d_set_tissue_params.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_uint]
    lib.dpm2d_set_cell_params.argtypes = [vp] + [vp] * 6
    lib.dpm2d_set_tissue_params.argtypes = [vp, vp, vp, vp, vp, ctypes.c_uint]
    lib.dpm3d_cell_stats.argtypes = [vp, vp, vp]
    lib.dpm3d_probe.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.dpm2d_probe.argtypes = [vp, vp, vp, vp]
    lib.dpm3d_probe_adhesion.argtypes = [vp, vp, vp]
    lib.dpm3d_probe_noalloc.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.dpm3d_cell_measures.argtypes = [vp, vp, vp, vp]
    lib.dpm_disperse.argtypes = [vp, i, fp, vp, vp, i, vp]
    lib.dpm2d_set_step_kernel.argtypes = [vp, i]
    lib.dpm3d_probe_contact.argtypes = [vp, vp, vp]
    lib.dpm2d_probe_attraction.argtypes = [vp, vp]
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != DPM_OK:
        msg = load().dpm_last_error().decode()
        if rc == DPM_EINVAL:
            raise ValueError(msg)
        raise DPMError(f"[{rc}] {msg}")


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    n = ctypes.c_int(0)
    load().dpm_device_count(ctypes.byref(n))
    return n.value


class _Desc3D(ctypes.Structure):
    _fields_ = [
        ("ntissues", ctypes.c_int), ("cell_begin", ctypes.c_void_p),
        ("nv", ctypes.c_int), ("nf", ctypes.c_int), ("faces", ctypes.c_void_p),
        ("Kv", ctypes.c_void_p), ("Ka", ctypes.c_void_p), ("Ks", ctypes.c_void_p),
        ("Kb", ctypes.c_void_p), ("v0", ctypes.c_void_p), ("a0", ctypes.c_void_p),
        ("L", ctypes.c_void_p), ("pbc", ctypes.c_void_p), ("Kre", ctypes.c_void_p),
        ("Kat", ctypes.c_void_p), ("ks_grad", ctypes.c_void_p), ("mask", ctypes.c_uint),
    ]


class _Desc2D(ctypes.Structure):
    _fields_ = [
        ("ntissues", ctypes.c_int), ("cell_begin", ctypes.c_void_p), ("nv", ctypes.c_void_p),
        ("Ka", ctypes.c_void_p), ("Kl", ctypes.c_void_p), ("Kb", ctypes.c_void_p),
        ("a0", ctypes.c_void_p), ("l0", ctypes.c_void_p), ("r0", ctypes.c_void_p),
        ("L", ctypes.c_void_p), ("pbc", ctypes.c_void_p), ("Kre", ctypes.c_void_p),
        ("Kat", ctypes.c_void_p), ("mask", ctypes.c_uint),
    ]


class FireParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dt0", "dt_max", "ftol", "f_inc", "f_dec", "alpha0", "f_alpha")] + \
               [("max_iter", ctypes.c_int), ("n_min", ctypes.c_int)]


class JamParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("phi0", "dphi", "phi_tol", "u_jam")] + \
               [("max_stages", ctypes.c_int), ("fire", FireParams)]


RELAX_STATS = ("iterations", "converged", "fmax", "energy", "e_int", "phi", "phi_lo", "phi_hi", "stages", "L",
               "unconverged", "status", "phi_target")


def fire_params(**kw) -> FireParams:
    p = FireParams()
    load().dpm_fire_defaults(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def jam_params(fire=None, **kw) -> JamParams:
    p = JamParams()
    load().dpm_jam_defaults(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    for k, v in (fire or {}).items():
        setattr(p.fire, k, v)
    return p


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class _Handle:
    def __init__(self):
        self._h = ctypes.c_void_p(None)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            load().dpm_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def euler(self, nsteps: int, dt: float) -> None:
        _check(load().dpm_euler(self._h, int(nsteps), float(dt)))

    def euler_async(self, nsteps: int, dt: float) -> None:
        _check(load().dpm_euler_async(self._h, int(nsteps), float(dt)))

    def synchronize(self) -> None:
        _check(load().dpm_synchronize(self._h))

    def stream(self) -> int:
        return load().dpm_stream(self._h) or 0

    def counters(self, reset: bool = False) -> np.ndarray:
        out = np.zeros(4, np.uint64)
        _check(load().dpm_counters(self._h, _p(out), int(reset)))
        return out


class Tissues3D(_Handle):
    """A batch of 3D tissues sharing one mesh (include/dpm.h, dpm3d_*).

    precision=32 is the production step (float positions and forces, double
    per-cell sums); precision=64 the all-double step (dpm3d_create_f64)."""

    def __init__(self, faces, cell_begin: Sequence[int], Kv, Ka, Ks, v0, a0, L, Kre, Kat=0.0,
                 pbc=1, Kb=None, ks_grad=None, mask: int = DPM3D_REFERENCE, nv: int = None,
                 precision: int = 32):
        super().__init__()
        if precision not in (32, 64):
            raise ValueError("precision must be 32 or 64")
        self.precision = precision
        self._faces = _i32(faces).reshape(-1, 3)
        self.nf = self._faces.shape[0]
        self.nv = int(nv if nv is not None else self._faces.max() + 1)
        self.cell_begin = _i32(cell_begin)
        self.T = len(self.cell_begin) - 1
        self.C = int(self.cell_begin[-1])
        bc = lambda v, n: _f32(np.broadcast_to(np.asarray(v, np.float32), (n,)))
        self._cell = [bc(Kv, self.C), bc(Ka, self.C), bc(Ks, self.C),
                      bc(0.0 if Kb is None else Kb, self.C), bc(v0, self.C), bc(a0, self.C)]
        self._tis = [bc(L, self.T), _i32(np.broadcast_to(np.asarray(pbc), (self.T,))),
                     bc(Kre, self.T), bc(Kat, self.T),
                     bc(0.0 if ks_grad is None else ks_grad, self.T)]
        d = _Desc3D()
        d.ntissues = self.T
        d.cell_begin = _p(self.cell_begin)
        d.nv, d.nf, d.faces = self.nv, self.nf, _p(self._faces)
        d.Kv, d.Ka, d.Ks, d.Kb, d.v0, d.a0 = [_p(a) for a in self._cell]
        d.L, d.pbc, d.Kre, d.Kat, d.ks_grad = [_p(a) for a in self._tis]
        d.mask = mask
        create = load().dpm3d_create_f64 if precision == 64 else load().dpm3d_create
        _check(create(ctypes.byref(d), ctypes.byref(self._h)))

    @property
    def nverts(self) -> int:
        return self.C * self.nv

    def set_positions(self, xyz) -> None:
        """float32 input goes through dpm3d_set_positions, anything else as
        float64 through dpm3d_set_positions_f64 (exact on a float64 handle)."""
        if np.asarray(xyz).dtype == np.float32:
            a = _f32(xyz).reshape(-1, 3)
            _check(load().dpm3d_set_positions(self._h, _p(a), a.shape[0]))
        else:
            a = np.ascontiguousarray(np.asarray(xyz, np.float64)).reshape(-1, 3)
            _check(load().dpm3d_set_positions_f64(self._h, _p(a), a.shape[0]))

    def _get(self, f32, f64, dtype):
        dtype = np.dtype(dtype or (np.float64 if self.precision == 64 else np.float32))
        out = np.empty((self.nverts, 3), dtype)
        fn = f64 if dtype == np.float64 else f32
        _check(fn(self._h, _p(out), self.nverts))
        return out

    def positions(self, dtype=None) -> np.ndarray:
        """Positions in the handle's precision (or `dtype`: float32/float64)."""
        return self._get(load().dpm3d_get_positions, load().dpm3d_get_positions_f64, dtype)

    def forces(self, dtype=None) -> np.ndarray:
        return self._get(load().dpm3d_get_forces, load().dpm3d_get_forces_f64, dtype)

    def set_cell_list(self, mode: int = -1) -> None:
        """Candidate search: 0 scan, 1 binned cell list, -1 automatic."""
        _check(load().dpm3d_set_cell_list(self._h, int(mode)))

    def set_bin_width(self, factor: float = 1.0) -> None:
        """Cell-list bin width over the largest candidate reach, in [1, 4]
        (dpm3d_set_bin_width)."""
        _check(load().dpm3d_set_bin_width(self._h, ctypes.c_float(factor)))

    def bin_width(self) -> float:
        return float(load().dpm3d_bin_width(self._h))

    def set_step_kernel(self, kind: int = 0) -> None:
        """0 automatic, 1 one workgroup per cell, 2 one wave per cell
        (dpm3d_set_step_kernel)."""
        _check(load().dpm3d_set_step_kernel(self._h, int(kind)))

    AUTOTUNE_CONFIGS = ("workgroup/scan", "workgroup/cell list", "wave x1/scan", "wave x2/scan", "wave x4/scan",
                        "wave x1/cell list", "wave x2/cell list", "wave x4/cell list")

    def autotune(self, steps: int = 10, dt: float = 0.002):
        """Times the step configurations on the current state and keeps the
        fastest (dpm3d_autotune).  Returns ({form, cells per workgroup, cell
        list}, {configuration: ms per step})."""
        best = np.zeros(3, np.int32)
        t = np.zeros(8, np.float32)
        _check(load().dpm3d_autotune(self._h, int(steps), ctypes.c_float(dt), _p(best), _p(t)))
        return ({"form": int(best[0]), "cells_per_workgroup": int(best[1]), "cell_list": int(best[2]),
                 "bin_width": self.bin_width()},
                {k: float(v) for k, v in zip(self.AUTOTUNE_CONFIGS, t) if v >= 0})

    def set_compensated(self, on: bool = True) -> None:
        """Compensated (x + residual) positions for a float tissue."""
        _check(load().dpm3d_set_compensated(self._h, int(bool(on))))

    def trajectory(self, nsteps: int, dt: float, every: int = 20) -> None:
        """nsteps Euler steps keeping checkpoints every `every` steps."""
        _check(load().dpm3d_trajectory(self._h, int(nsteps), float(dt), int(every)))

    def adjoint(self, lambda_final):
        """dL/d(Kv, Ka, Kb, Ks) per cell [C, 4] and dL/dx_0 [nverts, 3] of the
        recorded trajectory, given dL/dx_N."""
        dth = np.zeros((self.C, 4), np.float64)
        if self.precision == 64:
            lf = np.ascontiguousarray(np.asarray(lambda_final, np.float64).reshape(-1, 3))
            l0 = np.empty_like(lf)
            _check(load().dpm3d_adjoint_f64(self._h, _p(lf), _p(dth), _p(l0)))
        else:
            lf = _f32(lambda_final).reshape(-1, 3)
            l0 = np.empty_like(lf)
            _check(load().dpm3d_adjoint(self._h, _p(lf), _p(dth), _p(l0)))
        return dth, l0

    def energy(self) -> np.ndarray:
        """Per-cell elastic energy (volume + edge springs + bending, per mask)."""
        out = np.empty(self.C, np.float64)
        _check(load().dpm3d_energy(self._h, _p(out)))
        return out

    def set_tissue_params(self, L=None, pbc=None, Kre=None, Kat=None, ks_grad=None, mask=None):
        bc = lambda v: None if v is None else _f32(np.broadcast_to(np.asarray(v, np.float32), (self.T,)))
        arrs = [bc(L), None if pbc is None else _i32(np.broadcast_to(np.asarray(pbc), (self.T,))),
                bc(Kre), bc(Kat), bc(ks_grad)]
        self._keep = arrs
        _check(load().dpm3d_set_tissue_params(self._h, *[_p(a) for a in arrs],
                                              ctypes.c_uint(DPM3D_REFERENCE if mask is None else mask)))

    def set_cell_params(self, Kv=None, Ka=None, Ks=None, Kb=None, v0=None, a0=None):
        bc = lambda v: None if v is None else _f32(np.broadcast_to(np.asarray(v, np.float32), (self.C,)))
        arrs = [bc(Kv), bc(Ka), bc(Ks), bc(Kb), bc(v0), bc(a0)]
        self._keep2 = arrs
        _check(load().dpm3d_set_cell_params(self._h, *[_p(a) for a in arrs]))

    def cell_stats(self):
        vol = np.empty(self.C, np.float64)
        com = np.empty((self.C, 3), np.float32)
        _check(load().dpm3d_cell_stats(self._h, _p(vol), _p(com)))
        return vol, com

    def cell_measures(self):
        """Per cell (volume, surface area, centre of mass) by the reference's
        host reductions evaluated on the device (dpm3d_cell_measures)."""
        vol = np.empty(self.C, np.float32)
        sa = np.empty(self.C, np.float32)
        com = np.empty((self.C, 3), np.float32)
        _check(load().dpm3d_cell_measures(self._h, _p(vol), _p(sa), _p(com)))
        return vol, sa, com

    def eval_forces(self, dtype=None) -> np.ndarray:
        """Forces at the current positions, positions unchanged (dpm3d_probe
        with no pair tables), in the handle's precision or `dtype`."""
        _check(load().dpm3d_probe(self._h, None, None, None, None, None))
        return self.forces(dtype)

    def probe(self, winding: bool = False, contact: bool = True):
        """Forces at the current positions (handle precision) plus the
        recorded pair sets; contact=False skips the vertex x cell table."""
        ncmax = int(np.max(np.diff(self.cell_begin)))
        ct = np.zeros((self.nverts, ncmax), np.uint8) if contact else None
        w = np.zeros((self.nverts, ncmax), np.float32) if winding else None
        down = np.zeros((self.C, self.nf), np.uint8)
        adh = np.zeros(self.nverts, np.int32)
        _check(load().dpm3d_probe(self._h, None, _p(ct), _p(w), _p(down), _p(adh)))
        out = {"forces": self.forces(), "contact": ct, "winding": w, "down": down, "adhesion": adh}
        own = np.zeros(self.nverts, np.uint32)
        rev = np.zeros(self.nverts, np.uint32)
        _check(load().dpm3d_probe_adhesion(self._h, _p(own), _p(rev)))
        out["adhesion_hash"] = np.stack([own, rev], 1)
        return out


    def probe_contact(self):
        """Per vertex the number of contact partner cells and the key sum of
        their indices (dpm3d_probe_contact); forces as probe()."""
        n = np.zeros(self.nverts, np.uint32)
        k = np.zeros(self.nverts, np.uint32)
        _check(load().dpm3d_probe_contact(self._h, _p(n), _p(k)))
        return n, k

    def probe_noalloc(self, winding: bool = False, contact: bool = True):
        """probe() through dpm3d_probe_noalloc (device tables in the library's
        static scratch, no device allocation): forces as float, pair tables."""
        ncmax = int(np.max(np.diff(self.cell_begin)))
        f = np.zeros((self.nverts, 3), np.float32)
        ct = np.zeros((self.nverts, ncmax), np.uint8) if contact else None
        w = np.zeros((self.nverts, ncmax), np.float32) if winding else None
        down = np.zeros((self.C, self.nf), np.uint8)
        adh = np.zeros(self.nverts, np.int32)
        _check(load().dpm3d_probe_noalloc(self._h, _p(f), _p(ct), _p(w), _p(down), _p(adh)))
        return {"forces": f, "contact": ct, "winding": w, "down": down, "adhesion": adh}


class Tissues2D(_Handle):
    """A batch of 2D tissues (include/dpm.h, dpm2d_*); vertices packed.
    precision=64 makes an all-double tissue (dpm2d_create_f64)."""

    def __init__(self, cell_begin, nv, Ka, Kl, Kb, a0, l0, r0, L, Kre=1.0, Kat=0.0, pbc=1,
                 mask: int = DPM2D_ALL, precision: int = 32):
        super().__init__()
        if precision not in (32, 64):
            raise ValueError("precision must be 32 or 64")
        self.precision = precision
        self.cell_begin = _i32(cell_begin)
        self.T = len(self.cell_begin) - 1
        self.nv = _i32(nv)
        self.C = int(self.cell_begin[-1])
        self.V = int(self.nv.sum())
        bc = lambda v, n: _f32(np.broadcast_to(np.asarray(v, np.float32), (n,)))
        self._cell = [bc(Ka, self.C), bc(Kl, self.C), bc(Kb, self.C), bc(a0, self.C),
                      bc(l0, self.C), bc(r0, self.C)]
        self._tis = [bc(L, self.T), _i32(np.broadcast_to(np.asarray(pbc), (self.T,))),
                     bc(Kre, self.T), bc(Kat, self.T)]
        d = _Desc2D()
        d.ntissues = self.T
        d.cell_begin = _p(self.cell_begin)
        d.nv = _p(self.nv)
        d.Ka, d.Kl, d.Kb, d.a0, d.l0, d.r0 = [_p(a) for a in self._cell]
        d.L, d.pbc, d.Kre, d.Kat = [_p(a) for a in self._tis]
        d.mask = mask
        create = load().dpm2d_create_f64 if precision == 64 else load().dpm2d_create
        _check(create(ctypes.byref(d), ctypes.byref(self._h)))

    def set_positions(self, xy) -> None:
        if np.asarray(xy).dtype == np.float32:
            a = _f32(xy).reshape(-1, 2)
            _check(load().dpm2d_set_positions(self._h, _p(a), a.shape[0]))
        else:
            a = np.ascontiguousarray(np.asarray(xy, np.float64)).reshape(-1, 2)
            _check(load().dpm2d_set_positions_f64(self._h, _p(a), a.shape[0]))

    def _get(self, f32, f64, dtype):
        dtype = np.dtype(dtype or (np.float64 if self.precision == 64 else np.float32))
        out = np.empty((self.V, 2), dtype)
        _check((f64 if dtype == np.float64 else f32)(self._h, _p(out), self.V))
        return out

    def positions(self, dtype=None) -> np.ndarray:
        return self._get(load().dpm2d_get_positions, load().dpm2d_get_positions_f64, dtype)

    def forces(self, dtype=None) -> np.ndarray:
        return self._get(load().dpm2d_get_forces, load().dpm2d_get_forces_f64, dtype)

    def box(self) -> np.ndarray:
        out = np.empty(self.T, np.float64)
        _check(load().dpm2d_get_box(self._h, _p(out)))
        return out

    def set_tissue_params(self, L=None, pbc=None, Kre=None, Kat=None, mask=DPM2D_ALL):
        bc = lambda v: None if v is None else _f32(np.broadcast_to(np.asarray(v, np.float32), (self.T,)))
        arrs = [bc(L), None if pbc is None else _i32(np.broadcast_to(np.asarray(pbc), (self.T,))),
                bc(Kre), bc(Kat)]
        self._keep = arrs
        _check(load().dpm2d_set_tissue_params(self._h, *[_p(a) for a in arrs], ctypes.c_uint(mask)))

    @staticmethod
    def _stats(buf):
        a = buf.reshape(-1, len(RELAX_STATS))
        return {k: a[:, i].copy() for i, k in enumerate(RELAX_STATS)}

    def fire(self, **kw):
        """FIRE minimisation of every tissue (dpm2d_fire); per-tissue stats."""
        p = fire_params(**kw)
        buf = np.zeros(self.T * len(RELAX_STATS))
        _check(load().dpm2d_fire(self._h, ctypes.byref(p), _p(buf)))
        return self._stats(buf)

    def energy(self):
        buf = np.zeros(self.T * len(RELAX_STATS))
        _check(load().dpm2d_energy(self._h, _p(buf)))
        return self._stats(buf)

    def jam(self, fire=None, **kw):
        """Jamming protocol of every tissue in one (graph-replayed) launch."""
        p = jam_params(fire=fire, **kw)
        buf = np.zeros(self.T * len(RELAX_STATS))
        _check(load().dpm2d_jam(self._h, ctypes.byref(p), _p(buf)))
        return self._stats(buf)

    def jam_step(self, start: bool, stages: int, fire=None, **kw):
        """At most `stages` more minimisations of a jamming protocol
        (dpm2d_jam_step); status 2 while unfinished."""
        p = jam_params(fire=fire, **kw)
        buf = np.zeros(self.T * len(RELAX_STATS))
        _check(load().dpm2d_jam_step(self._h, ctypes.byref(p), int(bool(start)), int(stages), _p(buf)))
        return self._stats(buf)

    def set_relax_threads(self, n: int) -> None:
        _check(load().dpm2d_set_relax_threads(self._h, int(n)))

    def relax_threads(self) -> int:
        return int(load().dpm2d_relax_threads(self._h))

    def set_step_kernel(self, kind: int) -> None:
        """0 automatic, 1 one workgroup per tissue, 2 the wide form."""
        _check(load().dpm2d_set_step_kernel(self._h, int(kind)))

    def autotune(self, iters: int = 50):
        best = ctypes.c_int(0)
        times = np.zeros(5, np.float32)
        _check(load().dpm2d_autotune(self._h, int(iters), ctypes.byref(best), _p(times)))
        return best.value, dict(zip((64, 128, 256, 512, 1024), times.tolist()))

    def probe(self):
        f = np.empty((self.V, 2), np.float32)
        ov = np.zeros(self.V, np.int32)
        at = np.zeros(self.V, np.int32)
        _check(load().dpm2d_probe(self._h, _p(f), _p(ov), _p(at)))
        keys = np.zeros(self.V, np.uint32)
        _check(load().dpm2d_probe_attraction(self._h, _p(keys)))
        return {"forces": f, "overlap": ov, "attraction": at, "attraction_hash": keys}


def disperse(r, L, X, Y, max_count: int = 100000):
    """dpm_disperse: the reference's dispersal relaxation on the device.
    Returns the relaxed centres and (iterations, capped)."""
    r = _f32(r)
    n = len(r)
    X = np.array(X, np.float32)
    Y = np.array(Y, np.float32)
    res = np.zeros(2, np.int32)
    _check(load().dpm_disperse(_p(r), int(n), float(L), _p(X), _p(Y), int(max_count), _p(res)))
    return X, Y, (int(res[0]), bool(res[1]))
