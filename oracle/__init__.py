"""TEST INFRASTRUCTURE ONLY — the oracle.

ctypes wrapper of the CPU restatement of the reference kernels
(dpm_oracle.c: ``*_f32`` keeps the reference's float arithmetic, ``*_f64``
the same formulas in double) and of the reference's own host code built
from its sources into ``_ref/`` (set-up geometry and dispersal).  Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def _lib(name: str) -> ctypes.CDLL:
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise ImportError(f"{path} missing; run `make -C oracle`")
        _libs[name] = ctypes.CDLL(path)
    return _libs[name]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _t(prec):
    return (np.float32, ctypes.c_float, "_f32") if prec == 32 else (np.float64, ctypes.c_double, "_f64")


def forces2d(verts, stride, nv, Ka, Kl, Kb, a0, l0, r0, L, Kre, Kat, pbc=1, mask=31, prec=32,
             lib="liboracle.so", attr_hash=False):
    """One 2D force evaluation.  verts: padded [ncells*stride, 2].  Returns forces,
    overlap flags, attraction partner counts and, with attr_hash, each vertex's
    attraction partner-set key sum (vertex_keys)."""
    dt_, ct, sfx = _t(prec)
    nc = len(nv)
    V = np.ascontiguousarray(np.asarray(verts, dt_).reshape(nc * stride, 2))
    F = np.zeros_like(V)
    ov = np.zeros(nc * stride, np.int32)
    at = np.zeros(nc * stride, np.int32)
    ah = np.zeros(nc * stride, np.uint32) if attr_hash else None
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Ka, Kl, Kb, a0, l0, r0)]
    nvv = np.ascontiguousarray(np.asarray(nv, np.int32))
    f = getattr(_lib(lib), "ora2d_forces" + sfx)
    f(ctypes.c_int(nc), ctypes.c_int(stride), _p(nvv), _p(V), *[_p(a) for a in ps], ct(L),
      ctypes.c_int(pbc), ct(Kre), ct(Kat), ctypes.c_uint(mask), _p(F), _p(ov), _p(at), _p(ah))
    if attr_hash:
        return F, ov, at, ah
    return F, ov, at


def run2d(verts, stride, nv, Ka, Kl, Kb, a0, l0, r0, L, Kre, Kat, nsteps, dt, pbc=1, mask=31,
          prec=32, lib="liboracle.so"):
    dt_, ct, sfx = _t(prec)
    nc = len(nv)
    V = np.ascontiguousarray(np.asarray(verts, dt_).reshape(nc * stride, 2)).copy()
    F = np.zeros_like(V)
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Ka, Kl, Kb, a0, l0, r0)]
    nvv = np.ascontiguousarray(np.asarray(nv, np.int32))
    f = getattr(_lib(lib), "ora2d_run" + sfx)
    f(ctypes.c_int(nc), ctypes.c_int(stride), _p(nvv), _p(V), *[_p(a) for a in ps], ct(L),
      ctypes.c_int(pbc), ct(Kre), ct(Kat), ctypes.c_uint(mask), ctypes.c_int(nsteps), ct(dt), _p(F))
    return V, F


def energy2d(verts, stride, nv, Ka, Kl, Kb, a0, l0, L, Kre, pbc=1, mask=102, prec=64):
    """(total, vertex-repulsion part) of the conservative 2D energy."""
    dt_, ct, sfx = _t(prec)
    nc = len(nv)
    V = np.ascontiguousarray(np.asarray(verts, dt_).reshape(nc * stride, 2))
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Ka, Kl, Kb, a0, l0)]
    nvv = np.ascontiguousarray(np.asarray(nv, np.int32))
    out = np.zeros(2)
    getattr(_lib("liboracle.so"), "ora2d_energy" + sfx)(
        ctypes.c_int(nc), ctypes.c_int(stride), _p(nvv), _p(V), *[_p(a) for a in ps], ct(L), ctypes.c_int(pbc),
        ct(Kre), ctypes.c_uint(mask), _p(out))
    return float(out[0]), float(out[1])


FIRE_DEFAULTS = dict(dt0=0.01, dt_max=0.1, ftol=1e-10, f_inc=1.1, f_dec=0.5, alpha0=0.1, f_alpha=0.99,
                     max_iter=100000, n_min=5)


def fire2d(verts, stride, nv, Ka, Kl, Kb, a0, l0, r0, L, Kre, Kat, vel=None, pbc=1, mask=31, prec=64,
           u_clear=-1.0, **fire):
    """FIRE minimisation (ora2d_fire).  Returns positions, velocities and
    {iterations, converged, fmax, dt, alpha, n_pos, cleared}."""
    dt_, ct, sfx = _t(prec)
    p = dict(FIRE_DEFAULTS, **fire)
    nc = len(nv)
    V = np.ascontiguousarray(np.asarray(verts, dt_).reshape(nc * stride, 2)).copy()
    Vel = np.zeros_like(V) if vel is None else np.ascontiguousarray(np.asarray(vel, dt_).reshape(V.shape)).copy()
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Ka, Kl, Kb, a0, l0, r0)]
    nvv = np.ascontiguousarray(np.asarray(nv, np.int32))
    fp = np.array([p["dt0"], p["dt_max"], p["ftol"], p["f_inc"], p["f_dec"], p["alpha0"], p["f_alpha"],
                   u_clear])
    ip = np.array([p["max_iter"], p["n_min"]], np.int32)
    out = np.zeros(7)
    f = getattr(_lib("liboracle.so"), "ora2d_fire" + sfx)
    rc = f(ctypes.c_int(nc), ctypes.c_int(stride), _p(nvv), _p(V), _p(Vel), *[_p(a) for a in ps], ct(L),
           ctypes.c_int(pbc), ct(Kre), ct(Kat), ctypes.c_uint(mask), _p(fp), _p(ip), _p(out))
    assert rc == 0
    st = dict(zip(["iterations", "converged", "fmax", "dt", "alpha", "n_pos", "cleared"], out.tolist()))
    return V, Vel, st


JAM_DEFAULTS = dict(phi0=0.35, dphi=0.02, phi_tol=1e-3, u_jam=1e-16, max_stages=40)


def jam2d(verts, stride, nv, Ka, Kl, Kb, a0, l0, r0, L, Kre, Kat, pbc=1, mask=102, prec=64, fire=None, **jam):
    """Jamming protocol (ora2d_jam).  Returns positions, L and
    {phi_lo, phi_hi, stages, iterations, L, status, unconverged}."""
    dt_, ct, sfx = _t(prec)
    p = dict(FIRE_DEFAULTS, max_iter=20000)
    p.update(fire or {})
    j = dict(JAM_DEFAULTS, **jam)
    nc = len(nv)
    V = np.ascontiguousarray(np.asarray(verts, dt_).reshape(nc * stride, 2)).copy()
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Ka, Kl, Kb, a0, l0, r0)]
    nvv = np.ascontiguousarray(np.asarray(nv, np.int32))
    fp = np.array([p["dt0"], p["dt_max"], p["ftol"], p["f_inc"], p["f_dec"], p["alpha0"], p["f_alpha"]])
    ip = np.array([p["max_iter"], p["n_min"], j["max_stages"]], np.int32)
    jp = np.array([j["phi0"], j["dphi"], j["phi_tol"], j["u_jam"]])
    Lb = np.array([L], dt_)
    out = np.zeros(7)
    f = getattr(_lib("liboracle.so"), "ora2d_jam" + sfx)
    rc = f(ctypes.c_int(nc), ctypes.c_int(stride), _p(nvv), _p(V), *[_p(a) for a in ps], _p(Lb),
           ctypes.c_int(pbc), ct(Kre), ct(Kat), ctypes.c_uint(mask), _p(fp), _p(ip), _p(jp), _p(out))
    assert rc == 0
    st = dict(zip(["phi_lo", "phi_hi", "stages", "iterations", "L", "status", "unconverged"], out.tolist()))
    return V, float(Lb[0]), st


def adjacency(faces, nv, maxdeg=8):
    faces = np.asarray(faces, np.int64).reshape(-1, 3)
    nb = [set() for _ in range(nv)]
    for a, b, c in faces:
        nb[a] |= {b, c}
        nb[b] |= {a, c}
        nb[c] |= {a, b}
    adj = -np.ones((nv, maxdeg), np.int32)
    for v in range(nv):
        s = sorted(nb[v])
        adj[v, :len(s)] = s
    return adj


def forces3d(X, faces, nv, Kv, Ka, Ks, v0, a0, L, Kre, Kat=0.0, Kb=0.0, pbc=1, mask=15,
             ks_grad=0.0, prec=32, pairs=False, winding=False, lib="liboracle.so"):
    """One 3D force evaluation.  X: [ncells*nv, 3]."""
    dt_, ct, sfx = _t(prec)
    X = np.ascontiguousarray(np.asarray(X, dt_).reshape(-1, 3))
    nc = X.shape[0] // nv
    faces = np.ascontiguousarray(np.asarray(faces, np.int32).reshape(-1, 3))
    nf = faces.shape[0]
    adj = adjacency(faces, nv)
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Kv, Ka, Ks, Kb, v0, a0)]
    F = np.zeros_like(X)
    pr = np.zeros((nc * nv, nc), np.uint8) if pairs else None
    wn = np.zeros((nc * nv, nc), dt_) if winding else None
    down = np.zeros((nc, nf), np.uint8)
    adh = np.zeros(nc * nv, np.int32)
    ah = np.zeros((nc * nv, 2), np.uint32) if pairs else None
    f = getattr(_lib(lib), "ora3d_forces" + sfx)
    rc = f(ctypes.c_int(nc), ctypes.c_int(nv), ctypes.c_int(nf), _p(faces), _p(adj), ctypes.c_int(8),
           _p(X), *[_p(a) for a in ps], ct(L), ctypes.c_int(pbc), ct(Kre), ct(Kat), ctypes.c_uint(mask),
           ct(ks_grad), _p(F), _p(pr), _p(wn), _p(down), _p(adh), _p(ah))
    assert rc == 0
    return {"forces": F, "contact": pr, "winding": wn, "down": down, "adhesion": adh,
            "adhesion_hash": ah}


def run3d(X, faces, nv, Kv, Ka, Ks, v0, a0, L, Kre, nsteps, dt, Kat=0.0, Kb=0.0, pbc=1, mask=15,
          ks_grad=0.0, prec=32, lib="liboracle.so"):
    dt_, ct, sfx = _t(prec)
    X = np.ascontiguousarray(np.asarray(X, dt_).reshape(-1, 3)).copy()
    nc = X.shape[0] // nv
    faces = np.ascontiguousarray(np.asarray(faces, np.int32).reshape(-1, 3))
    adj = adjacency(faces, nv)
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Kv, Ka, Ks, Kb, v0, a0)]
    F = np.zeros_like(X)
    f = getattr(_lib(lib), "ora3d_run" + sfx)
    rc = f(ctypes.c_int(nc), ctypes.c_int(nv), ctypes.c_int(faces.shape[0]), _p(faces), _p(adj),
           ctypes.c_int(8), _p(X), *[_p(a) for a in ps], ct(L), ctypes.c_int(pbc), ct(Kre), ct(Kat),
           ctypes.c_uint(mask), ct(ks_grad), ctypes.c_int(nsteps), ct(dt), _p(F))
    assert rc == 0
    return X, F


def winding(X, faces, nv, cj, p, prec=64):
    dt_, ct, sfx = _t(prec)
    X = np.ascontiguousarray(np.asarray(X, dt_).reshape(-1, 3))
    faces = np.ascontiguousarray(np.asarray(faces, np.int32).reshape(-1, 3))
    pp = np.ascontiguousarray(np.asarray(p, dt_))
    f = getattr(_lib("liboracle.so"), "ora3d_winding" + sfx)
    f.restype = ct
    return f(ctypes.c_int(nv), ctypes.c_int(faces.shape[0]), _p(faces), _p(X), ctypes.c_int(cj), _p(pp))


def volumes(X, faces, nv, prec=64):
    dt_, ct, sfx = _t(prec)
    X = np.ascontiguousarray(np.asarray(X, dt_).reshape(-1, 3))
    nc = X.shape[0] // nv
    faces = np.ascontiguousarray(np.asarray(faces, np.int32).reshape(-1, 3))
    out = np.zeros(nc, dt_)
    getattr(_lib("liboracle.so"), "ora3d_volumes" + sfx)(
        ctypes.c_int(nc), ctypes.c_int(nv), ctypes.c_int(faces.shape[0]), _p(faces), _p(X), _p(out))
    return out


def energy3d(X, faces, nv, Kv, Ka, v0, a0, Kb=0.0, mask=35, prec=64):
    """Per-cell elastic energy (volume, edge springs, bending; mask bits 1, 2, 32)."""
    dt_, ct, sfx = _t(prec)
    X = np.ascontiguousarray(np.asarray(X, dt_).reshape(-1, 3))
    nc = X.shape[0] // nv
    faces = np.ascontiguousarray(np.asarray(faces, np.int32).reshape(-1, 3))
    adj = adjacency(faces, nv)
    arr = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt_), (nc,)))
    ps = [arr(a) for a in (Kv, Ka, Kb, v0, a0)]
    out = np.zeros(nc, dt_)
    getattr(_lib("liboracle.so"), "ora3d_energy" + sfx)(
        ctypes.c_int(nc), ctypes.c_int(nv), ctypes.c_int(faces.shape[0]), _p(faces), _p(adj),
        ctypes.c_int(8), _p(X), *[_p(a) for a in ps], ctypes.c_uint(mask), _p(out))
    return out


# ------------------------------------------------------ the reference's own host code

REF_LIB = os.path.join(_HERE, "_ref", "libdpmref.so")


def ref_available() -> bool:
    return os.path.exists(REF_LIB)
