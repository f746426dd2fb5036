"""Seeded synthetic inputs for the five benchmark configurations.

Every generator draws from numpy's PCG64(seed) and builds cells with the
drop-in constructors (clDPM.Cell2D / Cell3D, whose geometry is bit-identical
to the reference's cell.cpp), so the same arrays can be handed to the HIP
library, the pybind11 drop-in and the CPU oracle.

  monolayer3d   3D monolayer on a jittered square grid on the substrate
                plane z = 0 (256 cells x 162 vertices; 4096 x 642 for the
                precision form; 64 x 162 with a stiffness gradient for the
                differentiation form)
  tissue2d      2D bidisperse polygons relaxed to no overlap in a periodic box
  packings2d    many independent 2D packings of uneven size (batching form)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import clDPM


@dataclass
class Mesh:
    nv: int
    nf: int
    faces: np.ndarray        # [nf, 3] int32
    unit: np.ndarray         # [nv, 3] float32, the unit icosphere (r0 = 1 at origin)


_MESH_CACHE = {}


def icosphere(subdiv: int = 2) -> Mesh:
    if subdiv not in _MESH_CACHE:
        c = clDPM.Cell3D([0.0, 0.0, 0.0], 1.0, 1.0, subdiv)
        _MESH_CACHE[subdiv] = Mesh(c.NV, c.NF, np.asarray(c.GetFaces(), np.int32),
                                   np.asarray(c.Verts, np.float32))
    return _MESH_CACHE[subdiv]


@dataclass
class Monolayer:
    mesh: Mesh
    X: np.ndarray            # [C*nv, 3] float32 positions
    v0: np.ndarray           # [C]
    a0: np.ndarray           # [C]
    r0: np.ndarray           # [C]
    L: float
    Kv: float = 1.0
    Ka: float = 0.5
    Ks: float = 0.2
    Kb: float = 0.01
    Kre: float = 1.0
    Kat: float = 0.1
    ks_grad: float = 0.0
    dt: float = 0.002
    cells: List = field(default_factory=list)

    @property
    def C(self) -> int:
        return len(self.v0)


def monolayer3d(n_side: int = 16, spacing: float = 2.4, r_lo: float = 0.9, r_hi: float = 1.1,
                jitter: float = 0.1, subdiv: int = 2, seed: int = 0, keep_cells: bool = False,
                **params) -> Monolayer:
    """Cells of radius U(r_lo, r_hi) centred on an n_side x n_side grid with
    spacing `spacing` and uniform x/y jitter, resting on z = 0 (centre z = r0).
    The periodic box is L = n_side * spacing."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = n_side * n_side
    r0 = rng.uniform(r_lo, r_hi, C).astype(np.float32)
    jx = rng.uniform(-jitter, jitter, C)
    jy = rng.uniform(-jitter, jitter, C)
    mesh = icosphere(subdiv)
    j, i = np.arange(C) % n_side, np.arange(C) // n_side
    centres = np.stack([(j + 0.5) * spacing + jx, (i + 0.5) * spacing + jy, r0.astype(np.float64)],
                       1).astype(np.float32)
    # Cell3D's construction (cell.cpp:144-155) vectorised: unit icosphere
    # scaled by r0 then translated, in float; the sphere targets in double.
    X = (mesh.unit[None, :, :] * r0[:, None, None]) + centres[:, None, :]
    v0, a0 = sphere_targets(r0, 1.0, mesh.nf)
    cells = []
    if keep_cells:
        cells = [clDPM.Cell3D([float(x) for x in centres[c]], 1.0, float(r0[c]), subdiv) for c in range(C)]
    return Monolayer(mesh=mesh, X=X.astype(np.float32).reshape(-1, 3), v0=v0, a0=a0, r0=r0,
                     L=float(np.float32(n_side * spacing)), cells=cells, **params)


def sphere_targets(r0: np.ndarray, calA: float, nf: int):
    """v0 = 4/3 pi r0^3 and a0 = (6 sqrt(pi) v0 calA)^(2/3) / nf with the
    reference's float/double mix (cell.cpp:153-155)."""
    c43 = float(np.float32(4.0) / np.float32(3.0))
    e23 = float(np.float32(2.0) / np.float32(3.0))
    v0 = np.array([np.float32(c43 * math.pi * math.pow(float(r), 3)) for r in r0], np.float32)
    sa0 = np.array([np.float32(math.pow(6 * math.sqrt(math.pi) * float(v) * float(np.float32(calA)), e23))
                    for v in v0], np.float32)
    return v0, (sa0 / np.float32(nf)).astype(np.float32)


def monolayer_config(name: str, seed: int = 0) -> Monolayer:
    if name == "cfg2":   # plain 3D monolayer, 256 x 162
        return monolayer3d(16, 2.4, 0.9, 1.1, 0.1, 2, seed)
    if name == "cfg3":   # precision form, 4096 x 642
        return monolayer3d(64, 2.4, 0.85, 1.15, 0.05, 3, seed)
    if name == "cfg5":   # differentiation form, 64 x 162, Ks gradient 10x across x
        return monolayer3d(8, 2.4, 0.9, 1.1, 0.1, 2, seed, ks_grad=9.0)
    raise ValueError(name)


@dataclass
class Packing2D:
    nv: np.ndarray           # [C] int32
    X: np.ndarray            # [V, 2] float32, packed cell after cell
    a0: np.ndarray
    l0: np.ndarray
    r0: np.ndarray
    Ka: np.ndarray
    Kl: np.ndarray
    Kb: np.ndarray
    L: float
    Kre: float = 1.0
    Kat: float = 0.0

    @property
    def C(self) -> int:
        return len(self.nv)

    def padded(self):
        """Reference layout: [C*maxNV, 2] zero-padded, and the stride."""
        stride = int(self.nv.max())
        P = np.zeros((self.C * stride, 2), self.X.dtype)
        off = np.concatenate([[0], np.cumsum(self.nv)])
        for c in range(self.C):
            P[c * stride:c * stride + self.nv[c]] = self.X[off[c]:off[c + 1]]
        return P, stride


def _relax_disks(r: np.ndarray, L: float, rng, iters: int = 20000) -> np.ndarray:
    """Random centres in [0,L)^2 pushed apart (soft disks, minimum image)
    until no pair is closer than r_i + r_j."""
    n = len(r)
    P = rng.uniform(0.0, L, (n, 2))
    rsum = r[:, None] + r[None, :]
    for _ in range(iters):
        d = P[None, :, :] - P[:, None, :]
        d -= L * np.round(d / L)
        dist = np.sqrt((d ** 2).sum(-1))
        np.fill_diagonal(dist, np.inf)
        ov = np.clip(rsum - dist, 0.0, None)
        if ov.max() <= 1e-9:
            break
        f = -(ov / np.maximum(dist, 1e-12))[:, :, None] * d
        P = (P + 0.5 * f.sum(1)) % L
    return P


def tissue2d(ncells: int = 16, nv: int = 24, ratio: float = 1.4, phi: float = 0.6, Ka: float = 1.0,
             Kl: float = 1.0, Kb: float = 0.0, Kre: float = 1.0, Kat: float = 0.0, seed: int = 0,
             nv_range=None, calA: float = 1.0) -> Packing2D:
    """Bidisperse polygons (half r0 = 1, half r0 = sqrt(ratio)) at true
    packing fraction `phi` (the reference's L = sqrt(sum a)/phi0, so phi0 =
    sqrt(phi)), centres relaxed to no disk overlap; calA is the cells'
    shape parameter (Cell2D's third argument, cell.cpp:12-33)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = np.where(np.arange(ncells) % 2 == 0, 1.0, math.sqrt(ratio)).astype(np.float32)
    if nv_range is None:
        nvs = np.full(ncells, nv, np.int32)
    else:
        nvs = rng.integers(nv_range[0], nv_range[1] + 1, ncells).astype(np.int32)
    proto = [clDPM.Cell2D(0.0, 0.0, calA, int(nvs[c]), float(r[c])) for c in range(ncells)]
    T = clDPM.Tissue2D(proto, float(np.float32(math.sqrt(phi))))
    L = T.L
    P = _relax_disks(r.astype(np.float64), L, rng)
    Xs, a0, l0 = [], [], []
    for c in range(ncells):
        cell = clDPM.Cell2D(float(P[c, 0]), float(P[c, 1]), calA, int(nvs[c]), float(r[c]))
        Xs.append(np.asarray(cell.Verts, np.float32))
        a0.append(cell.a0)
        l0.append(cell.l0)
    f = lambda v: np.full(ncells, v, np.float32)
    return Packing2D(nv=nvs, X=np.concatenate(Xs), a0=np.asarray(a0, np.float32),
                     l0=np.asarray(l0, np.float32), r0=r, Ka=f(Ka), Kl=f(Kl), Kb=f(Kb), L=float(L),
                     Kre=Kre, Kat=Kat)


def packings2d(n: int = 512, cells=(8, 128), verts=(16, 64), seed: int = 0, phi: float = 0.4,
               calA: float = 1.02):
    """Batching form: n packings, cell counts U{cells}, vertices per cell
    U{verts}, bidisperse 1:1.4, Kb alternating 0 / 0.01, shape parameter
    calA (1.02: 2 % excess perimeter — with calA = 1 the regular polygon is a
    quartic minimum of the area + perimeter energy and FIRE converges only
    algebraically there)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n):
        nc = int(rng.integers(cells[0], cells[1] + 1))
        out.append(tissue2d(nc, 0, 1.4, phi, Kb=0.0 if k % 2 == 0 else 0.01,
                            seed=int(rng.integers(0, 2 ** 31)), nv_range=verts, calA=calA))
    return out
