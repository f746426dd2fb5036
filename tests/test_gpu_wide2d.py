"""The 2D step in its two forms: one workgroup per tissue (k2d_run) and the
wide form over the whole device (k2d_wide_*), for tissues beyond one
workgroup's LDS.  The forms must agree bit for bit (forces, overlap and
attraction pair sets, trajectories) on tissues both take, including the
crossing-parity decisions that k2d_run spreads over (vertex, candidate)
pairs and the wide form evaluates per vertex; a tissue beyond one workgroup
is checked against the oracle (pair sets bit-exact vs ref32, forces within
1e-5 of ref64)."""
import numpy as np
import pytest

import oracle
from opencl_dpm_amd import capi
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu


def handle(p):
    return capi.Tissues2D([0, p.C], p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, Kre=p.Kre, Kat=p.Kat)


CASES = {
    # name: (ncells, nv, phi, Kb, Kat, seed, nv_range)
    "cfg1": (16, 24, 0.6, 0.0, 0.0, 0, None),
    "dense_attr": (16, 24, 0.92, 0.01, 0.5, 1, None),
    "uneven": (40, 0, 0.88, 0.01, 0.3, 2, (16, 64)),
    "400_cells": (400, 0, 0.85, 0.05, 0.5, 3, (22, 25)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_wide_form_equals_one_workgroup(case):
    nc, nv, phi, Kb, Kat, seed, rng = CASES[case]
    p = W.tissue2d(nc, nv, 1.4, phi, Kb=Kb, Kat=Kat, seed=seed, nv_range=rng)
    out = []
    for kind in (1, 2):
        t = handle(p)
        t.set_step_kernel(kind)
        t.set_positions(p.X)
        g = t.probe()
        t.euler(200, 0.005)
        out.append((g, t.positions(), t.forces()))
    (g1, X1, F1), (g2, X2, F2) = out
    for key in ("forces", "overlap", "attraction", "attraction_hash"):
        np.testing.assert_array_equal(g1[key], g2[key], err_msg=f"{case} {key}")
    np.testing.assert_array_equal(X1, X2)
    np.testing.assert_array_equal(F1, F2)
    print(f"{case}: overlaps {int(g1['overlap'].sum())}, attraction {int(g1['attraction'].sum())}; "
          f"200 steps identical", flush=True)


def lattice_tissue(nx, ny, nv, spacing, jitter, seed):
    """Bidisperse polygons on a jittered square lattice (periodic box
    nx*spacing x ny*spacing), spacing below the mean diameter so that
    neighbours overlap: a large tissue built without a relaxation."""
    from opencl_dpm_amd import clDPM
    rng = np.random.Generator(np.random.PCG64(seed))
    C = nx * ny
    r = np.where(np.arange(C) % 2 == 0, 1.0, np.sqrt(1.4)).astype(np.float32)
    Xs, a0, l0 = [], [], []
    for c in range(C):
        cx = (c % nx + 0.5) * spacing + rng.uniform(-jitter, jitter)
        cy = (c // nx + 0.5) * spacing + rng.uniform(-jitter, jitter)
        cell = clDPM.Cell2D(float(cx), float(cy), 1.0, nv, float(r[c]))
        Xs.append(np.asarray(cell.Verts, np.float32))
        a0.append(cell.a0)
        l0.append(cell.l0)
    f = lambda v: np.full(C, v, np.float32)
    return W.Packing2D(nv=np.full(C, nv, np.int32), X=np.concatenate(Xs), a0=np.asarray(a0, np.float32),
                       l0=np.asarray(l0, np.float32), r0=r, Ka=f(1.0), Kl=f(1.0), Kb=f(0.01),
                       L=float(np.float32(nx * spacing)), Kre=1.0, Kat=0.5)


def test_tissue_beyond_one_workgroup_vs_oracle():
    # 600 cells x 16 vertices on a 30 x 20 lattice: beyond one workgroup's
    # LDS, so the handle takes the wide form
    p = lattice_tissue(30, 20, 16, 2.2, 0.15, seed=4)
    assert p.L == float(np.float32(30 * 2.2))
    t = handle(p)
    with pytest.raises(ValueError):
        t.set_step_kernel(1)
    t.set_positions(p.X)
    g = t.probe()
    P, stride = p.padded()
    keep = np.concatenate([np.arange(c * stride, c * stride + p.nv[c]) for c in range(p.C)])
    args = (P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat)
    print("oracle ref32 ...", flush=True)
    F32, ov, at, ah = oracle.forces2d(*args, prec=32, attr_hash=True, lib="liboracle_omp.so")
    print("oracle ref64 ...", flush=True)
    F64 = oracle.forces2d(*args, prec=64, lib="liboracle_omp.so")[0][keep]
    rel = float(np.linalg.norm(g["forces"] - F64) / np.linalg.norm(F64))
    print(f"600 cells: rel {rel:.3e}, overlaps {int(ov[keep].sum())}, attraction {int(at[keep].sum())}", flush=True)
    assert rel <= 1e-5
    np.testing.assert_array_equal(g["overlap"], ov[keep])
    np.testing.assert_array_equal(g["attraction"], at[keep])
    np.testing.assert_array_equal(g["attraction_hash"], ah[keep])
    assert ov[keep].sum() > 0 and at[keep].sum() > 0
    t.euler(150, 0.005)  # two 64-step graphs and eager steps
    X = t.positions()
    assert np.isfinite(X).all()
    t.set_positions(p.X)
    for _ in range(3):
        t.euler(50, 0.005)
    np.testing.assert_array_equal(t.positions(), X)
