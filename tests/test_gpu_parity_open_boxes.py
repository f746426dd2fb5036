"""Oracle parity on the code paths the headline 3D workload runs.

In a periodic box wider than about four cell diameters, neighbouring cells
need no periodic image ("no_image" in the step kernels).  Only there do the
step kernels skip the adhesion pair tests of candidate pairs whose
centre-line separation exceeds 2 l0, and test vertex pairs without the
image rounding.  The boxes here (6 x 6 and 8 x 8 monolayers, and the
256-cell monolayer itself) take those paths; every result is compared with
the oracle from the same float inputs:
  * forces: ||F_gpu - F_ref64|| / ||F_ref64|| <= 1e-5 and every vertex within
    1e-4 * max|F_ref64|, over the vertices where the float and the double
    restatement take every cut-off decision alike (at most 32 or 1e-3 of the
    vertices are set aside: the float step decides as ref32, and where ref64
    decides otherwise its force moves by O(Kat) or O(Ks));
  * contact (vertex, cell) pairs and facing-down faces: bit-exact vs ref32;
  * adhesion: per vertex the partner count and the key sums of the partners
    its own loop accepts and of the vertices whose loop accepts it
    (dpm3d_probe_adhesion), equal to ref32's.
Both step forms, with the cell list on (bins active) and off.
"""
import math

import numpy as np
import pytest

import oracle
from opencl_dpm_amd import capi
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
FORCE_VTOL = 1e-4


def log(*a):
    print(*a, flush=True)


def tissue3d(m, mask=capi.DPM3D_ALL):
    return capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat, Kb=m.Kb,
                          ks_grad=m.ks_grad, mask=mask, nv=m.mesh.nv)


def oracle_pair(m, X, lib="liboracle.so"):
    kw = dict(Kat=m.Kat, Kb=m.Kb, mask=capi.DPM3D_ALL, ks_grad=m.ks_grad, lib=lib)
    args = (X, m.mesh.faces, m.mesh.nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre)
    log(f"  oracle ref32 ({m.C} cells) ...")
    r32 = oracle.forces3d(*args, prec=32, pairs=True, **kw)
    log(f"  oracle ref64 ({m.C} cells) ...")
    r64 = oracle.forces3d(*args, prec=64, **kw)
    return r32, r64


def oracle_flips(m, X, r32, r64):
    """Vertices where the float and the double restatement take a cut-off
    decision of the force law differently at the same positions: adhesion
    partner counts (the vertex, and every vertex of another cell within 2 l0
    of it, which receives the pair's reaction), facing-down faces (their
    corners), and the substrate band |z| < 2 l0 at the corners of
    facing-down faces (l0 = sqrt(4 a0) / sqrt(3) rounded to float in one and
    to double in the other).  The float step decides as ref32 does (the pair
    sets below are compared bit for bit); at such a vertex its force moves
    by O(Kat) or O(Ks) from ref64's."""
    nv = m.mesh.nv
    faces = np.asarray(m.mesh.faces)
    flip = np.zeros(len(X), bool)
    fa = np.asarray(r32["adhesion"]) != np.asarray(r64["adhesion"])
    if fa.any():
        from scipy.spatial import cKDTree
        L = float(m.L)
        P = np.asarray(X, np.float64).copy()
        P[:, :2] = np.mod(P[:, :2], L)
        P[:, 2] -= P[:, 2].min()
        tree = cKDTree(P, boxsize=[L, L, P[:, 2].max() + 10.0])
        l0 = np.sqrt(4 * float(np.max(m.a0))) / np.sqrt(3)
        for i in np.nonzero(fa)[0]:
            flip[tree.query_ball_point(P[i], 2 * l0 * 1.01)] = True
    d32, d64 = np.asarray(r32["down"]) != 0, np.asarray(r64["down"]) != 0
    for c, f in zip(*np.nonzero(d32 != d64)):
        flip[c * nv + faces[f]] = True
    a0 = np.broadcast_to(np.asarray(m.a0, np.float32), (m.C,))
    band32 = (np.sqrt(np.float32(4) * a0) / np.sqrt(np.float32(3))) * np.float32(2)
    band64 = (np.sqrt(4.0 * a0.astype(np.float64)) / np.sqrt(3.0)) * 2.0
    z = np.abs(np.asarray(X, np.float32)[:, 2]).reshape(m.C, nv)
    differs = (z < band32[:, None]) != (z.astype(np.float64) < band64[:, None])
    corner = np.zeros((m.C, nv), bool)
    cd, fd = np.nonzero(d32 | d64)
    corner[cd[:, None], faces[fd]] = True
    flip |= (differs & corner).ravel()
    return flip


def check(g, r32, r64, what, flip):
    F64 = r64["forces"]
    Fg = np.asarray(g["forces"], np.float64)
    keep = ~flip
    r = float(np.linalg.norm(Fg[keep] - F64[keep]) / np.linalg.norm(F64[keep]))
    vmax = float(np.abs(Fg[keep] - F64[keep]).max())
    scale = float(np.abs(F64).max())
    log(f"  {what}: rel={r:.3e} vmax={vmax:.3e} scale={scale:.3e} contact={int(g['contact'].sum())} "
        f"down={int(g['down'].sum())} adhesion={int(g['adhesion'].sum())}; "
        f"vertices at a float/double cut-off flip {int(flip.sum())}")
    assert flip.sum() <= max(32, 1e-3 * len(flip)), (what, int(flip.sum()))
    assert r <= FORCE_RTOL, (what, r)
    assert vmax <= FORCE_VTOL * scale, (what, vmax, scale)
    np.testing.assert_array_equal(g["contact"], r32["contact"])
    np.testing.assert_array_equal(g["down"], r32["down"])
    np.testing.assert_array_equal(g["adhesion"], r32["adhesion"])
    np.testing.assert_array_equal(g["adhesion_hash"], r32["adhesion_hash"])


def coverage(m, X):
    """Candidate cell pairs (centres within R_i + R_j + 2 l0) that need no
    periodic image by the kernels' test, and those among them whose
    centre-line separation (the kernels' bound D - h_i - h_j, h the largest
    extent of each cell toward the other) is at least 2 l0: the pairs whose
    adhesion tests the step skips."""
    nv = m.mesh.nv
    P = np.asarray(X, np.float64).reshape(m.C, nv, 3)
    C = P.mean(1)
    R = np.sqrt(((P - C[:, None]) ** 2).sum(-1)).max(1)
    l0 = np.sqrt(4 * np.asarray(np.broadcast_to(m.a0, (m.C,)), np.float64)) / math.sqrt(3)
    L = float(m.L)
    no_image = skip = 0
    for i in range(m.C):
        for j in range(m.C):
            if i == j:
                continue
            d = C[j] - C[i]
            d -= L * np.round(d / L)
            l0m = max(l0[i], l0[j])
            D = float(np.linalg.norm(d))
            if D > R[i] + R[j] + 2 * l0m:
                continue
            if np.all(np.abs(C[i] - C[j]) + (R[i] + R[j]) < 0.48 * L):
                no_image += 1
                n = d / D
                hi = float(((P[i] - C[i]) @ n).max())
                hj = float(((C[j] - P[j]) @ n).max())
                skip += (D - hi - hj) - 1e-4 * (R[i] + R[j] + D) >= 2 * l0m * 1.001
    return no_image, skip


def expected_bins(m, X):
    """Bins per axis the cell list uses (k3d_bin: floor(L / reach), reach the
    largest candidate reach of the step, 0 below 3)."""
    nv = m.mesh.nv
    P = np.asarray(X, np.float64).reshape(m.C, nv, 3)
    C = P.mean(1)
    rmax = np.sqrt(((P - C[:, None]) ** 2).sum(-1)).max()
    F = np.asarray(m.mesh.faces)
    e = np.concatenate([P[:, F[:, 1]] - P[:, F[:, 0]], P[:, F[:, 2]] - P[:, F[:, 1]], P[:, F[:, 0]] - P[:, F[:, 2]]], 1)
    emax = np.sqrt((e ** 2).sum(-1)).max()
    l0max = math.sqrt(4 * float(np.max(m.a0))) / math.sqrt(3)
    reach = max(2 * rmax + 0.4004 * emax, 2 * rmax + 2 * l0max) * 1.0001 * 1.001
    nb = int(math.floor(float(m.L) / reach))
    return nb if nb >= 3 else 0


OPEN_BOXES = {
    # name: (n_side, spacing, seed)
    "6x6_spacing2.4": (6, 2.4, 60),
    "8x8_spacing2.4": (8, 2.4, 61),
    "8x8_spacing2.1": (8, 2.1, 62),  # squeezed: overlapping neighbours, contacts
}


@pytest.mark.parametrize("case", list(OPEN_BOXES))
def test_open_box_start_and_after_200_and_2000_steps(case):
    # as the cells spread on the substrate their bounding radii grow and
    # diagonal neighbours become candidates whose surfaces stay apart: the
    # separation skip appears along the trajectory
    n, sp, seed = OPEN_BOXES[case]
    m = W.monolayer3d(n, sp, seed=seed)
    run = tissue3d(m)
    run.set_positions(m.X)
    states = {"start": m.X}
    run.euler(200, 0.002)
    states["after 200 steps"] = run.positions()
    run.euler(1800, 0.002)
    states["after 2000 steps"] = run.positions()
    skipped = adhesion = contact = 0
    for label, X in states.items():
        ni, sk = coverage(m, X)
        nb = expected_bins(m, X)
        skipped += sk
        log(f"{case} {label}: candidate pairs without image {ni}, adhesion tests skipped by separation {sk}; "
            f"bins {nb}")
        assert ni > 0 and nb >= 3
        r32, r64 = oracle_pair(m, X)
        flip = oracle_flips(m, X, r32, r64)
        for kernel in (1, 2):
            for cl in (0, 1):
                t = tissue3d(m)
                t.set_step_kernel(kernel)
                t.set_cell_list(cl)
                t.set_positions(X)
                check(t.probe(), r32, r64, f"{case} {label} kernel={kernel} cell_list={cl}", flip)
        adhesion += int(r32["adhesion"].sum())
        contact += int(r32["contact"].sum())
    # along the trajectory: adhesion pairs, the separation skip on pairs the
    # oracle checked, and contacts where the cells start squeezed
    assert adhesion > 0 and skipped > 0
    if sp < 2.4:
        assert contact > 0


@pytest.mark.timeout(1200)
def test_headline_monolayer_start_and_after_200_steps():
    # the 256-cell x 162-vertex monolayer of the benchmark, all terms, at its
    # start and after 200 steps; the oracle runs all pairs (OpenMP build)
    m = W.monolayer_config("cfg2", seed=0)
    run = tissue3d(m)
    run.set_positions(m.X)
    run.euler(200, m.dt)
    for label, X in (("start", m.X), ("after 200 steps", run.positions())):
        ni, sk = coverage(m, X)
        log(f"cfg2 {label}: candidate pairs without image {ni}, adhesion tests skipped by separation {sk}; "
            f"bins {expected_bins(m, X)}")
        assert ni > 0
        r32, r64 = oracle_pair(m, X, lib="liboracle_omp.so")
        flip = oracle_flips(m, X, r32, r64)
        for kernel in (1, 2):
            for cl in (0, 1):
                t = tissue3d(m)
                t.set_step_kernel(kernel)
                t.set_cell_list(cl)
                t.set_positions(X)
                check(t.probe(), r32, r64, f"cfg2 {label} kernel={kernel} cell_list={cl}", flip)


@pytest.mark.parametrize("kernel", [1, 2])
def test_probe_instance_equals_production_step(kernel):
    # probes of 162-vertex meshes run the step's probe instance (the same
    # source with the diagnostic outputs compiled in); its forces are the
    # production kernel's bit for bit (dpm_euler's last-step forces)
    m = W.monolayer3d(6, 2.1, seed=63)
    t = tissue3d(m)
    t.set_step_kernel(kernel)
    t.set_positions(m.X)
    t.euler(50, 0.002)
    X = t.positions()
    for cl in (0, 1):
        t.set_cell_list(cl)
        t.set_positions(X)
        g = t.probe()
        t.set_positions(X)
        t.euler(1, 0.002)
        np.testing.assert_array_equal(g["forces"], t.forces())
        log(f"kernel={kernel} cell_list={cl}: probe forces = step forces (contacts {int(g['contact'].sum())}, "
            f"adhesion {int(g['adhesion'].sum())})")
