"""The drop-in module `clDPM` on the GPU: the reference's own driver scripts
replayed, dispersal on the device, positions kept on the device across
calls, the host reductions on the device, and plot.py's data access.

Bars: dispersal and host reductions bit-identical to the reference's own
host code (tests/golden/*.npz, written from src/cell.cpp, src/Tissue2D.cpp,
src/Tissue3D.cpp); pair sets bit-exact vs the oracle's ref32; forces vs
ref64 within 1e-5 relative; trajectories within the reference arithmetic's
own float drift (ref32 vs ref64) or the rounding bound of test_gpu_parity.
"""
import os

import numpy as np
import pytest

import oracle
from opencl_dpm_amd import capi, clDPM
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GLIBC_DRAND48_STATE = (0x330E, 0xABCD, 0x1234)  # what an unseeded reference run draws


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def log(*a):
    print(*a, flush=True)


def traj_bound(X0, X64, nsteps):
    return 1e-6 * float(np.abs(X64).max()) * np.sqrt(nsteps) + 1e-5 * float(np.abs(X64 - X0).max())


@pytest.fixture(scope="module")
def drivers():
    return np.load(os.path.join(ROOT, "tests", "golden", "driver_goldens.npz"))


# ------------------------------------------------------------ dispersal

def test_dispersal_on_device_matches_reference_host_code(goldens, drivers):
    # Tissue2D.Disperse / Tissue3D.Disperse2D run their relaxation on the
    # device; from the same drand48 state the centres are the reference
    # host code's, bit for bit
    cells = [clDPM.Cell2D(0, 0, float(a), int(n), float(r))
             for a, n, r in zip(goldens["t2d_calA"], goldens["t2d_nv"], goldens["t2d_r0"])]
    T = clDPM.Tissue2D(cells, float(goldens["t2d_phi0"]))
    clDPM.seed48(*GLIBC_DRAND48_STATE)
    T.Disperse()
    v = np.concatenate([np.asarray(c.Verts, np.float32) for c in T.Cells])
    np.testing.assert_array_equal(bits(v), bits(goldens["t2d_verts"]))
    cells = [clDPM.Cell3D([0, 0, 0], float(a), float(r)) for a, r in zip(goldens["t3d_calA"], goldens["t3d_r0"])]
    T3 = clDPM.Tissue3D(cells, float(goldens["t3d_phi0"]))
    clDPM.seed48(1, 2, 3)
    T3.Disperse2D()
    v3 = np.stack([np.asarray(c.Verts, np.float32) for c in T3.Cells])
    np.testing.assert_array_equal(bits(v3), bits(goldens["t3d_verts"]))


def test_dispersal_of_the_reference_drivers(drivers):
    # test2D.py: 400 cells (25 / 22 vertices) at 0.9; test3D.py: 32 cells at 0.35
    c, c2 = clDPM.Cell2D(0, 0, 1.2, 25, 1.0), clDPM.Cell2D(0, 0, 1.2, 22, 1.3)
    T = clDPM.Tissue2D([c, c2] * 200, 0.9)
    clDPM.seed48(*GLIBC_DRAND48_STATE)
    T.Disperse()
    v = np.concatenate([np.asarray(x.Verts, np.float32) for x in T.Cells])
    np.testing.assert_array_equal(bits(v), bits(drivers["d2_verts"]))
    c3 = clDPM.Cell3D([0.0] * 3, 1.0, 1.0)
    T3 = clDPM.Tissue3D([c3] * 32, 0.35)
    clDPM.seed48(*GLIBC_DRAND48_STATE)
    T3.Disperse2D()
    v3 = np.stack([np.asarray(x.Verts, np.float32) for x in T3.Cells])
    np.testing.assert_array_equal(bits(v3), bits(drivers["d3_verts"]))


def test_dispersal_entry_point_runs_to_the_reference_stopping_rule():
    # random starts in a box wide against the radii (bin grid) and in one
    # under 3 bins per axis (every centre scans all): deterministic, finite,
    # and max_count caps the loop at max_count + 1 iterations as the
    # reference's `count > 1e5` does
    rng = np.random.default_rng(3)
    for n, L in ((300, 40.0), (40, 6.0)):
        r = rng.uniform(0.8, 1.3, n).astype(np.float32)
        X = (rng.random(n) * L).astype(np.float32)
        Y = (rng.random(n) * L).astype(np.float32)
        X1, Y1, (it, capped) = capi.disperse(r, L, X, Y)
        log(f"n={n} L={L}: {it} iterations, capped={capped}")
        assert it >= 2 and np.isfinite(X1).all() and np.isfinite(Y1).all()
        X2, Y2, (it2, _) = capi.disperse(r, L, X, Y)
        np.testing.assert_array_equal(X1, X2)  # deterministic
        assert it2 == it
    X3, Y3, (it3, capped3) = capi.disperse(r, L, X, Y, max_count=5)
    assert it3 == 6 and capped3


# -------------------------------------------------- positions across calls

def make_3d(n=3, sp=1.8, seed=6):
    m = W.monolayer3d(n, sp, seed=seed, keep_cells=True)
    for c in m.cells:
        c.Kv, c.Ka, c.Ks, c.Kb = m.Kv, m.Ka, m.Ks, m.Kb
    T = clDPM.Tissue3D(m.cells, 0.35)
    T.Kre, T.Kat, T.VertexAdhesion, T.Verbose = m.Kre, m.Kat, True, False
    return m, T


def test_3d_calls_in_chunks_equal_one_call():
    # five calls of 20 steps keep the positions on the device between calls
    # (no re-upload) and give the bits of one call of 100 steps
    m, T = make_3d()
    for _ in range(5):
        T.CLEulerUpdate(20, 0.002)
    Xc = np.concatenate([np.asarray(c.Verts, np.float32) for c in T.Cells])
    Fc = np.concatenate([np.asarray(c.Forces, np.float32) for c in T.Cells])
    m2, T2 = make_3d()
    T2.CLEulerUpdate(100, 0.002)
    np.testing.assert_array_equal(Xc, np.concatenate([np.asarray(c.Verts, np.float32) for c in T2.Cells]))
    np.testing.assert_array_equal(Fc, np.concatenate([np.asarray(c.Forces, np.float32) for c in T2.Cells]))


def test_3d_positions_assigned_between_calls_are_used():
    m, T = make_3d()
    T.CLEulerUpdate(10, 0.002)
    v = np.asarray(T.Cells[2].Verts, np.float32)
    v[:, 0] += np.float32(0.05)
    T.Cells[2].Verts = v.tolist()  # marks the cell's positions changed: uploaded
    X = np.concatenate([np.asarray(c.Verts, np.float32) for c in T.Cells])
    T.CLEulerUpdate(10, 0.002)
    t = capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, T.L, m.Kre, Kat=m.Kat, Kb=m.Kb,
                       mask=capi.DPM3D_ALL, nv=m.mesh.nv)
    t.set_positions(X)
    t.euler(10, 0.002)
    np.testing.assert_array_equal(np.concatenate([np.asarray(c.Verts, np.float32) for c in T.Cells]),
                                  t.positions())


def test_2d_calls_in_chunks_equal_one_call():
    p = W.tissue2d(16, 24, 1.4, 0.8, Kb=0.01, Kat=0.3, seed=9)

    def tissue():
        cells = []
        off = 0
        for c in range(p.C):
            cell = clDPM.Cell2D(0.0, 0.0, 1.0, int(p.nv[c]), float(p.r0[c]))
            cell.Verts = p.X[off:off + p.nv[c]].tolist()
            cell.Ka, cell.Kl, cell.Kb = 1.0, 1.0, 0.01
            cells.append(cell)
            off += p.nv[c]
        T = clDPM.Tissue2D(cells, 0.8)
        T.Kat = 0.3
        return T

    A, B = tissue(), tissue()
    for _ in range(4):
        A.CLEulerUpdate(50, 0.005)
    B.CLEulerUpdate(200, 0.005)
    for key in ("Verts", "Forces"):
        np.testing.assert_array_equal(np.concatenate([np.asarray(getattr(c, key), np.float32) for c in A.Cells]),
                                      np.concatenate([np.asarray(getattr(c, key), np.float32) for c in B.Cells]))


# ------------------------------------------------------- host reductions

def test_cell_measures_are_the_host_reductions(goldens):
    # the reference's GetVolume / GetSurfaceArea / GetCOM on perturbed
    # vertices (golden from src/cell.cpp), evaluated on the device
    c = clDPM.Cell3D([0.0, 0.0, 0.0], 1.0, 1.0)
    faces = np.asarray(c.GetFaces())
    t = capi.Tissues3D(faces, [0, 1], 1.0, 1.0, 1.0, 4.0, 0.1, 100.0, 1.0, nv=162)
    t.set_positions(goldens["measure_verts"])
    vol, sa, com = t.cell_measures()
    out = goldens["measure_out"]
    assert bits(vol[0]) == bits(out[0]) and bits(sa[0]) == bits(out[1])
    np.testing.assert_array_equal(bits(com[0]), bits(out[2:5]))


def test_dropin_volume_and_area_are_the_host_functions():
    # after a call, Volume and SurfaceArea (Tissue3D.cpp:503-504) equal the
    # host functions on the positions handed back, bit for bit
    m, T = make_3d(4, 1.9, seed=8)
    T.CLEulerUpdate(30, 0.002)
    for c in T.Cells:
        assert bits(c.Volume) == bits(c.GetVolume())
        assert bits(c.SurfaceArea) == bits(c.GetSurfaceArea())


# ------------------------------------------- the reference's driver scripts

def plot_like_2d(T, path):
    """The data plot.py:7-23 (Plot2DTissue2D) draws: per cell the vertices
    wrapped into the box, coloured by the force sum, the outline unless it
    straddles the box."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig = plt.figure(figsize=(4, 4))
    drawn = 0
    for ci in range(T.NCELLS):
        pos = np.mod(np.array(T.Cells[ci].Verts).T, T.L)
        f = np.array(T.Cells[ci].Forces).T
        assert pos.shape == (2, T.Cells[ci].NV) and f.shape == (2, T.Cells[ci].NV)
        plt.scatter(pos[0], pos[1], c=f[0] + f[1], s=1)
        if np.ptp(pos[0]) <= T.L / 2 and np.ptp(pos[1]) <= T.L / 2:
            plt.plot(pos[0], pos[1], "-k", lw=0.3)
            drawn += 1
    fig.savefig(path)
    plt.close(fig)
    return drawn


def plot_like_3d(T, path):
    """The data plot.py:46-67 (Plot3DTissue2D) draws: the faces of cell 0's
    topology over every cell's positions wrapped into the box."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    faces = T.Cells[0].GetFaces()
    assert np.asarray(faces).shape == (320, 3)
    fig = plt.figure(figsize=(4, 4))
    for ci in range(T.NCELLS):
        pos = T.Cells[ci].GetPositions()
        assert np.asarray(pos).shape == (3, 162) and not np.isnan(pos).any()
        x, y = np.mod(pos[0], T.L), np.mod(pos[1], T.L)
        for face in faces[::16]:
            vx, vy = [x[i] for i in face], [y[i] for i in face]
            if max(vx) - min(vx) <= T.L / 2 and max(vy) - min(vy) <= T.L / 2:
                plt.plot(vx, vy, lw=0.2)
        plt.scatter(x, y, s=1)
    fig.savefig(path)
    plt.close(fig)


def test_reference_test3d_driver(drivers, tmp_path):
    # test3D.py:8-22: [Cell3D([0,0,0],1,1)] * 32 with Ka=2, Kv=5, Ks=3,
    # Tissue3D(cells, 0.35), Kre=25, Disperse2D (cells end at negative x, y
    # and half below the substrate), CLEulerUpdate(25, 0.01), then the plot
    c = clDPM.Cell3D([0.0] * 3, 1.0, 1.0)
    c.Ka, c.Kv, c.Ks = 2.0, 5.0, 3.0
    T = clDPM.Tissue3D([c] * 32, 0.35)
    T.Kre = 25.0
    T.Verbose = False
    clDPM.seed48(*GLIBC_DRAND48_STATE)
    T.Disperse2D()
    X0 = np.concatenate([np.asarray(x.Verts, np.float32) for x in T.Cells])
    np.testing.assert_array_equal(bits(X0.reshape(32, 162, 3)), bits(drivers["d3_verts"]))
    faces = np.asarray(T.Cells[0].GetFaces())
    nv = 162
    v0 = np.float32(T.Cells[0].v0)
    a0 = np.float32(T.Cells[0].a0)
    args = (faces, nv, 5.0, 2.0, 3.0, v0, a0, T.L, 25.0)
    # pair sets and forces at the dispersed start
    t = capi.Tissues3D(faces, [0, 32], 5.0, 2.0, 3.0, v0, a0, T.L, 25.0, mask=capi.DPM3D_REFERENCE, nv=nv)
    t.set_positions(X0)
    g = t.probe()
    r32 = oracle.forces3d(X0, *args, mask=15, prec=32, pairs=True)
    r64 = oracle.forces3d(X0, *args, mask=15, prec=64)
    rel = float(np.linalg.norm(g["forces"] - r64["forces"]) / np.linalg.norm(r64["forces"]))
    log(f"test3D start: rel {rel:.3e}, contacts {int(g['contact'].sum())}, down {int(g['down'].sum())}")
    assert rel <= 1e-5
    np.testing.assert_array_equal(g["contact"], r32["contact"])
    np.testing.assert_array_equal(g["down"], r32["down"])
    assert g["down"].sum() > 0
    # the driver's call, against both oracle trajectories
    T.CLEulerUpdate(25, 0.01)
    X = np.concatenate([np.asarray(x.Verts, np.float32) for x in T.Cells])
    X64, _ = oracle.run3d(X0, *args, 25, 0.01, mask=15, prec=64)
    X32, _ = oracle.run3d(X0, *args, 25, 0.01, mask=15, prec=32)
    eg = np.abs(X - X64).max(1)
    e32 = np.abs(X32 - X64).max(1)
    log(f"test3D 25 steps: gpu vs ref64 max {eg.max():.3e}; ref32 vs ref64 max {e32.max():.3e}")
    assert eg.max() <= max(2 * e32.max(), traj_bound(X0, X64, 25))
    for x in T.Cells:
        assert x.Volume > 0 and x.SurfaceArea > 0
    plot_like_3d(T, str(tmp_path / "test3d.png"))
    assert (tmp_path / "test3d.png").stat().st_size > 0


def test_reference_test2d_driver(drivers, tmp_path):
    # test2D.py:9-31: [Cell2D(0,0,1.2,25,1.0), Cell2D(0,0,1.2,22,1.3)] * 200,
    # Ka=0.1, Kl=1, Kb=0.05, Tissue2D(cells, 0.9), Kre=1, Kat=0.5, Disperse,
    # then 500-step calls of CLEulerUpdate(500, 0.005) with a plot after each
    c, c2 = clDPM.Cell2D(0, 0, 1.2, 25, 1.0), clDPM.Cell2D(0, 0, 1.2, 22, 1.3)
    for x in (c, c2):
        x.Ka, x.Kl, x.Kb = 0.1, 1.0, 0.05
    T = clDPM.Tissue2D([c, c2] * 200, 0.9)
    T.Kre, T.Kat = 1.0, 0.5
    clDPM.seed48(*GLIBC_DRAND48_STATE)
    T.Disperse()
    X0 = np.concatenate([np.asarray(x.Verts, np.float32) for x in T.Cells])
    np.testing.assert_array_equal(bits(X0), bits(drivers["d2_verts"]))
    nvs = np.array([x.NV for x in T.Cells], np.int32)
    a0 = np.array([x.a0 for x in T.Cells], np.float32)
    l0 = np.array([x.l0 for x in T.Cells], np.float32)
    r0 = np.array([x.r0 for x in T.Cells], np.float32)
    # pair sets and forces at the dispersed start, against the oracle
    t = capi.Tissues2D([0, 400], nvs, 0.1, 1.0, 0.05, a0, l0, r0, T.L, Kre=1.0, Kat=0.5)
    t.set_positions(X0)
    g = t.probe()
    stride = int(nvs.max())
    P = np.zeros((400 * stride, 2), np.float32)
    off = 0
    for ci in range(400):
        P[ci * stride:ci * stride + nvs[ci]] = X0[off:off + nvs[ci]]
        off += nvs[ci]
    keep = np.concatenate([np.arange(ci * stride, ci * stride + nvs[ci]) for ci in range(400)])
    kw = dict(Ka=0.1, Kl=1.0, Kb=0.05, a0=a0, l0=l0, r0=r0, L=T.L, Kre=1.0, Kat=0.5)
    log("test2D: oracle at the start ...")
    F32, ov32, at32, ah32 = oracle.forces2d(P, stride, nvs, prec=32, attr_hash=True, **kw)
    F64, _, _ = oracle.forces2d(P, stride, nvs, prec=64, **kw)
    F64 = F64[keep]
    rel = float(np.linalg.norm(g["forces"] - F64) / np.linalg.norm(F64))
    log(f"test2D start: rel {rel:.3e}, overlaps {int(ov32[keep].sum())}, attraction {int(at32[keep].sum())}")
    assert rel <= 1e-5
    np.testing.assert_array_equal(g["overlap"], ov32[keep])
    np.testing.assert_array_equal(g["attraction"], at32[keep])
    np.testing.assert_array_equal(g["attraction_hash"], ah32[keep])
    # two of the driver's 500-step calls, each followed by the plot's reads;
    # the chunked calls equal one 1000-step run of the C ABI, bit for bit
    for i in range(2):
        T.CLEulerUpdate(500, 0.005)
        drawn = plot_like_2d(T, str(tmp_path / f"2d_{i}.png"))
        log(f"test2D call {i}: outlines drawn {drawn} of 400")
    X = np.concatenate([np.asarray(x.Verts, np.float32) for x in T.Cells])
    assert np.isfinite(X).all()
    t.set_positions(X0)
    t.euler(1000, 0.005)
    np.testing.assert_array_equal(X, t.positions())
