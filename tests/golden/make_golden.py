"""Writes tests/golden/setup_goldens.npz from the reference's own host code.

The reference's src/cell.cpp, src/Tissue2D.cpp and src/Tissue3D.cpp are
compiled where they lie by `make -C oracle ref` into oracle/_ref/libdpmref.so
(with the C-ABI shim oracle/ref_shim.cpp).  This script runs those set-up
paths — Cell2D/Cell3D construction, the host reductions and the two
dispersal routines with a fixed drand48 state — and stores inputs and
outputs as arrays.  Run it where the reference checkout exists:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFDIR = os.environ.get("DPM_REFERENCE_DIR", "/root/reference")
LIB = os.path.join(ROOT, "oracle", "_ref", "libdpmref.so")


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def main():
    L = ctypes.CDLL(LIB)
    out = {}
    # Cell3D(sp, calA, r0)
    c3 = np.array([[1, 2, 3, 1.0, 1.0], [7.0, 6.0, 1.3, 1.05, 1.8], [0, 0, 0, 1.2, 0.93],
                   [-4.5, 12.25, 0.95, 1.0, 0.95], [3.0, 3.0, 1.1, 1.0, 1.1]], np.float32)
    V, F, S = [], [], []
    for row in c3:
        v = np.zeros(162 * 3, np.float32)
        f = np.zeros(320 * 3, np.int32)
        s = np.zeros(7, np.float32)
        sp = np.ascontiguousarray(row[:3])
        L.refshim_cell3d(p(sp), ctypes.c_float(row[3]), ctypes.c_float(row[4]), p(v), p(f), p(s))
        V.append(v.reshape(162, 3))
        F.append(f.reshape(320, 3))
        S.append(s)
    out["cell3d_args"] = c3
    out["cell3d_verts"] = np.stack(V)
    out["cell3d_faces"] = F[0]
    out["cell3d_scalars"] = np.stack(S)  # v0, sa0, a0, Volume, SurfaceArea, calA0, r0
    # host reductions on perturbed vertices
    rng = np.random.Generator(np.random.PCG64(7))
    mv = (V[1] + rng.normal(0, 0.05, V[1].shape)).astype(np.float32)
    m5 = np.zeros(5, np.float32)
    L.refshim_cell3d_measure(p(np.ascontiguousarray(mv)), p(m5))
    out["measure_verts"] = mv
    out["measure_out"] = m5  # GetVolume, GetSurfaceArea, GetCOM xyz
    # Cell2D(x0, y0, calA, NV, r0)
    c2 = [(0, 0, 1.2, 25, 1.0), (0, 0, 1.2, 22, 1.3), (3.5, -2.25, 1.0, 24, 1.0),
          (10.0, 4.0, 1.05, 32, 1.0), (1.0, 1.0, 1.05, 40, 1.5), (0, 0, 1.0, 16, 1.1832159566199232)]
    for i, (x0, y0, ca, nv, r0) in enumerate(c2):
        v = np.zeros(nv * 2, np.float32)
        s = np.zeros(4, np.float32)
        L.refshim_cell2d(ctypes.c_float(x0), ctypes.c_float(y0), ctypes.c_float(ca), ctypes.c_uint(nv),
                         ctypes.c_float(r0), p(v), p(s))
        out[f"cell2d_{i}_args"] = np.array([x0, y0, ca, nv, r0], np.float64)
        out[f"cell2d_{i}_verts"] = v.reshape(nv, 2)
        out[f"cell2d_{i}_scalars"] = s  # a0, l0, calA0, GetArea
    # Tissue2D(cells, phi0).Disperse() with drand48 state (0x330E, 0xABCD, 0x1234)
    # = glibc's default, i.e. what an unseeded reference run draws.
    refdir = REFDIR.encode()
    n2 = 12
    calA = np.full(n2, 1.0, np.float32)
    nv = np.array([24, 20] * (n2 // 2), np.uint32)
    r0 = np.array([1.0, 1.2] * (n2 // 2), np.float32)
    Lout = np.zeros(1, np.float32)
    verts = np.zeros(int(nv.sum()) * 2, np.float32)
    L.refshim_seed48(ctypes.c_ushort(0x330E), ctypes.c_ushort(0xABCD), ctypes.c_ushort(0x1234))
    rc = L.refshim_tissue2d_disperse(n2, p(calA), p(nv), p(r0), ctypes.c_float(0.8), refdir,
                                     p(Lout), p(verts))
    assert rc == 0, rc
    out["t2d_calA"], out["t2d_nv"], out["t2d_r0"] = calA, nv.astype(np.int32), r0
    out["t2d_phi0"] = np.float32(0.8)
    out["t2d_L"] = Lout.copy()
    out["t2d_verts"] = verts.reshape(-1, 2)
    # Tissue3D(cells, phi0).Disperse2D() with seed48(1, 2, 3)
    n3 = 6
    calA3 = np.ones(n3, np.float32)
    r03 = np.array([1.0, 0.9, 1.1, 1.0, 0.95, 1.05], np.float32)
    verts3 = np.zeros(n3 * 162 * 3, np.float32)
    L.refshim_seed48(ctypes.c_ushort(1), ctypes.c_ushort(2), ctypes.c_ushort(3))
    rc = L.refshim_tissue3d_disperse2d(n3, p(calA3), p(r03), ctypes.c_float(0.35), refdir, p(Lout),
                                       p(verts3))
    assert rc == 0, rc
    out["t3d_calA"], out["t3d_r0"], out["t3d_phi0"] = calA3, r03, np.float32(0.35)
    out["t3d_L"] = Lout.copy()
    out["t3d_verts"] = verts3.reshape(n3, 162, 3)
    np.savez_compressed(os.path.join(HERE, "setup_goldens.npz"), **out)
    print("wrote", os.path.join(HERE, "setup_goldens.npz"))
    drivers(L, refdir)


def drivers(L, refdir):
    """The set-up of the reference's own driver scripts, with an unseeded
    drand48 stream (glibc's default state 0x1234ABCD330E):
      test2D.py:9-21  [Cell2D(0,0,1.2,25,1.0), Cell2D(0,0,1.2,22,1.3)] * 200,
                      Tissue2D(cells, 0.9), Disperse();
      test3D.py:8-15  [Cell3D([0,0,0],1.0,1.0)] * 32, Tissue3D(cells, 0.35),
                      Disperse2D().
    Stored: box lengths and the dispersed vertices."""
    out = {}
    n2 = 400
    calA = np.full(n2, 1.2, np.float32)
    nv = np.array([25, 22] * (n2 // 2), np.uint32)
    r0 = np.array([1.0, 1.3] * (n2 // 2), np.float32)
    Lout = np.zeros(1, np.float32)
    verts = np.zeros(int(nv.sum()) * 2, np.float32)
    L.refshim_seed48(ctypes.c_ushort(0x330E), ctypes.c_ushort(0xABCD), ctypes.c_ushort(0x1234))
    rc = L.refshim_tissue2d_disperse(n2, p(calA), p(nv), p(r0), ctypes.c_float(0.9), refdir, p(Lout), p(verts))
    assert rc == 0, rc
    out["d2_L"] = Lout.copy()
    out["d2_verts"] = verts.reshape(-1, 2)
    n3 = 32
    calA3 = np.ones(n3, np.float32)
    r03 = np.ones(n3, np.float32)
    verts3 = np.zeros(n3 * 162 * 3, np.float32)
    L.refshim_seed48(ctypes.c_ushort(0x330E), ctypes.c_ushort(0xABCD), ctypes.c_ushort(0x1234))
    rc = L.refshim_tissue3d_disperse2d(n3, p(calA3), p(r03), ctypes.c_float(0.35), refdir, p(Lout), p(verts3))
    assert rc == 0, rc
    out["d3_L"] = Lout.copy()
    out["d3_verts"] = verts3.reshape(n3, 162, 3)
    np.savez_compressed(os.path.join(HERE, "driver_goldens.npz"), **out)
    print("wrote", os.path.join(HERE, "driver_goldens.npz"))


if __name__ == "__main__":
    sys.exit(main())
