"""The drop-in Python surface and the C ABI, on the CPU (no GPU calls).

Geometry and dispersal are compared bit-for-bit with golden vectors that
the reference's own host code produced (tests/golden/make_golden.py).
"""
import ctypes
import os
import re

import numpy as np
import pytest

from opencl_dpm_amd import capi, clDPM

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_cell3d_matches_reference_constructor(goldens):
    for k, row in enumerate(goldens["cell3d_args"]):
        c = clDPM.Cell3D([float(x) for x in row[:3]], float(row[3]), float(row[4]))
        np.testing.assert_array_equal(bits(c.Verts), bits(goldens["cell3d_verts"][k]))
        np.testing.assert_array_equal(np.asarray(c.GetFaces()), goldens["cell3d_faces"])
        s = goldens["cell3d_scalars"][k]
        assert [c.v0, c.sa0, c.a0, c.Volume, c.SurfaceArea, c.calA0, c.r0] == list(s)
        assert np.asarray(c.GetPositions()).shape == (3, 162)
        assert np.asarray(c.GetForces()).shape == (3, 162)


def test_cell3d_host_reductions_match_reference(goldens):
    c = clDPM.Cell3D([0.0, 0.0, 0.0], 1.0, 1.0)
    c.Verts = goldens["measure_verts"].tolist()
    out = goldens["measure_out"]
    assert c.GetVolume() == out[0]
    assert c.GetSurfaceArea() == out[1]
    assert list(c.GetCOM()) == list(out[2:5])


def test_cell2d_matches_reference_constructor(goldens):
    k = 0
    while f"cell2d_{k}_args" in goldens:
        x0, y0, ca, nv, r0 = goldens[f"cell2d_{k}_args"]
        c = clDPM.Cell2D(float(x0), float(y0), float(ca), int(nv), float(r0))
        np.testing.assert_array_equal(bits(c.Verts), bits(goldens[f"cell2d_{k}_verts"]))
        assert [c.a0, c.l0, c.calA0, c.GetArea()] == list(goldens[f"cell2d_{k}_scalars"])
        k += 1
    assert k >= 5


def test_icosphere_counts_and_orientation():
    for s in range(4):
        c = clDPM.Cell3D([0, 0, 0], 1.0, 1.0, s)
        assert c.NV == 10 * 4 ** s + 2 and c.NF == 20 * 4 ** s
        X = np.asarray(c.Verts, np.float64)
        F = np.asarray(c.GetFaces())
        n = np.cross(X[F[:, 1]] - X[F[:, 0]], X[F[:, 2]] - X[F[:, 0]])
        cen = X[F].mean(1)
        assert np.all(np.einsum("ij,ij->i", n, cen) > 0)  # outward
        # closed manifold: every edge shared by exactly two faces
        e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)


def test_cells_are_references_and_verts_are_copies():
    c = clDPM.Cell3D([0, 0, 0], 1.0, 1.0)
    T = clDPM.Tissue3D([c] * 3, 0.35)
    T.Cells[1].Kv = 2.5            # propagates (reference_internal elements)
    assert T.Cells[1].Kv == 2.5 and T.Cells[0].Kv == 0.0
    assert c.Kv == 0.0             # the constructor copied the cells
    v = T.Cells[0].Verts
    v[0][0] = 99.0                 # a fresh list: no effect
    assert T.Cells[0].Verts[0][0] != 99.0
    T.Cells[0].Verts = v           # assignment does propagate
    assert T.Cells[0].Verts[0][0] == np.float32(99.0)


def make_tissue3d(n=2, **kw):
    cells = []
    for i in range(n):
        c = clDPM.Cell3D([2.5 * i, 0.0, 1.0], 1.0, 1.0)
        c.Kv, c.Ka, c.Ks = kw.get("Kv", 1.0), kw.get("Ka", 1.0), kw.get("Ks", 1.0)
        cells.append(c)
    return clDPM.Tissue3D(cells, 0.35)


@pytest.mark.parametrize("nsteps,dt", [(0, 0.01), (-3, 0.01), (5, 0.0), (5, -0.1), (5, 0.2)])
def test_3d_argument_validation_raises_value_error(nsteps, dt):
    T = make_tissue3d()
    with pytest.raises(ValueError):
        T.CLEulerUpdate(nsteps, dt)


@pytest.mark.parametrize("field", ["Kv", "Ka", "Ks"])
def test_3d_nonpositive_constants_raise_runtime_error(field):
    T = make_tissue3d(**{field: 0.0})
    with pytest.raises(RuntimeError, match="Invalid spring constants"):
        T.CLEulerUpdate(5, 0.01)


def test_3d_nonfinite_input_raises_runtime_error():
    T = make_tissue3d()
    v = T.Cells[1].Verts
    v[3][2] = float("nan")
    T.Cells[1].Verts = v
    with pytest.raises(RuntimeError, match="Non-finite vertex"):
        T.CLEulerUpdate(5, 0.01)


def test_2d_zero_steps_returns_zero_forces():
    c = clDPM.Cell2D(0, 0, 1.0, 16, 1.0)
    c.Forces = [[1.0, 2.0]] * 16
    T = clDPM.Tissue2D([c], 0.8)
    before = T.Cells[0].Verts
    T.CLEulerUpdate(0, 0.005)
    assert T.Cells[0].Verts == before
    assert np.all(np.asarray(T.Cells[0].Forces) == 0)


def header_functions():
    src = open(os.path.join(ROOT, "include", "dpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dpm\w+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    lib = capi.load()
    names = header_functions()
    assert len(names) >= 20
    assert sorted(capi.EXPORTS) == names
    for n in names:
        assert hasattr(lib, n), n


def test_library_fails_loudly_without_device():
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    mesh_faces = np.asarray(clDPM.Cell3D([0, 0, 0], 1.0, 1.0, 1).GetFaces())
    with pytest.raises(capi.DPMError):
        capi.Tissues3D(mesh_faces, [0, 1], 1, 1, 1, 4.0, 0.1, 10.0, 1.0)
    T = make_tissue3d()
    with pytest.raises(RuntimeError):
        T.CLEulerUpdate(2, 0.01)


def test_short_verts_lists_raise_value_error_before_any_device_call():
    # Verts is read-write from Python; the drop-in reads NV entries, so a
    # shorter list is an error (the reference would read past its end)
    T = clDPM.Tissue2D([clDPM.Cell2D(0, 0, 1.0, 16, 1.0), clDPM.Cell2D(3, 0, 1.0, 16, 1.0)], 0.8)
    T.Cells[1].Verts = T.Cells[1].Verts[:10]
    with pytest.raises(ValueError, match="fewer than NV"):
        T.CLEulerUpdate(3, 0.005)
    with pytest.raises(ValueError, match="fewer than NV"):
        T.Disperse()
    c = clDPM.Cell2D(0, 0, 1.0, 16, 1.0)
    c.Verts = c.Verts[:10]
    with pytest.raises(ValueError, match="fewer than NV"):
        c.GetArea()
    with pytest.raises(ValueError, match="fewer than NV"):
        clDPM.Tissue2D([c], 0.8)
    c3 = clDPM.Cell3D([0, 0, 0], 1.0, 1.0)
    c3.Verts = c3.Verts[:100]
    for f in (c3.GetPositions, c3.GetVolume, c3.GetSurfaceArea, c3.GetCOM, lambda: c3.GetVesselPositions(10.0)):
        with pytest.raises(ValueError, match="fewer than NV"):
            f()
    c3 = clDPM.Cell3D([0, 0, 0], 1.0, 1.0)
    c3.Forces = c3.Forces[:5]
    with pytest.raises(ValueError, match="fewer than NV"):
        c3.GetForces()
    T3 = clDPM.Tissue3D([clDPM.Cell3D([0, 0, 0], 1.0, 1.0)], 0.35)
    T3.Cells = [T3.Cells[0], clDPM.Cell3D([3, 0, 0], 1.0, 1.0)]
    v = T3.Cells[1].Verts
    T3.Cells[1].Verts = v[:50]
    with pytest.raises(ValueError, match="fewer than NV"):
        T3.Disperse2D()


def test_longer_verts_list_is_packed_to_nv():
    # a longer Verts list is packed to its first NV entries, as the reference
    # packs (Tissue2D.cpp:137-139): without a device the call fails in the
    # library, after packing, with the library's error
    c = clDPM.Cell2D(0, 0, 1.0, 16, 1.0)
    c.Verts = c.Verts + [[5.0, 5.0]] * 4
    T = clDPM.Tissue2D([c], 0.8)
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible (covered by the GPU drop-in tests)")
    with pytest.raises(RuntimeError):
        T.CLEulerUpdate(3, 0.005)
    assert len(T.Cells[0].Verts) == 20


def test_box_lengths_match_reference_host_code(goldens):
    # Tissue2D.L = sqrt(sum a0) / phi0, Tissue3D.L = cbrt(sum v0) / phi0
    # (Tissue2D.cpp:31, Tissue3D.cpp:28), as the reference's own host code
    # computed them for the golden tissues and the driver scripts
    g = np.load(os.path.join(ROOT, "tests", "golden", "driver_goldens.npz"))
    cells = [clDPM.Cell2D(0, 0, float(a), int(n), float(r))
             for a, n, r in zip(goldens["t2d_calA"], goldens["t2d_nv"], goldens["t2d_r0"])]
    T = clDPM.Tissue2D(cells, float(goldens["t2d_phi0"]))
    assert T.L == goldens["t2d_L"][0]
    assert T.NCELLS == len(cells) and T.PBC is True and T.Kre == 1.0 and T.Kat == 0.0
    cells = [clDPM.Cell3D([0, 0, 0], float(a), float(r)) for a, r in zip(goldens["t3d_calA"], goldens["t3d_r0"])]
    assert clDPM.Tissue3D(cells, float(goldens["t3d_phi0"])).L == goldens["t3d_L"][0]
    T2 = clDPM.Tissue2D([clDPM.Cell2D(0, 0, 1.2, 25, 1.0), clDPM.Cell2D(0, 0, 1.2, 22, 1.3)] * 200, 0.9)
    assert T2.L == g["d2_L"][0]
    assert clDPM.Tissue3D([clDPM.Cell3D([0.0] * 3, 1.0, 1.0)] * 32, 0.35).L == g["d3_L"][0]


def test_dispersal_fails_loudly_without_device():
    # the dispersal relaxation runs on the device (dpm_disperse); without one
    # it raises instead of falling back to the host
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible (the dispersal goldens are checked by the GPU tests)")
    T = clDPM.Tissue2D([clDPM.Cell2D(0, 0, 1.0, 16, 1.0)] * 4, 0.8)
    with pytest.raises(RuntimeError):
        T.Disperse()
    T3 = clDPM.Tissue3D([clDPM.Cell3D([0.0] * 3, 1.0, 1.0)] * 4, 0.35)
    with pytest.raises(RuntimeError):
        T3.Disperse2D()
