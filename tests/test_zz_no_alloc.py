"""dpm3d_probe_noalloc: dpm3d_probe with its device tables in the library's
static scratch instead of per-call allocations.  Its outputs must equal
dpm3d_probe's bit for bit, on consecutive calls with a smaller input the
second time (the scratch keeps the previous call's contents), and sizes
beyond the scratch must be refused before any device work."""
import numpy as np
import pytest

from opencl_dpm_amd import capi
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu


def handle(m):
    return capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat, Kb=m.Kb,
                          mask=capi.DPM3D_ALL, nv=m.mesh.nv)


def test_probe_noalloc_equals_probe_on_consecutive_calls():
    # 4 x 4 squeezed monolayer (contacts and adhesion), then a 3 x 3 one
    for n, sp, seed in ((4, 1.8, 70), (3, 1.9, 71)):
        m = W.monolayer3d(n, sp, seed=seed)
        t = handle(m)
        t.set_positions(m.X)
        t.euler(30, 0.002)
        for kernel in (1, 2):
            t.set_step_kernel(kernel)
            ref = t.probe(winding=True)
            got = t.probe_noalloc(winding=True)
            for key in ("forces", "contact", "winding", "down", "adhesion"):
                np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{n}x{n} kernel={kernel} {key}")
            print(f"{n}x{n} kernel={kernel}: contacts {int(ref['contact'].sum())}, "
                  f"adhesion {int(ref['adhesion'].sum())}, identical", flush=True)
        assert int(ref["contact"].sum()) > 0 and int(ref["adhesion"].sum()) > 0
        # the positions are untouched by either probe
        X = t.positions()
        t.probe_noalloc()
        np.testing.assert_array_equal(t.positions(), X)


def test_probe_noalloc_refuses_tables_beyond_its_scratch():
    # 24 x 24 cells: the winding table alone is 576 * 162 * 576 floats (215 MB)
    m = W.monolayer3d(24, 2.4, seed=72)
    t = handle(m)
    t.set_positions(m.X)
    with pytest.raises(capi.DPMError, match="scratch"):
        t.probe_noalloc(winding=True)
    # without the winding table (54 MB) it fits and matches the probe
    got = t.probe_noalloc()
    ref = t.probe()
    for key in ("forces", "contact", "down", "adhesion"):
        np.testing.assert_array_equal(got[key], ref[key])
