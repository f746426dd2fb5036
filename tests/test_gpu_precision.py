"""The all-double 3D step (dpm3d_create_f64) and the energy diagnostic.

Bars:
  * FP64 handle vs the FP64 restatement (ref64) from identical inputs
    (parameters float-rounded on both sides, as the device stores them):
    forces within 1e-13 relative (per vertex 1e-13 of the largest), pair
    sets bit-exact against ref64's decisions;
  * FP64 trajectories: 40 steps within 1e-12 of ref64 (both double; only the
    summation order differs);
  * energy (both handles): within 1e-13 relative of the oracle's double
    energy at the handle's positions;
  * the float production step against the FP64 step at identical positions,
    at the full 4096-cell x 642-vertex precision workload: <= 1e-5 relative
    force error (the production bar), and the float step's energy stays
    within a stated band of the FP64 run's over a trajectory.
"""
import numpy as np
import pytest

import oracle
from opencl_dpm_amd import capi
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu

F64_RTOL = 1e-13
F64_VTOL = 1e-13


def tissue3d(m, mask=capi.DPM3D_ALL, precision=64):
    return capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat,
                          Kb=m.Kb, ks_grad=m.ks_grad, mask=mask, nv=m.mesh.nv, precision=precision)


def f32(v):
    # the device holds every parameter as float: the double restatement is
    # given the same (float-rounded) values
    a = np.asarray(v, np.float32)
    return float(a) if a.ndim == 0 else a


def ref3d(m, X, mask, prec=64, pairs=False):
    return oracle.forces3d(X, m.mesh.faces, m.mesh.nv, f32(m.Kv), f32(m.Ka), f32(m.Ks), f32(m.v0), f32(m.a0),
                           f32(m.L), f32(m.Kre), Kat=f32(m.Kat), Kb=f32(m.Kb), mask=mask,
                           ks_grad=f32(m.ks_grad), prec=prec, pairs=pairs)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


CASES = {
    "grid3x3": (3, 2.4, 2, 0),
    "compressed3x3": (3, 1.7, 2, 1),
    "subdiv3_2x2": (2, 1.8, 3, 3),
}


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("mask", [capi.DPM3D_REFERENCE, capi.DPM3D_ALL])
def test_f64_forces_and_pairs_vs_ref64(case, mask, kernel):
    # kernel 1: one workgroup per cell; 2: one wave per cell (<= 256 vertices)
    n, sp, sd, seed = CASES[case]
    m = W.monolayer3d(n, sp, subdiv=sd, seed=seed)
    t = tissue3d(m, mask)
    if kernel == 2 and m.mesh.nv > 256:
        pytest.skip("the wave-per-cell step takes meshes of at most 256 vertices")
    t.set_step_kernel(kernel)
    X = m.X.astype(np.float64)
    t.set_positions(X)
    np.testing.assert_array_equal(t.positions(), X)
    g = t.probe()
    Fg = t.forces()
    assert Fg.dtype == np.float64
    r = ref3d(m, X, mask, pairs=True)
    e = rel(Fg, r["forces"])
    vmax = float(np.abs(Fg - r["forces"]).max())
    print(f"f64 {case} mask={mask} kernel={kernel}: rel={e:.3e} vmax={vmax:.3e}", flush=True)
    assert e <= F64_RTOL
    assert vmax <= F64_VTOL * float(np.abs(r["forces"]).max())
    np.testing.assert_array_equal(g["contact"], r["contact"])
    np.testing.assert_array_equal(g["down"], r["down"])
    np.testing.assert_array_equal(g["adhesion"], r["adhesion"])


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("mask", [capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_BENDING,
                                  capi.DPM3D_ALL])
def test_f64_trajectory_vs_ref64(mask, kernel):
    m = W.monolayer3d(3, 1.7, seed=4)
    t = tissue3d(m, mask)
    t.set_step_kernel(kernel)
    t.set_positions(m.X.astype(np.float64))
    nsteps, dt = 40, 0.002
    t.euler(nsteps, dt)
    X64, _ = oracle.run3d(m.X, m.mesh.faces, m.mesh.nv, f32(m.Kv), f32(m.Ka), f32(m.Ks), f32(m.v0), f32(m.a0),
                          f32(m.L), f32(m.Kre), nsteps, f32(dt), Kat=f32(m.Kat), Kb=f32(m.Kb), mask=mask, prec=64)
    err = float(np.abs(t.positions() - X64).max())
    print(f"f64 40-step trajectory mask={mask} kernel={kernel}: max|dx|={err:.3e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("precision", [32, 64])
def test_energy_vs_oracle(precision):
    m = W.monolayer3d(3, 1.7, seed=5)
    t = tissue3d(m, capi.DPM3D_ALL, precision)
    t.set_positions(m.X)
    t.euler(30, 0.002)  # away from the rest shape
    X = t.positions(np.float64)
    Eg = t.energy()
    Eo = oracle.energy3d(X, m.mesh.faces, m.mesh.nv, f32(m.Kv), f32(m.Ka), f32(m.v0), f32(m.a0), Kb=f32(m.Kb),
                         mask=35)
    print(f"energy f{precision}: rel={np.abs(Eg - Eo).max() / np.abs(Eo).max():.3e}", flush=True)
    np.testing.assert_allclose(Eg, Eo, rtol=1e-13)
    # the mask selects the terms
    t.set_tissue_params(mask=capi.DPM3D_BENDING)
    Eb = oracle.energy3d(X, m.mesh.faces, m.mesh.nv, f32(m.Kv), f32(m.Ka), f32(m.v0), f32(m.a0), Kb=f32(m.Kb),
                         mask=32)
    np.testing.assert_allclose(t.energy(), Eb, rtol=1e-13)


def test_f32_step_vs_f64_step_full_precision_workload():
    # 4096 cells x 642 vertices (icosphere s=3): float forces with double
    # per-cell sums against the all-double step, at identical positions
    m = W.monolayer_config("cfg3")
    t32 = tissue3d(m, capi.DPM3D_ALL, 32)
    t64 = tissue3d(m, capi.DPM3D_ALL, 64)
    t32.set_positions(m.X)
    t32.euler(50, 0.002)
    X = t32.positions()
    t64.set_positions(X.astype(np.float64))
    # one more step of each: the forces it reports are those at X
    t32.euler(1, 0.002)
    t64.euler(1, 0.002)
    F32, F64 = t32.forces(), t64.forces()
    e = rel(F32, F64)
    vmax = float(np.abs(F32 - F64).max())
    print(f"cfg3 force error f32 vs f64: rel={e:.3e} vmax={vmax:.3e} scale={np.abs(F64).max():.3e}", flush=True)
    assert e <= 1e-5
    assert vmax <= 1e-4 * float(np.abs(F64).max())


def test_energy_drift_f32_vs_f64():
    # the same start advanced by both steps; the elastic energy of the float
    # run follows the double run's
    m = W.monolayer3d(4, 2.0, seed=6)
    t32 = tissue3d(m, capi.DPM3D_ALL, 32)
    t64 = tissue3d(m, capi.DPM3D_ALL, 64)
    t32.set_positions(m.X)
    t64.set_positions(m.X.astype(np.float64))
    E0 = t64.energy().sum()
    worst = 0.0
    for _ in range(10):
        t32.euler(100, 0.002)
        t64.euler(100, 0.002)
        e32, e64 = t32.energy().sum(), t64.energy().sum()
        worst = max(worst, abs(e32 - e64) / abs(E0))
    print(f"1000-step energy drift |E32 - E64|/E0 <= {worst:.3e} (E0={E0:.4e})", flush=True)
    assert worst <= 1e-4


def test_compensated_positions_track_the_double_run():
    # smooth terms only: the float run with compensated positions stays far
    # closer to the double run than the plain float run, and its update is
    # the plain one where no rounding occurs (same forces from the same x)
    m = W.monolayer3d(4, 2.0, seed=7)
    mask = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_BENDING
    t32, t32c, t64 = tissue3d(m, mask, 32), tissue3d(m, mask, 32), tissue3d(m, mask, 64)
    t32c.set_compensated(True)
    t32.set_positions(m.X)
    t32c.set_positions(m.X)
    t64.set_positions(m.X.astype(np.float64))
    for t in (t32, t32c, t64):
        t.euler(1000, 0.002)
    X64 = t64.positions()
    e_plain = float(np.abs(t32.positions(np.float64) - X64).max())
    e_comp = float(np.abs(t32c.positions(np.float64) - X64).max())
    E64 = t64.energy().sum()
    d_plain = abs(t32.energy().sum() - E64) / E64
    d_comp = abs(t32c.energy().sum() - E64) / E64
    print(f"1000 steps: max|x - x64| plain {e_plain:.3e} compensated {e_comp:.3e}; energy drift plain "
          f"{d_plain:.3e} compensated {d_comp:.3e}", flush=True)
    assert e_comp < 0.2 * e_plain
    assert d_comp < d_plain
    # the float accessor returns the float part, the double one x + residual
    np.testing.assert_array_equal(t32c.positions(), t32c.positions(np.float64).astype(np.float32))
