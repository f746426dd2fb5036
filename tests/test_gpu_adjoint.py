"""The adjoint of the 3D step (dpm3d_trajectory / dpm3d_adjoint) and its
torch.autograd wrapper.

Bars:
  * one step, double: J_x^T lambda along random directions (Richardson
    extrapolated central differences, steps 2e-6 and 1e-6) and J_theta^T
    lambda for every parameter kind (central differences, step 1e-6), of the
    FP64 restatement's forces: 1e-6 relative;
  * trajectories (loss = mean |COM_N - COM_0|), double adjoint against
    central differences of the double forward on the device with the
    parameter step 1e-6 (taken between float-representable values, as the
    device stores parameters in float): 1e-6 relative (gradient norm);
    float adjoint against the same differences: 1e-4 relative;
  * the autograd path returns the adjoint's gradients.
"""
import numpy as np
import pytest
import torch

import oracle
from opencl_dpm_amd import autograd as AG
from opencl_dpm_amd import capi
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu

MASK = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_SUBSTRATE | capi.DPM3D_CONTACT | capi.DPM3D_BENDING
NAMES = ("Kv", "Ka", "Kb", "Ks")


def f32(v):
    a = np.asarray(v, np.float32)
    return float(a) if a.ndim == 0 else a


def params(m):
    C = m.C
    b = lambda v: np.broadcast_to(np.asarray(v, np.float32), (C,)).astype(np.float64)
    return {"Kv": b(m.Kv), "Ka": b(m.Ka), "Kb": b(m.Kb), "Ks": b(m.Ks)}


def tissue(m, precision, th=None, mask=MASK):
    th = th or params(m)
    return capi.Tissues3D(m.mesh.faces, [0, m.C], th["Kv"], th["Ka"], th["Ks"], m.v0, m.a0, m.L, m.Kre, Kat=0.0,
                          Kb=th["Kb"], ks_grad=m.ks_grad, mask=mask, nv=m.mesh.nv, precision=precision)


def ref_forces(m, X, th, mask=MASK):
    return oracle.forces3d(X, m.mesh.faces, m.mesh.nv, th["Kv"], th["Ka"], th["Ks"], f32(m.v0), f32(m.a0),
                           f32(m.L), f32(m.Kre), Kat=0.0, Kb=th["Kb"], mask=mask, ks_grad=f32(m.ks_grad),
                           prec=64)["forces"]


@pytest.mark.parametrize("mask", [capi.DPM3D_VOLUME, capi.DPM3D_SURFACE, capi.DPM3D_SUBSTRATE, capi.DPM3D_CONTACT,
                                  capi.DPM3D_BENDING, MASK])
def test_step_vjp_vs_oracle_differences(mask):
    m = W.monolayer3d(3, 1.7, seed=1, ks_grad=9.0)  # compressed: contact pairs present
    th = params(m)
    # the compressed start has vertices deep inside neighbouring cells (the
    # winding number is locally constant there); lifted by 1e-3 so that no
    # vertex sits on the |z| kink of the substrate force at z = 0.  (After
    # some steps contact pushes vertices onto the neighbours' surfaces, where
    # the winding number jumps and differences are meaningless.)
    X = m.X.astype(np.float64) + [0.0, 0.0, 1e-3]
    t = tissue(m, 64, mask=mask)
    t.set_positions(X)
    dt = 0.002
    t.trajectory(1, dt, 1)
    rng = np.random.Generator(np.random.PCG64(0))
    lam = rng.normal(size=X.shape)
    dth, l0 = t.adjoint(lam)
    mu = (l0 - lam) / dt  # J_x^T lambda
    eps = 1e-6
    bad = []
    # Richardson-extrapolated central differences (steps 2e-6, 1e-6): the
    # contact term's winding numbers vary sharply near the skipped faces
    cd = lambda v, e: np.sum(lam * (ref_forces(m, X + e * v, th, mask) - ref_forces(m, X - e * v, th, mask))) / (2 * e)
    for k in range(3):
        v = rng.normal(size=X.shape)
        fd = (4 * cd(v, eps) - cd(v, 2 * eps)) / 3
        ad = float(np.sum(mu * v))
        print(f"mask {mask} J_x^T lambda . v: adjoint {ad:.10e} differences {fd:.10e}", flush=True)
        bad.append(abs(ad - fd) > 1e-6 * abs(fd) + 1e-9)
    for c in (0, 4):
        for k, name in enumerate(NAMES):
            tp = {q: a.copy() for q, a in th.items()}
            tm = {q: a.copy() for q, a in th.items()}
            h = eps * max(abs(th[name][c]), 1e-3)
            tp[name][c] += h
            tm[name][c] -= h
            fd = np.sum(lam * (ref_forces(m, X, tp, mask) - ref_forces(m, X, tm, mask))) / (2 * h)
            ad = dth[c, k] / dt
            print(f"mask {mask} cell {c} d/d{name}: adjoint {ad:.10e} differences {fd:.10e}", flush=True)
            bad.append(abs(ad - fd) > 1e-6 * abs(fd) + 1e-6)
    assert not any(bad)


def loss_and_lambda(t, m, X0):
    XN = t.positions(np.float64)
    d = XN.reshape(m.C, -1, 3).mean(1) - X0.reshape(m.C, -1, 3).mean(1)
    n = np.linalg.norm(d, axis=1)
    lam = np.repeat(d / n[:, None] / m.C / m.mesh.nv, m.mesh.nv, axis=0)
    return float(n.mean()), lam


def fd_gradient(m, X0, nsteps, dt, cells, eps=1e-6):
    """Central differences of the loss with the double forward; the step is
    taken between the float values the device stores."""
    th = params(m)
    out = {}
    for c in cells:
        for k, name in enumerate(NAMES):
            vals = []
            for sgn in (+1, -1):
                tp = {q: a.copy() for q, a in th.items()}
                tp[name][c] = float(np.float32(th[name][c] * (1 + sgn * eps)))
                vals.append(tp[name][c])
                t = tissue(m, 64, tp)
                t.set_positions(X0)
                t.euler(nsteps, dt)
                vals.append(loss_and_lambda(t, m, X0)[0])
            out[(c, k)] = (vals[1] - vals[3]) / (vals[0] - vals[2])
    return out


@pytest.mark.parametrize("precision,tol", [(64, 1e-6), (32, 1e-4), ("32c", 1e-4)])
def test_trajectory_gradient_vs_differences(precision, tol):
    # cells 2.4 apart on the graded substrate: the trajectory has no contact
    # event, so the loss is smooth in the parameters and differences apply
    m = W.monolayer3d(2, 2.4, seed=3, ks_grad=9.0)
    X0 = m.X.astype(np.float64)
    nsteps, dt, every = 60, 0.002, 20
    t = tissue(m, 64 if precision == 64 else 32)
    if precision == "32c":
        t.set_compensated(True)
    t.set_positions(X0 if precision == 64 else m.X)
    t.counters(reset=True)
    t.trajectory(nsteps, dt, every)
    assert t.counters()[2] == 0  # no contact pair along the way
    L, lam = loss_and_lambda(t, m, X0)
    dth, _ = t.adjoint(lam)
    fd = fd_gradient(m, X0, nsteps, dt, range(m.C))
    a = np.array([dth[c, k] for (c, k) in fd])
    f = np.array(list(fd.values()))
    err = float(np.linalg.norm(a - f) / np.linalg.norm(f))
    print(f"f{precision} trajectory gradient: loss {L:.6e}, relative error {err:.3e} over {len(f)} parameters",
          flush=True)
    for name, k in zip(NAMES, range(4)):
        print(f"  d/d{name}: adjoint {dth[:, k]} differences {[fd[(c, k)] for c in range(m.C)]}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("precision", [64, "32c"])
def test_adjoint_same_with_either_kernel_form(precision):
    # the adjoint recomputes the forward with the handle's kernel form; the
    # two forms differ by float rounding only, so do their gradients
    m = W.monolayer3d(3, 1.75, seed=7, ks_grad=9.0)  # contact events along the way
    X0 = m.X.astype(np.float64)
    res = []
    for kind in (1, 2):
        t = tissue(m, 64 if precision == 64 else 32)
        t.set_step_kernel(kind)
        if precision == "32c":
            t.set_compensated(True)
        t.set_positions(X0 if precision == 64 else m.X)
        t.trajectory(40, 0.002, 10)
        L, lam = loss_and_lambda(t, m, X0)
        res.append((L,) + tuple(t.adjoint(lam)))
    (L1, d1, l1), (L2, d2, l2) = res
    e = float(np.linalg.norm(d1 - d2) / np.linalg.norm(d1))
    ex = float(np.linalg.norm(l1 - l2) / np.linalg.norm(l1))
    print(f"adjoint f{precision}: kernels differ by {e:.3e} (parameters), {ex:.3e} (positions)", flush=True)
    tol = 1e-9 if precision == 64 else 1e-3
    assert abs(L1 - L2) <= tol * abs(L1)
    assert e <= tol and ex <= tol


def test_autograd_function():
    m = W.monolayer3d(2, 2.4, seed=3, ks_grad=9.0)
    t = tissue(m, 64)
    th = params(m)
    tens = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in th.items()}
    X0 = torch.tensor(m.X.astype(np.float64), requires_grad=True)
    XN = AG.trajectory(t, tens["Kv"], tens["Ka"], tens["Kb"], tens["Ks"], X0, 40, 0.002, 10)
    loss = AG.com_displacement_loss(XN, X0.detach(), m.C)
    loss.backward()
    # the same gradients from the C ABI directly
    t2 = tissue(m, 64)
    t2.set_positions(m.X.astype(np.float64))
    t2.trajectory(40, 0.002, 10)
    L, lam = loss_and_lambda(t2, m, m.X.astype(np.float64))
    dth, l0 = t2.adjoint(lam)
    assert abs(loss.item() - L) <= 1e-12 * L
    for k, name in enumerate(NAMES):
        np.testing.assert_allclose(tens[name].grad.numpy(), dth[:, k], rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(X0.grad.numpy(), l0, rtol=1e-10, atol=1e-10 * np.abs(l0).max())


@pytest.mark.timeout(600)
def test_differentiation_workload_gradients():
    # the differentiation workload itself: 64 cells x 162 vertices on the
    # substrate whose stiffness grows 10x across x, loss = mean |COM_200 -
    # COM_0|, checkpoints every 20 steps.  Reference values: FP64 forward
    # runs; per parameter the median of the difference quotients over the
    # consecutive float values theta_0 + k ulp, k = -8..8 (the force law's
    # cut-offs put isolated jumps in the loss, which a single central
    # difference can straddle).  Bar: 1e-4 relative (gradient norm) for the
    # double adjoint and for the float adjoint with compensated positions;
    # the plain float adjoint is reported (its positions drop displacements
    # below one float ulp of the coordinates, so its own loss drifts).
    m = W.monolayer_config("cfg5")
    th = params(m)
    X0 = m.X.astype(np.float64)
    nsteps, dt, every = 200, 0.002, 20
    cells = range(4)
    grads = {}
    for prec in (64, "32c", 32):
        t = tissue(m, 64 if prec == 64 else 32, th)
        if prec == "32c":
            t.set_compensated(True)
        t.set_positions(X0 if prec == 64 else m.X)
        t.trajectory(nsteps, dt, every)
        L, lam = loss_and_lambda(t, m, X0)
        grads[prec] = t.adjoint(lam)[0]
        print(f"f{prec}: loss {L:.9e}", flush=True)

    def loss_at(c, name, value):
        tp = {q: a.copy() for q, a in th.items()}
        tp[name][c] = value
        t = tissue(m, 64, tp)
        t.set_positions(X0)
        t.euler(nsteps, dt)
        return loss_and_lambda(t, m, X0)[0]

    ref = np.zeros((len(cells), 4))
    for c in cells:
        for k, name in enumerate(NAMES):
            x0 = np.float32(th[name][c])
            xs = [float(x0)]
            up = dn = x0
            for _ in range(8):
                up = np.nextafter(up, np.float32(np.inf))
                dn = np.nextafter(dn, np.float32(-np.inf))
                xs = [float(dn)] + xs + [float(up)]
            Ls = [loss_at(c, name, x) for x in xs]
            ref[c, k] = float(np.median(np.diff(Ls) / np.diff(xs)))
        print(f"cell {c}: reference {ref[c]}", flush=True)
    for prec, tol in ((64, 1e-4), ("32c", 1e-4), (32, None)):
        g = grads[prec][:len(cells)]
        e = float(np.linalg.norm(g - ref) / np.linalg.norm(ref))
        print(f"f{prec} adjoint vs FP64 1-ulp difference quotients: relative error {e:.3e}", flush=True)
        if tol is not None:
            assert e <= tol
