"""The 2D relaxation kernel (double tissues, FIRE, jamming) against the oracle.

Bars:
  * forces of a double tissue vs the FP64 restatement at identical inputs
    (parameters float-rounded on both sides, as stored): 1e-13 relative,
    for the reference terms and the conservative terms; a float tissue's
    conservative forces within 1e-5;
  * double Euler trajectories: 100 steps within 1e-12;
  * FIRE iterates (double): 300 iterations within 1e-9 of the restatement's
    (the only difference is the order of the double power/norm sums);
    converged runs reach max |F| < 1e-10 at the same minimum energy (1e-8);
  * jamming (double): the same bracket [phi_lo, phi_hi] and stage count as
    the restatement;
  * batches, graph replays and repeated runs: bit-identical.
"""
import numpy as np
import pytest

import oracle
from opencl_dpm_amd import capi
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu


def f32(v):
    a = np.asarray(v, np.float32)
    return float(a) if a.ndim == 0 else a


def handle(ps, mask=capi.DPM2D_CONSERVATIVE, precision=64, threads=None):
    ps = ps if isinstance(ps, list) else [ps]
    cb = np.concatenate([[0], np.cumsum([p.C for p in ps])])
    cat = lambda f: np.concatenate([np.asarray(f(p)) for p in ps])
    t = capi.Tissues2D(cb, cat(lambda p: p.nv), cat(lambda p: p.Ka), cat(lambda p: p.Kl), cat(lambda p: p.Kb),
                       cat(lambda p: p.a0), cat(lambda p: p.l0), cat(lambda p: p.r0), [p.L for p in ps],
                       Kre=[p.Kre for p in ps], Kat=[p.Kat for p in ps], mask=mask, precision=precision)
    x = cat(lambda p: p.X)
    t.set_positions(x.astype(np.float64) if precision == 64 else x)
    if threads:
        t.set_relax_threads(threads)
    return t


def padded(p, X):
    return W.Packing2D(p.nv, np.asarray(X), p.a0, p.l0, p.r0, p.Ka, p.Kl, p.Kb, p.L).padded()


def unpad(p, P, stride):
    keep = np.concatenate([np.arange(c * stride, c * stride + p.nv[c]) for c in range(p.C)])
    return P[keep]


def ref_forces(p, X, mask, L=None):
    P, stride = padded(p, np.asarray(X, np.float64))
    F, _, _ = oracle.forces2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, f32(p.L if L is None else L),
                              f32(p.Kre), f32(p.Kat), mask=mask, prec=64)
    return unpad(p, F, stride)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


PACKS = {
    "dense_ref": dict(ncells=8, nv=24, phi=0.85, Kb=0.01, Kat=0.5, seed=2),
    "uneven": dict(ncells=10, nv=0, nv_range=(16, 40), phi=0.7, Kb=0.01, seed=3, calA=1.02),
}


@pytest.mark.parametrize("case", list(PACKS))
@pytest.mark.parametrize("mask", [capi.DPM2D_ALL, capi.DPM2D_CONSERVATIVE, capi.DPM2D_ALL_TERMS])
def test_forces_vs_ref64(case, mask):
    p = W.tissue2d(**PACKS[case])
    t64 = handle(p, mask)
    t64.energy()  # forces at the current positions
    F64 = t64.forces()
    r = ref_forces(p, p.X, mask)
    e = rel(F64, r)
    print(f"2d f64 {case} mask={mask}: rel={e:.3e}", flush=True)
    assert e <= 1e-13
    if mask == capi.DPM2D_CONSERVATIVE:
        t32 = handle(p, mask, precision=32)
        t32.energy()
        e32 = rel(t32.forces(), r)
        print(f"2d f32 {case} conservative: rel={e32:.3e}", flush=True)
        assert e32 <= 1e-5


def test_energy_vs_oracle():
    p = W.tissue2d(**PACKS["uneven"])
    t = handle(p)
    st = t.energy()
    P, stride = padded(p, p.X)
    e, ei = oracle.energy2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, f32(p.L), f32(p.Kre))
    print(f"2d energy {st['energy'][0]:.12e} vs {e:.12e}; e_int {st['e_int'][0]:.6e} vs {ei:.6e}", flush=True)
    assert ei > 0
    assert abs(st["energy"][0] - e) <= 1e-12 * e
    assert abs(st["e_int"][0] - ei) <= 1e-11 * ei


def test_f64_euler_vs_ref64():
    p = W.tissue2d(**PACKS["dense_ref"])
    t = handle(p, capi.DPM2D_ALL)
    t.euler(100, 0.005)
    P, stride = padded(p, p.X.astype(np.float64))
    V, _ = oracle.run2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, f32(p.L), f32(p.Kre), f32(p.Kat),
                        100, f32(0.005), prec=64)
    err = float(np.abs(t.positions() - unpad(p, V, stride)).max())
    print(f"2d f64 100 Euler steps: max|dx|={err:.3e}", flush=True)
    assert err <= 1e-12


SMALL = dict(ncells=4, nv=16, phi=0.4, seed=11, calA=1.02, Kb=0.01)


def test_fire_iterates_vs_oracle():
    p = W.tissue2d(**SMALL)
    t = handle(p)
    st = t.fire(max_iter=300, ftol=0.0)
    assert st["iterations"][0] == 300
    P, stride = padded(p, p.X)
    V, _, so = oracle.fire2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, f32(p.L), f32(p.Kre), 0.0,
                             mask=capi.DPM2D_CONSERVATIVE, max_iter=300, ftol=0.0)
    err = float(np.abs(t.positions() - unpad(p, V, stride)).max())
    print(f"FIRE 300 iterations: max|dx|={err:.3e}", flush=True)
    assert err <= 1e-9


def test_fire_converges_to_oracle_minimum():
    p = W.tissue2d(**SMALL)
    t = handle(p)
    st = t.fire(max_iter=50000)
    P, stride = padded(p, p.X)
    V, _, so = oracle.fire2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, f32(p.L), f32(p.Kre), 0.0,
                             mask=capi.DPM2D_CONSERVATIVE, max_iter=50000)
    print(f"FIRE to 1e-10: gpu {int(st['iterations'][0])} iterations fmax {st['fmax'][0]:.2e}; "
          f"oracle {int(so['iterations'])} fmax {so['fmax']:.2e}", flush=True)
    assert st["converged"][0] == 1 and st["fmax"][0] < 1e-10 and so["converged"] == 1
    assert np.abs(t.forces()).max() < 1e-10
    # cells out of contact keep rigid-body zero modes, so the minimiser is a
    # set: after ~2e4 iterations the two summation orders end at different
    # points of it, at the same energy
    e, _ = oracle.energy2d(V, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, f32(p.L), f32(p.Kre))
    print(f"minimum energy: gpu {st['energy'][0]:.12e} oracle {e:.12e}", flush=True)
    assert abs(st["energy"][0] - e) <= 1e-8 * e


def test_jam_vs_oracle():
    p = W.tissue2d(**SMALL)
    t = handle(p)
    kw = dict(phi0=0.4, dphi=0.05, phi_tol=0.01)
    st = t.jam(fire=dict(max_iter=20000), **kw)
    P, stride = padded(p, p.X)
    V, L, so = oracle.jam2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, f32(p.L), f32(p.Kre), 0.0,
                            fire=dict(max_iter=20000), **kw)
    print(f"jamming: gpu phi in [{st['phi_lo'][0]:.5f}, {st['phi_hi'][0]:.5f}] {int(st['stages'][0])} stages "
          f"{int(st['iterations'][0])} iterations; oracle [{so['phi_lo']:.5f}, {so['phi_hi']:.5f}] "
          f"{int(so['stages'])} stages {int(so['iterations'])} iterations", flush=True)
    assert st["status"][0] == 0 and so["status"] == 0
    assert st["phi_lo"][0] == so["phi_lo"] and st["phi_hi"][0] == so["phi_hi"]
    assert st["stages"][0] == so["stages"]
    assert abs(t.box()[0] - L) <= 1e-12 * L
    assert st["e_int"][0] / p.C <= 1e-16


def test_fire_batch_equals_individual():
    ps = W.packings2d(3, cells=(4, 10), verts=(16, 32), seed=5)
    b = handle(ps, threads=128)
    sb = b.fire(max_iter=2000)
    Xb = b.positions()
    off = 0
    for k, p in enumerate(ps):
        t = handle(p, threads=128)
        s1 = t.fire(max_iter=2000)
        np.testing.assert_array_equal(Xb[off:off + len(p.X)], t.positions())
        assert s1["iterations"][0] == sb["iterations"][k]
        off += len(p.X)


def test_fire_on_the_full_512_packing_ensemble():
    """The batching form's whole ensemble (512 packings of 8-128 cells with
    16-64 vertices, 1.39 M vertices) in one double handle: 2,000 FIRE
    iterations leave every packing finite and no higher in energy, and three
    packings (the smallest, one from the middle, the largest) give the same
    bits and iteration counts run alone as in the batch (equal workgroup
    size, so the same reduction order)."""
    ps = W.packings2d(512, seed=0)
    sizes = np.array([len(p.X) for p in ps])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    b = handle(ps, threads=512)
    e0 = b.energy()["energy"]
    sb = b.fire(max_iter=2000)
    Xb = b.positions()
    print(f"512 packings, {int(offs[-1])} vertices: iterations {int(sb['iterations'].min())}-"
          f"{int(sb['iterations'].max())}, converged {int(sb['converged'].sum())}", flush=True)
    assert len(sb["energy"]) == 512
    assert np.all(np.isfinite(Xb))
    assert np.all(sb["energy"] <= e0 * (1 + 1e-9) + 1e-12)
    for k in (int(np.argmin(sizes)), 256, int(np.argmax(sizes))):
        t = handle(ps[k], threads=512)
        s1 = t.fire(max_iter=2000)
        np.testing.assert_array_equal(Xb[offs[k]:offs[k + 1]], t.positions())
        assert s1["iterations"][0] == sb["iterations"][k]


def test_jam_graph_replay_and_autotune():
    ps = W.packings2d(4, cells=(4, 8), verts=(16, 24), seed=6)
    t = handle(ps)
    X0 = t.positions()
    best, times = t.autotune(iters=20)
    print(f"autotune: {best} threads, {times}", flush=True)
    assert best in (64, 128, 256, 512, 1024) and t.relax_threads() == best
    np.testing.assert_array_equal(t.positions(), X0)  # positions restored
    kw = dict(phi0=0.4, dphi=0.05, phi_tol=0.02, fire=dict(max_iter=5000))
    s1 = t.jam(**kw)
    X1, L1 = t.positions(), t.box()
    t.set_positions(X0)
    t.set_tissue_params(L=[p.L for p in ps], mask=capi.DPM2D_CONSERVATIVE)
    s2 = t.jam(**kw)  # replays the captured graph
    np.testing.assert_array_equal(t.positions(), X1)
    np.testing.assert_array_equal(t.box(), L1)
    for k in ("phi_lo", "phi_hi", "stages", "iterations"):
        np.testing.assert_array_equal(s1[k], s2[k])
    # the same protocol over launches of at most 2 minimisations each
    t.set_positions(X0)
    t.set_tissue_params(L=[p.L for p in ps], mask=capi.DPM2D_CONSERVATIVE)
    st = t.jam_step(True, 2, **kw)
    launches = 1
    while (st["status"] == 2).any():
        st = t.jam_step(False, 2, **kw)
        launches += 1
    print(f"stepped jamming: {launches} launches", flush=True)
    assert launches > 1
    np.testing.assert_array_equal(t.positions(), X1)
    for k in ("phi_lo", "phi_hi", "stages", "iterations", "status"):
        np.testing.assert_array_equal(s1[k], st[k])
