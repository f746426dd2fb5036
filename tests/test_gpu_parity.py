"""Parity of the HIP timestep with the CPU oracle, through the C ABI.

Bars (stated per test):
  * forces: ||F_gpu - F_ref64|| / ||F_ref64|| <= 1e-5 and every vertex within
    1e-4 * max|F_ref64|, from identical float inputs;
  * pair sets (3D contact vertex-cell pairs, facing-down faces, adhesion
    partners; 2D overlap and attraction partners): bit-exact against the
    float restatement ref32, whose decisions use the same IEEE operation
    order as the kernels (partner sets as counts plus per-vertex key sums,
    dpm3d_probe_adhesion / dpm2d_probe_attraction);
  * batching, graph replay and repeated runs: bit-identical to the plain path.
"""
import numpy as np
import pytest

import oracle
from opencl_dpm_amd import capi, clDPM
from opencl_dpm_amd import workloads as W

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
FORCE_VTOL = 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def assert_forces(Fg, F64, what):
    r = rel(Fg, F64)
    vmax = float(np.abs(np.asarray(Fg, np.float64) - F64).max())
    scale = float(np.abs(F64).max())
    print(f"{what}: rel={r:.3e} vmax={vmax:.3e} scale={scale:.3e}", flush=True)
    assert r <= FORCE_RTOL, (what, r)
    assert vmax <= FORCE_VTOL * scale, (what, vmax, scale)


def fmix32(h):
    h = np.asarray(h, np.uint64) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h


def contact_signature(table):
    """Per vertex the contact partner count and the 32-bit key sum of the
    partner cells (the form dpm3d_probe_contact reports) from the table."""
    keys = fmix32(np.arange(table.shape[1], dtype=np.uint64) + np.uint64(0x9E3779B9))
    on = table.astype(bool)
    return on.sum(1).astype(np.uint32), ((on * keys[None, :]).sum(1) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def tissue3d(m, mask=capi.DPM3D_ALL, Kat=None, Kre=None):
    return capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L,
                          m.Kre if Kre is None else Kre, Kat=m.Kat if Kat is None else Kat,
                          Kb=m.Kb, ks_grad=m.ks_grad, mask=mask, nv=m.mesh.nv)


def ref3d(m, X, mask, prec, pairs=False):
    return oracle.forces3d(X, m.mesh.faces, m.mesh.nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat,
                           Kb=m.Kb, mask=mask, ks_grad=m.ks_grad, prec=prec, pairs=pairs)


CASES_3D = {
    # name: (n_side, spacing, subdiv, seed)
    "grid3x3": (3, 2.4, 2, 0),
    "compressed3x3": (3, 1.7, 2, 1),
    "compressed4x4": (4, 1.8, 2, 2),
    "subdiv3_2x2": (2, 1.8, 3, 3),
    "subdiv1_3x3": (3, 1.6, 1, 5),   # 42 vertices: fewer than a wave's lanes
    "subdiv0_3x3": (3, 1.6, 0, 6),   # 12 vertices (icosahedra)
}


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("case", list(CASES_3D))
@pytest.mark.parametrize("mask", [capi.DPM3D_REFERENCE, capi.DPM3D_ALL])
def test_3d_forces_and_pairs(case, mask, kernel):
    # kernel 1: one workgroup per cell; 2: one wave per cell (<= 256 vertices)
    n, sp, sd, seed = CASES_3D[case]
    m = W.monolayer3d(n, sp, subdiv=sd, seed=seed)
    t = tissue3d(m, mask)
    if kernel == 2 and m.mesh.nv > 256:
        with pytest.raises(ValueError):
            t.set_step_kernel(2)
        return
    t.set_step_kernel(kernel)
    t.set_positions(m.X)
    g = t.probe()
    r64 = ref3d(m, m.X, mask, 64)
    r32 = ref3d(m, m.X, mask, 32, pairs=True)
    assert_forces(g["forces"], r64["forces"], f"3d {case} mask={mask} kernel={kernel}")
    np.testing.assert_array_equal(g["contact"], r32["contact"])
    np.testing.assert_array_equal(g["down"], r32["down"])
    np.testing.assert_array_equal(g["adhesion"], r32["adhesion"])
    np.testing.assert_array_equal(g["adhesion_hash"], r32["adhesion_hash"])
    n, k = t.probe_contact()  # the compact form of the same contact set
    n32, k32 = contact_signature(r32["contact"])
    np.testing.assert_array_equal(n, n32)
    np.testing.assert_array_equal(k, k32)
    print(f"pairs: contact={int(g['contact'].sum())} down={int(g['down'].sum())} "
          f"adhesion={int(g['adhesion'].sum())}", flush=True)
    if case.startswith("compressed") and mask & capi.DPM3D_CONTACT:
        assert g["contact"].sum() > 0


@pytest.mark.parametrize("kernel", [1, 2])
def test_3d_pair_sets_across_the_separation_cutoffs(kernel):
    # neighbours from overlapping to well apart (periodic 3 x 3 monolayer,
    # spacing 1.9 ... 2.7): the centre-line separation crosses the contact
    # culling margin (~0.4 e_max) and the adhesion reach 2 l0 on the way; the
    # pairs it lets the step skip must be exactly the ones ref32 rejects
    for k, sp in enumerate(np.linspace(1.9, 2.7, 17)):
        m = W.monolayer3d(3, float(sp), seed=40 + k)
        t = tissue3d(m)
        t.set_step_kernel(kernel)
        t.set_positions(m.X)
        g = t.probe()
        r32 = ref3d(m, m.X, capi.DPM3D_ALL, 32, pairs=True)
        assert_forces(g["forces"], ref3d(m, m.X, capi.DPM3D_ALL, 64)["forces"], f"3d spacing={sp:.2f} kernel={kernel}")
        np.testing.assert_array_equal(g["contact"], r32["contact"])
        np.testing.assert_array_equal(g["adhesion"], r32["adhesion"])
        np.testing.assert_array_equal(g["adhesion_hash"], r32["adhesion_hash"])
        print(f"spacing {sp:.2f}: contact={int(g['contact'].sum())} adhesion={int(g['adhesion'].sum())}", flush=True)


@pytest.mark.parametrize("spacing,seed", [(1.7, 21), (1.6, 22)])
def test_3d_pairs_after_deformation(spacing, seed):
    # cells squeezed against each other for 300 steps: deformed meshes with
    # stretched faces, then every pair set vs ref32 at those positions
    m = W.monolayer3d(3, spacing, seed=seed)
    t = tissue3d(m)
    t.set_positions(m.X)
    t.euler(300, 0.002)
    X = t.positions()
    g = t.probe()
    r32 = ref3d(m, X, capi.DPM3D_ALL, 32, pairs=True)
    assert_forces(g["forces"], ref3d(m, X, capi.DPM3D_ALL, 64)["forces"], f"3d deformed spacing={spacing}")
    np.testing.assert_array_equal(g["contact"], r32["contact"])
    np.testing.assert_array_equal(g["down"], r32["down"])
    np.testing.assert_array_equal(g["adhesion"], r32["adhesion"])
    np.testing.assert_array_equal(g["adhesion_hash"], r32["adhesion_hash"])
    print(f"deformed pairs: contact={int(g['contact'].sum())}", flush=True)
    assert g["contact"].sum() > 0


@pytest.mark.parametrize("mask", [capi.DPM3D_REFERENCE, capi.DPM3D_ALL])
def test_3d_wave_and_workgroup_kernels_agree(mask):
    # the two kernel forms of the step: one evaluation on squeezed, deformed
    # cells (equal pair sets, forces to float rounding: the wave form groups
    # the edge-spring and adhesion sums differently), then a 60-step
    # trajectory from the same start
    m = W.monolayer3d(4, 1.8, seed=2)
    t = tissue3d(m, mask)
    t.set_positions(m.X)
    t.euler(100, 0.002)
    X = t.positions()
    probes = []
    for kind in (1, 2):
        t.set_step_kernel(kind)
        t.set_positions(X)
        probes.append(t.probe())
    g1, g2 = probes
    r = rel(g2["forces"], g1["forces"].astype(np.float64))
    print(f"kernels: mask={mask} force rel {r:.3e} contacts {int(g1['contact'].sum())} "
          f"adhesion {int(g1['adhesion'].sum())}", flush=True)
    assert r <= 1e-6
    for key in ("contact", "down", "adhesion", "adhesion_hash"):
        np.testing.assert_array_equal(g1[key], g2[key])
    out = []
    for kind in (1, 2):
        t.set_step_kernel(kind)
        t.set_positions(m.X)
        t.euler(60, 0.002)
        out.append(t.positions())
    d = float(np.abs(out[0] - out[1]).max())
    print(f"kernels: mask={mask} 60-step position diff {d:.3e}", flush=True)
    assert d <= 1e-5 * float(np.abs(out[0]).max())


def test_3d_wave_kernel_degenerate_edge():
    # a zero-length edge: the reference skips the edge springs of its faces;
    # the wave form takes its per-corner path for that cell, whose arithmetic
    # is the workgroup form's face pass
    m = W.monolayer3d(2, 2.4, seed=4)
    nv = m.mesh.nv
    X = m.X.copy()
    u = int(m.mesh.faces[0, 1])
    X[u] = X[int(m.mesh.faces[0, 0])]  # cell 0: vertex u on vertex faces[0,0]
    shape = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_SUBSTRATE | capi.DPM3D_BENDING
    t = tissue3d(m, shape)
    F = []
    for kind in (1, 2):
        t.set_step_kernel(kind)
        t.set_positions(X)
        F.append(t.probe(contact=False)["forces"])
    np.testing.assert_array_equal(F[0][:nv], F[1][:nv])
    assert_forces(F[1], ref3d(m, X, shape, 64)["forces"], "3d degenerate edge, wave kernel")


def test_3d_substrate_gradient_forces():
    m = W.monolayer_config("cfg5")
    t = tissue3d(m)
    t.set_positions(m.X)
    assert_forces(t.probe()["forces"], ref3d(m, m.X, capi.DPM3D_ALL, 64)["forces"], "3d cfg5 ks_grad=9")


def traj_bound(X0, X64, nsteps):
    # float positions: one rounding of x + F dt per step (random walk of
    # half-ulps) plus the 1e-5 force bar integrated over the displacement
    return 1e-6 * float(np.abs(X64).max()) * np.sqrt(nsteps) + 1e-5 * float(np.abs(X64 - X0).max())


SMOOTH3D = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_BENDING


def test_3d_trajectory_smooth_terms_vs_ref64():
    m = W.monolayer3d(3, 1.7, seed=4)
    t = tissue3d(m, SMOOTH3D)
    t.set_positions(m.X)
    nsteps, dt = 40, 0.002
    t.euler(nsteps, dt)
    X64, F64 = oracle.run3d(m.X, m.mesh.faces, m.mesh.nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, nsteps, dt,
                            Kat=m.Kat, Kb=m.Kb, mask=SMOOTH3D, prec=64)
    err = float(np.abs(t.positions() - X64).max())
    print(f"40-step smooth-term positions: max|dx|={err:.3e} bound {traj_bound(m.X, X64, nsteps):.3e}", flush=True)
    assert err <= traj_bound(m.X, X64, nsteps)
    # the next step's forces, against ref64 at the same (GPU) positions
    Xn = t.positions()
    t.euler(1, dt)
    assert_forces(t.forces(), ref3d(m, Xn, SMOOTH3D, 64)["forces"], "3d step-41 forces (smooth terms)")


def test_3d_trajectory_all_terms_within_reference_float_drift():
    # With the cut-offs of contact, adhesion and substrate, a trajectory
    # in float departs from the exact one wherever a borderline decision
    # flips; the yardstick is the reference arithmetic's own departure
    # (ref32 vs ref64) over the same steps.
    m = W.monolayer3d(3, 1.7, seed=4)
    t = tissue3d(m)
    t.set_positions(m.X)
    nsteps, dt = 40, 0.002
    t.euler(nsteps, dt)
    kw = dict(Kat=m.Kat, Kb=m.Kb, mask=capi.DPM3D_ALL)
    X64, _ = oracle.run3d(m.X, m.mesh.faces, m.mesh.nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, nsteps, dt,
                          prec=64, **kw)
    X32, _ = oracle.run3d(m.X, m.mesh.faces, m.mesh.nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, nsteps, dt,
                          prec=32, **kw)
    eg = np.abs(t.positions() - X64).max(1)
    e32 = np.abs(X32 - X64).max(1)
    print(f"40-step all-term positions vs ref64: gpu max {eg.max():.3e} p99 {np.quantile(eg, 0.99):.3e}; "
          f"ref32 max {e32.max():.3e} p99 {np.quantile(e32, 0.99):.3e}", flush=True)
    assert np.quantile(eg, 0.99) <= max(2 * np.quantile(e32, 0.99), traj_bound(m.X, X64, nsteps))
    assert eg.max() <= max(2 * e32.max(), traj_bound(m.X, X64, nsteps))


def test_3d_full_size_shape_terms_and_determinism():
    m = W.monolayer_config("cfg2")
    shape = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_SUBSTRATE | capi.DPM3D_BENDING
    t = tissue3d(m, shape)
    t.set_positions(m.X)
    assert_forces(t.probe()["forces"], ref3d(m, m.X, shape, 64)["forces"], "3d cfg2 shape terms")
    full = tissue3d(m)
    full.set_positions(m.X)
    full.euler(100, m.dt)
    a, fa = full.positions(), full.forces()
    full.set_positions(m.X)
    full.euler(100, m.dt)
    np.testing.assert_array_equal(a, full.positions())
    np.testing.assert_array_equal(fa, full.forces())
    assert np.isfinite(a).all()


@pytest.mark.parametrize("kernel", [1, 2])
def test_3d_graph_replay_equals_eager(kernel):
    m = W.monolayer3d(4, 1.9, seed=5)
    t = tissue3d(m)
    t.set_step_kernel(kernel)
    t.set_positions(m.X)
    t.euler(150, 0.002)          # 149 steps through two 64-step graphs + eager
    a = t.positions()
    t.set_positions(m.X)
    for _ in range(150):
        t.euler(1, 0.002)        # eager, one launch per call
    np.testing.assert_array_equal(a, t.positions())


@pytest.mark.parametrize("kernel", [1, 2])
def test_3d_batch_equals_individual(kernel):
    ms = [W.monolayer3d(3, 1.8, seed=10), W.monolayer3d(2, 1.9, seed=11), W.monolayer3d(4, 2.1, seed=12)]
    ms[1].Kre, ms[1].Kat = 2.0, 0.3
    cb = np.concatenate([[0], np.cumsum([m.C for m in ms])])
    cat = lambda f: np.concatenate([np.broadcast_to(np.asarray(f(m), np.float32), (m.C,)) for m in ms])
    b = capi.Tissues3D(ms[0].mesh.faces, cb, cat(lambda m: m.Kv), cat(lambda m: m.Ka), cat(lambda m: m.Ks),
                       cat(lambda m: m.v0), cat(lambda m: m.a0), [m.L for m in ms], [m.Kre for m in ms],
                       Kat=[m.Kat for m in ms], Kb=cat(lambda m: m.Kb), mask=capi.DPM3D_ALL, nv=ms[0].mesh.nv)
    b.set_step_kernel(kernel)
    b.set_positions(np.concatenate([m.X for m in ms]))
    b.euler(30, 0.002)
    Xb = b.positions()
    off = 0
    for m in ms:
        t = tissue3d(m)
        t.set_step_kernel(kernel)
        t.set_positions(m.X)
        t.euler(30, 0.002)
        n = m.C * m.mesh.nv
        np.testing.assert_array_equal(Xb[off:off + n], t.positions())
        off += n


@pytest.mark.parametrize("kernel", [1, 2])
def test_3d_single_cell_tissue(kernel):
    # one cell in its box: no candidate cells, so no contact or adhesion pair,
    # only shape, substrate and bending forces; alone and batched beside a
    # 3 x 3 tissue (the batched step gives it the same bits as its own run)
    m = W.monolayer3d(1, 2.4, seed=7)
    t = tissue3d(m)
    t.set_step_kernel(kernel)
    t.set_positions(m.X)
    g = t.probe()
    r32 = ref3d(m, m.X, capi.DPM3D_ALL, 32, pairs=True)
    assert_forces(g["forces"], ref3d(m, m.X, capi.DPM3D_ALL, 64)["forces"], f"3d single cell kernel={kernel}")
    np.testing.assert_array_equal(g["down"], r32["down"])
    assert int(g["contact"].sum()) == 0 and int(g["adhesion"].sum()) == 0
    t.euler(30, 0.002)
    X1 = t.positions()
    assert np.isfinite(X1).all()
    ms = [m, W.monolayer3d(3, 1.8, seed=8)]
    cb = np.concatenate([[0], np.cumsum([q.C for q in ms])])
    cat = lambda f: np.concatenate([np.broadcast_to(np.asarray(f(q), np.float32), (q.C,)) for q in ms])
    b = capi.Tissues3D(m.mesh.faces, cb, cat(lambda q: q.Kv), cat(lambda q: q.Ka), cat(lambda q: q.Ks),
                       cat(lambda q: q.v0), cat(lambda q: q.a0), [q.L for q in ms], [q.Kre for q in ms],
                       Kat=[q.Kat for q in ms], Kb=cat(lambda q: q.Kb), mask=capi.DPM3D_ALL, nv=m.mesh.nv)
    b.set_step_kernel(kernel)
    b.set_positions(np.concatenate([q.X for q in ms]))
    b.euler(30, 0.002)
    np.testing.assert_array_equal(b.positions()[:m.mesh.nv], X1)


def test_3d_open_mesh_takes_the_per_corner_path():
    # a mesh with a face removed is not closed: the wave form's ring-edge
    # sum does not apply and it evaluates every (face, corner) term, with
    # the workgroup form's arithmetic
    m = W.monolayer3d(2, 2.4, seed=8)
    faces = np.ascontiguousarray(m.mesh.faces[1:])
    shape = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_SUBSTRATE | capi.DPM3D_BENDING
    F = []
    for kind in (1, 2):
        t = capi.Tissues3D(faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat, Kb=m.Kb,
                           mask=shape, nv=m.mesh.nv)
        t.set_step_kernel(kind)
        t.set_positions(m.X)
        F.append(t.probe(contact=False)["forces"])
    np.testing.assert_array_equal(F[0], F[1])


def test_3d_unoriented_mesh_takes_the_per_corner_path():
    # every edge still on two faces, but one face listed in the opposite
    # corner order: no vertex ring of that face's corners is one oriented
    # cycle, so the wave form takes the per-corner path (the workgroup
    # form's arithmetic, bit for bit)
    m = W.monolayer3d(2, 2.4, seed=9)
    faces = np.array(m.mesh.faces)
    faces[5] = faces[5][::-1]
    faces = np.ascontiguousarray(faces)
    shape = capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_SUBSTRATE | capi.DPM3D_BENDING
    F = []
    for kind in (1, 2):
        t = capi.Tissues3D(faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat, Kb=m.Kb,
                           mask=shape, nv=m.mesh.nv)
        t.set_step_kernel(kind)
        t.set_positions(m.X)
        F.append(t.probe(contact=False)["forces"])
    np.testing.assert_array_equal(F[0], F[1])


def test_3d_autotune_restores_state_and_keeps_a_configuration():
    ms = [W.monolayer3d(6, 1.9, seed=30 + k) for k in range(3)]
    cb = np.concatenate([[0], np.cumsum([m.C for m in ms])])
    cat = lambda f: np.concatenate([np.broadcast_to(np.asarray(f(m), np.float32), (m.C,)) for m in ms])

    def handle():
        return capi.Tissues3D(ms[0].mesh.faces, cb, cat(lambda m: m.Kv), cat(lambda m: m.Ka), cat(lambda m: m.Ks),
                              cat(lambda m: m.v0), cat(lambda m: m.a0), [m.L for m in ms], [m.Kre for m in ms],
                              Kat=[m.Kat for m in ms], Kb=cat(lambda m: m.Kb), mask=capi.DPM3D_ALL,
                              nv=ms[0].mesh.nv)

    X0 = np.concatenate([m.X for m in ms])
    t = handle()
    t.set_positions(X0)
    t.euler(5, 0.002)
    X5, c5 = t.positions(), t.counters()
    best, times = t.autotune(steps=3, dt=0.002)
    print(f"autotune: best {best}; ms per step {times}", flush=True)
    assert len(times) == 8 and all(v > 0 for v in times.values())
    np.testing.assert_array_equal(t.positions(), X5)
    np.testing.assert_array_equal(t.counters(), c5)
    t.euler(20, 0.002)
    # the same steps on a handle set to the chosen configuration explicitly
    u = handle()
    u.set_step_kernel(best["form"])
    u.set_cell_list(best["cell_list"])
    assert best["bin_width"] in (1.0, 1.5, 2.0)
    u.set_bin_width(best["bin_width"])
    u.set_positions(X0)
    u.euler(5, 0.002)
    u.euler(20, 0.002)
    np.testing.assert_array_equal(t.positions(), u.positions())


def test_3d_dropin_class_matches_c_abi():
    m = W.monolayer3d(3, 1.8, seed=6, keep_cells=True)
    for c in m.cells:
        c.Kv, c.Ka, c.Ks, c.Kb = m.Kv, m.Ka, m.Ks, m.Kb
    T = clDPM.Tissue3D(m.cells, 0.35)
    T.Kre, T.Kat, T.VertexAdhesion, T.Verbose = m.Kre, m.Kat, True, False
    L = T.L
    T.CLEulerUpdate(20, 0.002)
    Xd = np.concatenate([np.asarray(c.Verts, np.float32) for c in T.Cells])
    Fd = np.concatenate([np.asarray(c.Forces, np.float32) for c in T.Cells])
    t = capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, L, m.Kre, Kat=m.Kat, Kb=m.Kb,
                       mask=capi.DPM3D_ALL, nv=m.mesh.nv)
    t.set_positions(m.X)
    t.euler(20, 0.002)
    np.testing.assert_array_equal(Xd, t.positions())
    np.testing.assert_array_equal(Fd, t.forces())
    vol, com = t.cell_stats()
    host = np.array([c.GetVolume() for c in T.Cells])
    np.testing.assert_allclose(vol, host, rtol=1e-5)


# ----------------------------------------------------------------------- 2D

def tissue2d_handle(p: W.Packing2D, mask=capi.DPM2D_ALL):
    return capi.Tissues2D([0, p.C], p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, Kre=p.Kre, Kat=p.Kat,
                          mask=mask)


def ref2d(p: W.Packing2D, X, prec, mask=31, keys=False):
    P, stride = W.Packing2D(p.nv, X, p.a0, p.l0, p.r0, p.Ka, p.Kl, p.Kb, p.L).padded()
    F, ov, at, ah = oracle.forces2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat,
                                    mask=mask, prec=prec, attr_hash=True)
    keep = np.concatenate([np.arange(c * stride, c * stride + p.nv[c]) for c in range(p.C)])
    if keys:
        return F[keep], ov[keep], at[keep], ah[keep]
    return F[keep], ov[keep], at[keep]


CASES_2D = {
    # name: (ncells, nv, phi, Kb, Kat, seed, nv_range)
    "cfg1": (16, 24, 0.6, 0.0, 0.0, 0, None),
    "dense_attr": (16, 24, 0.92, 0.01, 0.5, 1, None),
    "uneven": (40, 0, 0.88, 0.01, 0.3, 2, (16, 64)),
}


@pytest.mark.parametrize("case", list(CASES_2D))
def test_2d_forces_and_pairs(case):
    nc, nv, phi, Kb, Kat, seed, rng = CASES_2D[case]
    p = W.tissue2d(nc, nv, 1.4, phi, Kb=Kb, Kat=Kat, seed=seed, nv_range=rng)
    t = tissue2d_handle(p)
    t.set_positions(p.X)
    g = t.probe()
    F64, _, _ = ref2d(p, p.X, 64)
    F32, ov32, at32, ah32 = ref2d(p, p.X, 32, keys=True)
    assert_forces(g["forces"], F64, f"2d {case}")
    np.testing.assert_array_equal(g["overlap"], ov32)
    np.testing.assert_array_equal(g["attraction"], at32)
    np.testing.assert_array_equal(g["attraction_hash"], ah32)
    print(f"pairs: overlap={int(ov32.sum())} attraction={int(at32.sum())}", flush=True)


def takes_parity_pass(p: W.Packing2D) -> bool:
    """k2d_run's split_over predicate for a one-tissue float handle, at the
    block size dpm2d_create gives it (four threads per vertex, up to 1024)."""
    nvt, nc = int(np.sum(p.nv)), p.C
    nt = min(1024, max(64, (4 * nvt + 63) // 64 * 64))
    return nvt <= 256 * nc and nvt * nc <= 16 * nt


@pytest.mark.parametrize("case", ["cfg1", "dense_attr"])
def test_2d_parity_pass_matches_vertex_loop(case):
    """The step's crossing-parity pass ((vertex, candidate) pairs over the
    whole workgroup) and the relaxation kernel's per-vertex crossing loop
    (its energy evaluation) give bit-identical forces on the same positions,
    at the start state and after 300 steps."""
    nc, nv, phi, Kb, Kat, seed, rng = CASES_2D[case]
    p = W.tissue2d(nc, nv, 1.4, phi, Kb=Kb, Kat=Kat, seed=seed, nv_range=rng)
    assert takes_parity_pass(p)
    t = tissue2d_handle(p)
    t.set_positions(p.X)
    for steps in (0, 300):
        if steps:
            t.euler(steps, 0.005)
        g = t.probe()
        t.energy()
        np.testing.assert_array_equal(t.forces(), g["forces"])
        print(f"2d {case} after {steps} steps: overlapping vertices {int((g['overlap'] > 0).sum())}", flush=True)


def test_2d_cfg1_1000_steps_vs_ref64():
    # cfg1 starts at mechanical rest (regular polygons at their own a0, l0),
    # so its forces are rounding-level; positions are the meaningful check.
    p = W.tissue2d(16, 24, 1.4, 0.6, seed=0)
    t = tissue2d_handle(p)
    t.set_positions(p.X)
    t.euler(1000, 0.005)
    P, stride = p.padded()
    run = lambda prec: oracle.run2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat,
                                    1000, 0.005, prec=prec)
    X64, _ = run(64)
    X32, _ = run(32)
    keep = np.concatenate([np.arange(c * stride, c * stride + p.nv[c]) for c in range(p.C)])
    X64, X32 = X64[keep], X32[keep]
    eg = float(np.abs(t.positions() - X64).max())
    e32 = float(np.abs(X32 - X64).max())
    print(f"2d cfg1 1000 steps: gpu max|dx|={eg:.3e}, ref32 {e32:.3e}, "
          f"bound {traj_bound(p.X, X64, 1000):.3e}", flush=True)
    assert eg <= max(2 * e32, traj_bound(p.X, X64, 1000))


def test_2d_dense_trajectory_vs_ref64():
    p = W.tissue2d(16, 24, 1.4, 0.92, Kb=0.01, Kat=0.5, seed=1)
    t = tissue2d_handle(p)
    t.set_positions(p.X)
    t.euler(100, 0.005)
    P, stride = p.padded()
    run = lambda prec: oracle.run2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat,
                                    100, 0.005, prec=prec)
    (X64, F64), (X32, _) = run(64), run(32)
    eg = float(np.abs(t.positions() - X64).max())
    e32 = float(np.abs(X32 - X64).max())
    print(f"2d dense 100 steps: gpu max|dx|={eg:.3e}, ref32 {e32:.3e}", flush=True)
    assert eg <= max(2 * e32, traj_bound(P, X64, 100))


def test_2d_batch_equals_individual():
    ps = W.packings2d(6, cells=(8, 40), verts=(16, 40), seed=3, phi=0.85)
    cb = np.concatenate([[0], np.cumsum([p.C for p in ps])])
    cat = lambda f: np.concatenate([f(p) for p in ps])
    b = capi.Tissues2D(cb, cat(lambda p: p.nv), cat(lambda p: p.Ka), cat(lambda p: p.Kl), cat(lambda p: p.Kb),
                       cat(lambda p: p.a0), cat(lambda p: p.l0), cat(lambda p: p.r0), [p.L for p in ps])
    b.set_positions(cat(lambda p: p.X))
    b.euler(200, 0.005)
    Xb = b.positions()
    off = 0
    for p in ps:
        t = tissue2d_handle(p)
        t.set_positions(p.X)
        t.euler(200, 0.005)
        np.testing.assert_array_equal(Xb[off:off + len(p.X)], t.positions())
        off += len(p.X)


def test_2d_single_cell_tissue():
    # one cell (its vertices jittered by 5 % of l0, away from the rest
    # shape): no overlap or attraction partner, shape forces only
    p = W.tissue2d(1, 24, 1.4, 0.6, Kb=0.01, Kat=0.5, seed=3)
    rng = np.random.default_rng(5)
    X = (p.X + rng.normal(0, 0.05 * float(np.mean(p.l0)), p.X.shape)).astype(np.float32)
    t = tissue2d_handle(p)
    t.set_positions(X)
    g = t.probe()
    F64, _, _ = ref2d(p, X, 64)
    assert_forces(g["forces"], F64, "2d single cell")
    assert int(g["overlap"].sum()) == 0 and int(g["attraction"].sum()) == 0
    t.euler(100, 0.005)
    assert np.isfinite(t.positions()).all()


def test_2d_dropin_class_matches_c_abi():
    p = W.tissue2d(16, 24, 1.4, 0.6, Kat=0.2, seed=7)
    cells = []
    off = 0
    for c in range(p.C):
        cell = clDPM.Cell2D(0.0, 0.0, 1.0, int(p.nv[c]), float(p.r0[c]))
        cell.Verts = p.X[off:off + p.nv[c]].tolist()
        cell.Ka, cell.Kl, cell.Kb = 1.0, 1.0, 0.0
        cells.append(cell)
        off += p.nv[c]
    T = clDPM.Tissue2D(cells, 0.7745966692414834)
    T.Kat = 0.2
    T.CLEulerUpdate(300, 0.005)
    Xd = np.concatenate([np.asarray(c.Verts, np.float32) for c in T.Cells])
    a0 = np.array([c.a0 for c in cells], np.float32)
    l0 = np.array([c.l0 for c in cells], np.float32)
    t = capi.Tissues2D([0, p.C], p.nv, 1.0, 1.0, 0.0, a0, l0, p.r0, T.L, Kre=1.0, Kat=0.2)
    t.set_positions(p.X)
    t.euler(300, 0.005)
    np.testing.assert_array_equal(Xd, t.positions())


@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("n_side,spacing,subdiv", [(20, 1.9, 2), (3, 1.7, 2), (8, 1.8, 3)])
def test_3d_cell_list_equals_scan(n_side, spacing, subdiv, kernel):
    # binned candidates (counting sort of the centres into hashed periodic
    # bins) against scanning every cell: identical pair sets, forces and
    # trajectories, bit for bit (the 3 x 3 box is too small for bins: the
    # cell list falls back to the scan)
    # at every bin width the autotuner tries, and a wider one
    m = W.monolayer3d(n_side, spacing, subdiv=subdiv, seed=9)
    if kernel == 2 and m.mesh.nv > 256:
        pytest.skip("the wave-per-cell step takes meshes of at most 256 vertices")
    out = []
    for mode, width in ((0, 1.0), (1, 1.0), (1, 1.5), (1, 2.0), (1, 3.0)):
        t = tissue3d(m)
        t.set_step_kernel(kernel)
        t.set_cell_list(mode)
        t.set_bin_width(width)
        assert t.bin_width() == width
        t.set_positions(m.X)
        g = t.probe()
        t.euler(20, 0.002)
        out.append((g, t.positions(), t.forces()))
    g0, X0, F0 = out[0]
    for g1, X1, F1 in out[1:]:
        np.testing.assert_array_equal(g0["forces"], g1["forces"])
        np.testing.assert_array_equal(g0["contact"], g1["contact"])
        np.testing.assert_array_equal(g0["adhesion"], g1["adhesion"])
        np.testing.assert_array_equal(g0["adhesion_hash"], g1["adhesion_hash"])
        np.testing.assert_array_equal(X0, X1)
        np.testing.assert_array_equal(F0, F1)
    print(f"cell list {n_side}x{n_side} s={subdiv} kernel={kernel}: contacts {int(g0['contact'].sum())}, "
          f"adhesion {int(g0['adhesion'].sum())}", flush=True)


@pytest.mark.parametrize("kernel", [1, 2])
def test_3d_cell_list_batch_of_uneven_tissues(kernel):
    ms = [W.monolayer3d(n, 1.9, seed=s) for n, s in ((12, 1), (3, 2), (9, 3))]
    cb = np.concatenate([[0], np.cumsum([m.C for m in ms])])
    cat = lambda f: np.concatenate([np.broadcast_to(np.asarray(f(m), np.float32), (m.C,)) for m in ms])
    res = []
    for mode in (0, 1):
        t = capi.Tissues3D(ms[0].mesh.faces, cb, cat(lambda m: m.Kv), cat(lambda m: m.Ka), cat(lambda m: m.Ks),
                           cat(lambda m: m.v0), cat(lambda m: m.a0), [m.L for m in ms], [m.Kre for m in ms],
                           Kat=[m.Kat for m in ms], Kb=cat(lambda m: m.Kb), mask=capi.DPM3D_ALL,
                           nv=ms[0].mesh.nv)
        t.set_step_kernel(kernel)
        t.set_cell_list(mode)
        t.set_positions(np.concatenate([m.X for m in ms]))
        t.euler(70, 0.002)  # past one 64-step graph
        res.append((t.positions(), t.forces()))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
