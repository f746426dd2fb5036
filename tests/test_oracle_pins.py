"""Pins the CPU oracle (oracle/dpm_oracle.c) to what the mathematics fixes.

The reference kernels cannot execute here (no OpenCL device, no CPU OpenCL
runtime), so each oracle term is checked against a closed form, an
invariant, or a finite-difference derivative of an energy evaluated by an
independent numpy expression — chosen so that a dropped term, a wrong sign
or index, or a transposed operand fails.  The reference's quirks are pinned
the same way (the 2D area force's y sign, the 3D adhesion sign).
"""
import math

import numpy as np
import pytest

import oracle
from opencl_dpm_amd import workloads as W


def regular_polygon(n, r=1.0, cx=0.0, cy=0.0):
    th = 2 * np.pi * (np.arange(n) + 1) / n
    return np.stack([r * np.cos(th) + cx, r * np.sin(th) + cy], 1)


def shoelace(P):  # reference orientation (cell.cpp:35-47), absolute value
    Pm = np.roll(P, 1, axis=0)
    return abs(0.5 * np.sum((Pm[:, 0] + P[:, 0]) * (Pm[:, 1] - P[:, 1])))


def fd_grad(E, P, h=1e-6):
    G = np.zeros_like(P)
    for idx in np.ndindex(P.shape):
        Pp, Pm = P.copy(), P.copy()
        Pp[idx] += h
        Pm[idx] -= h
        G[idx] = (E(Pp) - E(Pm)) / (2 * h)
    return G


def f2d(P, mask, Ka=1.3, Kl=0.7, Kb=0.4, a0=2.5, l0=0.5, r0=1.0, prec=64, L=1000.0, Kre=1.0, Kat=0.0):
    n = P.shape[0]
    F, ov, at = oracle.forces2d(P, n, [n], Ka, Kl, Kb, a0, l0, r0, L, Kre, Kat, mask=mask, prec=prec)
    return F


@pytest.fixture
def poly():
    rng = np.random.Generator(np.random.PCG64(3))
    P = regular_polygon(12, 1.0, 0.3, -0.2)
    return P + rng.normal(0, 0.05, P.shape)


def test_area_force_x_is_gradient_y_has_reference_sign(poly):
    Ka, a0 = 1.3, 2.5
    E = lambda P: Ka * math.sqrt(a0) * 0.5 * (shoelace(P) / a0 - 1.0) ** 2
    G = fd_grad(E, poly)
    F = f2d(poly, 1, Ka=Ka, a0=a0)
    # Cell2D_kernel.cl:38-41: F.x = -dE/dx, F.y = +dE/dy (the y sign quirk)
    np.testing.assert_allclose(F[:, 0], -G[:, 0], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(F[:, 1], +G[:, 1], rtol=1e-6, atol=1e-9)


def test_perimeter_force_is_gradient(poly):
    Kl, a0, l0 = 0.7, 2.5, 0.5
    C = Kl * math.sqrt(a0 / l0)

    def E(P):
        l = np.linalg.norm(np.roll(P, -1, 0) - P, axis=1)
        return C * l0 / 2 * np.sum((l / l0 - 1.0) ** 2)

    np.testing.assert_allclose(f2d(poly, 2, Kl=Kl, a0=a0, l0=l0), -fd_grad(E, poly), rtol=1e-6, atol=1e-9)


def test_bending_closed_form_regular_polygon():
    n, Kb = 24, 0.4
    P = regular_polygon(n, 1.7, 2.0, -1.0)
    c = 2 - 2 * math.cos(2 * math.pi / n)
    F = f2d(P, 4, Kb=Kb)
    np.testing.assert_allclose(F, -Kb * c * c * (P - [2.0, -1.0]), rtol=1e-9, atol=1e-12)


def test_bending_is_gradient(poly):
    Kb = 0.4

    def E(P):
        s = np.roll(P, -1, 0) - 2 * P + np.roll(P, 1, 0)
        return Kb / 2 * np.sum(s ** 2)

    np.testing.assert_allclose(f2d(poly, 4, Kb=Kb), -fd_grad(E, poly), rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("mask", [1, 2, 4, 7])
def test_shape_forces_sum_to_zero(poly, mask):
    for prec, tol in [(64, 1e-13), (32, 1e-5)]:
        F = f2d(poly, mask, prec=prec)
        assert np.abs(F.sum(0)).max() < tol


def test_area_strain_zero_on_fresh_polygon():
    from opencl_dpm_amd import clDPM

    c = clDPM.Cell2D(0.0, 0.0, 1.0, 24, 1.0)
    P = np.array(c.Verts)
    F = f2d(P, 3, a0=c.a0, l0=c.l0, prec=32)
    # A = a0 exactly (same shoelace order) and l ~= l0 to rounding
    assert np.abs(F).max() < 2e-6


# ------------------------------------------ conservative 2D terms, FIRE, jamming

@pytest.mark.parametrize("orient", [1, -1])
def test_gradient_area_force_is_gradient_either_orientation(poly, orient):
    P = poly[::orient].copy()
    Ka, a0 = 1.3, 2.5
    E = lambda Q: Ka * math.sqrt(a0) * 0.5 * (shoelace(Q) / a0 - 1.0) ** 2
    np.testing.assert_allclose(f2d(P, 32, Ka=Ka, a0=a0), -fd_grad(E, P), rtol=1e-6, atol=1e-9)
    if orient == 1:  # counter-clockwise: the reference's x, the flipped y
        Fr, Fg = f2d(P, 1, Ka=Ka, a0=a0), f2d(P, 32, Ka=Ka, a0=a0)
        np.testing.assert_allclose(Fg[:, 0], Fr[:, 0], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(Fg[:, 1], -Fr[:, 1], rtol=1e-12, atol=1e-15)


def overlapping_pair(offset=1.9, L=1000.0):
    A = regular_polygon(16, 1.0, 0.0, 0.0)
    B = regular_polygon(16, 1.0, offset, 0.05)
    P = np.concatenate([A, B])
    return P, 16


def vrep_energy_np(P, stride, nv, l0, Kre, L):
    # independent restatement: pairs of vertices of different cells
    U = 0.0
    off = [0, stride]
    for ci in range(2):
        for cj in range(ci + 1, 2):
            sig = 0.5 * (l0[ci] + l0[cj])
            A = P[off[ci]:off[ci] + nv[ci]]
            B = P[off[cj]:off[cj] + nv[cj]]
            d = B[None] - A[:, None]
            d -= L * np.round(d / L)
            r = np.sqrt((d ** 2).sum(-1))
            U += np.sum(np.where(r < sig, Kre / 2 * (1 - r / sig) ** 2, 0.0))
    return U


@pytest.mark.parametrize("L", [1000.0, 3.5])
def test_vertex_repulsion_is_gradient_and_pairwise(L):
    P, stride = overlapping_pair(1.9 if L > 100 else 3.5 - 1.6)
    nv, l0, Kre = [16, 16], [0.5, 0.3], 0.8
    F, _, _ = oracle.forces2d(P, stride, nv, 1, 1, 0, 3, l0, 1, L, Kre, 0.0, mask=64, prec=64)
    E = lambda Q: vrep_energy_np(Q, stride, nv, l0, Kre, L)
    assert E(P) > 0
    np.testing.assert_allclose(F, -fd_grad(E, P), rtol=1e-6, atol=1e-9)
    assert np.abs(F.sum(0)).max() < 1e-12  # equal and opposite pair forces
    e, ei = oracle.energy2d(P, stride, nv, 1, 1, 0, 3, l0, L, Kre, mask=64)
    assert e == ei and abs(ei - E(P)) < 1e-12 * E(P)


def test_energy2d_is_potential_of_conservative_forces():
    P, stride = overlapping_pair(1.85)
    rng = np.random.Generator(np.random.PCG64(4))
    P = P + rng.normal(0, 0.02, P.shape)
    nv, Ka, Kl, Kb, a0, l0, Kre = [16, 16], [1.3, 0.9], [0.7, 1.1], [0.05, 0.02], [2.9, 3.3], [0.4, 0.45], 0.8
    F, _, _ = oracle.forces2d(P, stride, nv, Ka, Kl, Kb, a0, l0, 1, 1000.0, Kre, 0.0, mask=102, prec=64)
    E = lambda Q: oracle.energy2d(Q, stride, nv, Ka, Kl, Kb, a0, l0, 1000.0, Kre, mask=102)[0]
    np.testing.assert_allclose(F, -fd_grad(E, P), rtol=1e-6, atol=1e-9)


def test_fire_minimises_to_tolerance_and_lowers_energy():
    P, stride = overlapping_pair(1.8)
    nv, Ka, Kl, Kb, a0, l0 = [16, 16], 1.0, 1.0, 0.01, [3.06, 3.06], [0.39, 0.39]
    e0 = oracle.energy2d(P, stride, nv, Ka, Kl, Kb, a0, l0, 1000.0, 1.0)[0]
    V, Vel, st = oracle.fire2d(P, stride, nv, Ka, Kl, Kb, a0, l0, 1, 1000.0, 1.0, 0.0, mask=102, max_iter=50000)
    assert st["converged"] == 1 and st["fmax"] < 1e-10
    F, _, _ = oracle.forces2d(V, stride, nv, Ka, Kl, Kb, a0, l0, 1, 1000.0, 1.0, 0.0, mask=102, prec=64)
    assert np.abs(F).max() < 1e-10
    assert oracle.energy2d(V, stride, nv, Ka, Kl, Kb, a0, l0, 1000.0, 1.0)[0] < e0


def test_fire_restart_rule():
    # with ftol 0 and one iteration from rest: P = 0 (v = 0) changes neither
    # dt nor alpha, and the step is v = F dt0, x += v dt0
    P, stride = overlapping_pair(1.8)
    nv = [16, 16]
    F, _, _ = oracle.forces2d(P, stride, nv, 1, 1, 0, 3.06, 0.39, 1, 1000.0, 1.0, 0.0, mask=102, prec=64)
    V, Vel, st = oracle.fire2d(P, stride, nv, 1, 1, 0, 3.06, 0.39, 1, 1000.0, 1.0, 0.0, mask=102, max_iter=1,
                               ftol=0.0)
    assert st["dt"] == 0.01 and st["alpha"] == 0.1 and st["n_pos"] == 0
    np.testing.assert_allclose(Vel, F * 0.01, rtol=1e-15)
    np.testing.assert_allclose(V, P + F * 0.01 * 0.01, rtol=1e-14)


def test_jamming_brackets_and_ends_unjammed():
    m = W.tissue2d(4, 16, 1.4, 0.4, seed=11, calA=1.02)
    P, stride = m.padded()
    V, L, st = oracle.jam2d(P, stride, m.nv, m.Ka, m.Kl, m.Kb, m.a0, m.l0, m.r0, m.L, m.Kre, 0.0,
                            fire=dict(max_iter=20000), phi0=0.4, dphi=0.1, phi_tol=0.02)
    assert st["status"] == 0 and 0.4 <= st["phi_lo"] < st["phi_hi"] <= st["phi_lo"] + 0.02 + 1e-12
    A0 = float(np.sum(m.a0.astype(np.float64)))
    assert abs(A0 / L ** 2 - st["phi_lo"]) < 1e-12
    e, ei = oracle.energy2d(V, stride, m.nv, m.Ka, m.Kl, m.Kb, m.a0, m.l0, L, m.Kre)
    assert ei / 4 <= 1e-16  # the returned state is the unjammed one


def two_cells_2d(offset):
    A = regular_polygon(16, 1.0, 0.0, 0.0)
    B = regular_polygon(16, 1.0, offset, 0.1)
    return np.concatenate([A, B])


def _inside(poly, p):
    # even-odd ray casting toward -x (Cell2D_kernel.cl:191-193 in d = p - v)
    x, y = p
    odd = False
    n = len(poly)
    for i in range(n):
        xi, yi = poly[i]
        xj, yj = poly[i - 1]
        if (yi > y) != (yj > y):
            xc = xi + (y - yi) * (xj - xi) / (yj - yi)
            if xc < x:
                odd = not odd
    return odd


def test_repulsion_pairs_and_direction():
    P = two_cells_2d(1.5)
    n = 16
    Kre, r0 = 2.0, 1.0
    F, ov, at = oracle.forces2d(P, n, [n, n], 0, 0, 0, 3.0, 0.4, r0, 1000.0, Kre, 0.0, mask=16, prec=64)
    A, B = P[:n], P[n:]
    want = [_inside(B, p) for p in A] + [_inside(A, p) for p in B]
    assert list(ov.astype(bool)) == want
    assert sum(want) > 0
    com = [A.mean(0), B.mean(0)]
    for k in np.nonzero(ov)[0]:
        c = com[k // n]
        d = c - P[k]
        dist = np.linalg.norm(d)
        expect = 0.5 * Kre * (1 - dist / (2 * r0)) * d / dist
        np.testing.assert_allclose(F[k], expect, rtol=1e-12)
    assert np.all(F[ov == 0] == 0)


def test_attraction_pairs_and_magnitude():
    P = two_cells_2d(2.05)
    n, l0, Kat = 16, 0.4, 0.9
    F, ov, at = oracle.forces2d(P, n, [n, n], 0, 0, 0, 3.0, l0, 1.0, 1000.0, 0.0, Kat, mask=8, prec=64)
    for k in range(2 * n):
        others = P[n:] if k < n else P[:n]
        d = others - P[k]
        dist = np.linalg.norm(d, axis=1)
        sel = dist < l0
        assert at[k] == sel.sum()
        expect = np.sum((Kat / n * dist[sel] / l0)[:, None] * d[sel] / dist[sel][:, None], 0) if sel.any() else 0
        np.testing.assert_allclose(F[k], expect, rtol=1e-12, atol=1e-15)
    assert at.sum() > 0


def test_attraction_uses_minimum_image():
    # two cells at opposite box edges attract through the boundary
    L = 10.0
    P = np.concatenate([regular_polygon(16, 1.0, 0.5, 5.0), regular_polygon(16, 1.0, L - 1.55, 5.0)])
    F, ov, at = oracle.forces2d(P, 16, [16, 16], 0, 0, 0, 3.0, 0.4, 1.0, L, 0.0, 1.0, mask=8, prec=64)
    assert at.sum() > 0


# ----------------------------------------------------------------------- 3D

@pytest.fixture(scope="module")
def small3d():
    m = W.monolayer3d(n_side=2, spacing=2.4, seed=1)
    return m


def ico(subdiv=2, r=1.0, centre=(0.0, 0.0, 0.0), seed=None, noise=0.0):
    mesh = W.icosphere(subdiv)
    X = mesh.unit.astype(np.float64) * r + np.asarray(centre)
    if seed is not None:
        X = X + np.random.Generator(np.random.PCG64(seed)).normal(0, noise, X.shape)
    return mesh, X


def volume_np(X, faces):
    P0, P1, P2 = X[faces[:, 0]], X[faces[:, 1]], X[faces[:, 2]]
    return abs(np.sum(np.einsum("ij,ij->i", np.cross(P0, P1), P2)) / 6.0)


def test_volume_matches_closed_mesh_formula():
    mesh, X = ico(2, 1.3, (2.0, -1.0, 0.5), seed=0, noise=0.02)
    v = oracle.volumes(X, mesh.faces, mesh.nv, prec=64)[0]
    assert abs(v - volume_np(X, mesh.faces)) < 1e-12 * v
    # icosphere volume approaches the sphere from below
    m2, X2 = ico(3, 1.0)
    assert 0.95 * 4 / 3 * math.pi < oracle.volumes(X2, m2.faces, m2.nv)[0] < 4 / 3 * math.pi


def F3(X, mesh, mask, prec=64, **kw):
    p = dict(Kv=1.3, Ka=0.7, Ks=0.5, v0=3.0, a0=0.05, L=1000.0, Kre=1.0, Kat=0.0, Kb=0.3)
    p.update(kw)
    return oracle.forces3d(X, mesh.faces, mesh.nv, p["Kv"], p["Ka"], p["Ks"], p["v0"], p["a0"], p["L"],
                           p["Kre"], Kat=p["Kat"], Kb=p["Kb"], mask=mask, prec=prec)


def test_volume_force_is_gradient():
    mesh, X = ico(1, 1.1, (0.3, 0.2, 1.5), seed=2, noise=0.03)
    Kv, v0 = 1.3, 3.0
    E = lambda Y: Kv * v0 / 2 * (volume_np(Y, mesh.faces) / v0 - 1.0) ** 2
    F = F3(X, mesh, 1, Kv=Kv, v0=v0)["forces"]
    np.testing.assert_allclose(F, -fd_grad(E, X), rtol=1e-6, atol=1e-9)


def test_surface_force_is_gradient_of_edge_springs():
    mesh, X = ico(1, 1.1, (0.3, 0.2, 1.5), seed=3, noise=0.03)
    Ka, a0 = 0.7, 0.05
    l0 = math.sqrt(4 * a0) / math.sqrt(3)  # Tissue3D.cpp:177
    sc = Ka * math.sqrt(a0) / l0 * 0.3

    def E(Y):
        f = mesh.faces
        e = [Y[f[:, 1]] - Y[f[:, 0]], Y[f[:, 2]] - Y[f[:, 1]], Y[f[:, 0]] - Y[f[:, 2]]]
        return sum(sc * l0 / 2 * np.sum((np.linalg.norm(x, axis=1) / l0 - 1.0) ** 2) for x in e)

    F = F3(X, mesh, 2, Ka=Ka, a0=a0)["forces"]
    np.testing.assert_allclose(F, -fd_grad(E, X), rtol=1e-6, atol=1e-9)


def vertex_rings(faces, nv):
    # u_0 .. u_(k-1) per vertex with face q = (v, u_q, u_(q+1)) in its corner order
    nxt = [dict() for _ in range(nv)]
    for f in faces:
        for c in range(3):
            nxt[f[c]][f[(c + 1) % 3]] = f[(c + 2) % 3]
    rings = []
    for v in range(nv):
        u0 = next(iter(nxt[v]))
        r, u = [u0], nxt[v][u0]
        while u != u0:
            r.append(u)
            u = nxt[v][u]
        assert len(r) == len(nxt[v])
        rings.append(r)
    return rings


def test_ring_walk_gives_the_volume_and_edge_spring_forces():
    # The wave-per-cell step walks each vertex's ring once: volume
    # kvol sum_q d_q x d_(q+1) (d_q = u_q - p; the centre-relative corner
    # terms sum to this on a closed ring) and edge springs 2 sc sum_q s(d_q).
    mesh, X = ico(2, 1.2, (0.4, -0.3, 1.3), seed=11, noise=0.03)
    Kv, v0, Ka, a0 = 1.3, 3.0, 0.7, 0.05
    rings = vertex_rings(mesh.faces, mesh.nv)
    strain = volume_np(X, mesh.faces) / v0 - 1.0
    kvol = -Kv * strain / 6.0
    l0 = math.sqrt(4 * a0) / math.sqrt(3)
    sc = Ka * math.sqrt(a0) / l0 * float(np.float32(0.3))  # the kernels' 0.3f
    Fv, Fs = np.zeros_like(X), np.zeros_like(X)
    for v, r in enumerate(rings):
        d = X[r] - X[v]
        Fv[v] = kvol * np.cross(d, np.roll(d, -1, axis=0)).sum(0)
        n = np.linalg.norm(d, axis=1, keepdims=True)
        Fs[v] = 2 * sc * (d * (n / l0 - 1.0) / n).sum(0)
    np.testing.assert_allclose(Fv, F3(X, mesh, 1, Kv=Kv, v0=v0)["forces"], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(Fs, F3(X, mesh, 2, Ka=Ka, a0=a0)["forces"], rtol=1e-10, atol=1e-13)
    # the opposite ring direction flips the volume term: orientation matters
    Fr = np.array([kvol * np.cross(X[r[::-1]] - X[v], np.roll(X[r[::-1]] - X[v], -1, axis=0)).sum(0)
                   for v, r in enumerate(rings)])
    assert np.abs(Fr + Fv).max() < 1e-12 and np.abs(Fv).max() > 1e-3


def test_bending_is_gradient_and_invariant():
    mesh, X = ico(1, 1.1, (0.3, 0.2, 1.5), seed=4, noise=0.05)
    Kb = 0.3
    adj = oracle.adjacency(mesh.faces, mesh.nv)

    def lap(Y):
        out = np.zeros_like(Y)
        for v in range(mesh.nv):
            nb = adj[v][adj[v] >= 0]
            out[v] = Y[nb].sum(0) - len(nb) * Y[v]
        return out

    E = lambda Y: Kb / 2 * np.sum(lap(Y) ** 2)
    F = F3(X, mesh, 32, Kb=Kb)["forces"]
    np.testing.assert_allclose(F, -fd_grad(E, X), rtol=1e-6, atol=1e-9)
    # translation invariant, linear in positions
    np.testing.assert_allclose(F3(X + [5.0, -3.0, 2.0], mesh, 32, Kb=Kb)["forces"], F, atol=1e-9)
    np.testing.assert_allclose(F3(2.0 * X, mesh, 32, Kb=Kb)["forces"], 2.0 * F, rtol=1e-9, atol=1e-9)


def test_bending_on_a_ring_is_the_2d_reference_bending(poly):
    # The 3D bending F = -Kb L^T L x (L the graph Laplacian of the mesh's
    # edge graph) on a cycle graph is the reference's 2D bending
    # Kb (2 s_i - s_(i-1) - s_(i+1)), s_i = x_(i+1) - 2 x_i + x_(i-1)
    # (Cell2D_kernel.cl:44-90): the operator the 3D oracle applies to the
    # icosphere graph in the test above.
    Kb, n = 0.4, poly.shape[0]
    Lc = -2.0 * np.eye(n) + np.roll(np.eye(n), 1, axis=1) + np.roll(np.eye(n), -1, axis=1)
    F_ring = -Kb * Lc.T @ Lc @ poly
    np.testing.assert_allclose(f2d(poly, 4, Kb=Kb), F_ring, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mask", [1, 2, 32, 35])
def test_energy_is_potential_of_shape_forces(mask):
    # the oracle's energy (used for the drift study) against the oracle's
    # forces, themselves pinned above against independent energies
    mesh, X1 = ico(1, 1.1, (0.3, 0.2, 1.5), seed=7, noise=0.04)
    _, X2 = ico(1, 0.9, (3.0, -1.0, 0.7), seed=8, noise=0.04)
    X = np.concatenate([X1, X2])
    p = dict(Kv=[1.3, 0.8], Ka=[0.7, 0.4], v0=[3.0, 2.5], a0=[0.05, 0.07], Kb=[0.3, 0.1])
    E = lambda Y: oracle.energy3d(Y, mesh.faces, mesh.nv, p["Kv"], p["Ka"], p["v0"], p["a0"],
                                  Kb=p["Kb"], mask=mask).sum()
    F = oracle.forces3d(X, mesh.faces, mesh.nv, p["Kv"], p["Ka"], 0.5, p["v0"], p["a0"], 1000.0, 0.0,
                        Kb=p["Kb"], mask=mask, prec=64)["forces"]
    np.testing.assert_allclose(F, -fd_grad(E, X), rtol=1e-6, atol=1e-8)
    # zero for the disabled terms, translation invariant
    assert oracle.energy3d(X, mesh.faces, mesh.nv, 1, 1, 3, 0.05, Kb=1, mask=0).max() == 0.0
    np.testing.assert_allclose(E(X + [4.0, -2.0, 1.0]), E(X), rtol=1e-12)


@pytest.mark.parametrize("mask", [1, 2, 32, 35])
def test_3d_internal_forces_sum_to_zero(mask):
    mesh, X = ico(2, 1.0, (1.0, 2.0, 3.0), seed=5, noise=0.02)
    F = F3(X, mesh, mask)["forces"]
    assert np.abs(F.sum(0)).max() < 1e-12


def test_winding_number_inside_outside():
    mesh, X = ico(2, 1.0, (0.0, 0.0, 0.0), seed=6, noise=0.01)
    assert abs(oracle.winding(X, mesh.faces, mesh.nv, 0, [0.0, 0.0, 0.0]) - 1.0) < 1e-12
    assert abs(oracle.winding(X, mesh.faces, mesh.nv, 0, [0.1, -0.2, 0.3]) - 1.0) < 1e-12
    assert abs(oracle.winding(X, mesh.faces, mesh.nv, 0, [3.0, 0.0, 0.0])) < 1e-12
    assert abs(oracle.winding(X, mesh.faces, mesh.nv, 0, [0.0, 1.5, 0.2])) < 1e-12


def test_winding_matches_independent_solid_angles():
    # Van Oosterom-Strackee in unnormalised form (no hemisphere skip) on
    # points off the surface: equals the oracle's sum wherever no face
    # subtends more than a hemisphere.
    mesh, X = ico(1, 1.0, (0.0, 0.0, 0.0), seed=7, noise=0.02)
    rng = np.random.Generator(np.random.PCG64(8))
    for _ in range(20):
        p = rng.uniform(-2, 2, 3)
        r = np.linalg.norm(p)
        if 0.8 < r < 1.2:
            continue
        a, b, c = (X[mesh.faces[:, k]] - p for k in range(3))
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", b, c) * la + \
            np.einsum("ij,ij->i", c, a) * lb
        if np.any(den <= 0):
            continue
        w = np.sum(2 * np.arctan2(num, den)) / (4 * np.pi)
        assert abs(oracle.winding(X, mesh.faces, mesh.nv, 0, p) - w) < 1e-12


@pytest.mark.parametrize("subdiv,noise", [(1, 0.03), (2, 0.02), (2, 0.05), (3, 0.01)])
def test_contact_culling_margin(subdiv, noise):
    # The step skips the winding number of a point farther than
    # R_j + 0.4 e_max from a cell's centre, and of a point beyond
    # h + 0.4 e_max along its own direction d from the centre, h the support
    # max_v (x_v - C).d / |d| (outside_by_margin).  Lemma behind both: every point at
    # least 0.4 e_max from the convex hull (here: beyond the support plane in
    # its own direction) sees each face with denominator >= 0.34, so no face
    # is skipped and the float winding number is noise far below the 1e-6
    # pair threshold.  Points on the boundary of the culled region, in float.
    mesh, X = ico(subdiv, 1.0, (0.3, -0.2, 1.1), seed=subdiv + 17, noise=noise)
    C = X.mean(0)
    F = mesh.faces
    e = max(np.linalg.norm(X[F[:, a]] - X[F[:, b]], axis=1).max() for a, b in ((0, 1), (1, 2), (2, 0)))
    rng = np.random.Generator(np.random.PCG64(subdiv))
    u = rng.normal(size=(300, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    h = ((X - C) @ u.T).max(0)  # support of the vertex set
    def worst(f):
        dmin, wmax = np.inf, 0.0
        for k in range(len(u)):
            p = C + u[k] * (h[k] + f * e)
            a, b, c = (X[F[:, i]] - p for i in range(3))
            ua, ub, uc = (x / np.linalg.norm(x, axis=1, keepdims=True) for x in (a, b, c))
            den = 1 + np.einsum("ij,ij->i", ua, ub) + np.einsum("ij,ij->i", ub, uc) + \
                np.einsum("ij,ij->i", uc, ua)
            dmin = min(dmin, den.min())
            w = oracle.winding(X.astype(np.float32), F, mesh.nv, 0, p.astype(np.float32), prec=32)
            wmax = max(wmax, abs(w))
        return dmin, wmax

    dmin, wmax = worst(0.4)
    print(f"subdiv {subdiv}: at 0.4 e_max min denominator {dmin:.3f}, max |float winding| {wmax:.2e}")
    assert dmin >= 0.34
    assert wmax < 1e-6
    # closer than that, faces are skipped and the winding number is not 0
    dmin, wmax = worst(0.05)
    assert dmin < 1e-8 and wmax > 1e-2


def test_contact_pairs_of_overlapping_cells():
    mesh = W.icosphere(2)
    X = np.concatenate([mesh.unit * 1.0, mesh.unit * 1.0 + [1.6, 0.0, 0.0]]).astype(np.float64)
    res = oracle.forces3d(X, mesh.faces, mesh.nv, 1, 1, 1, 3.0, 0.05, 1000.0, 2.0, mask=8, prec=64,
                          pairs=True, winding=True)
    pairs = res["contact"]
    # a vertex of cell 0 is in contact with cell 1 iff it lies inside the
    # sphere of cell 1 (up to the faceted surface)
    d = np.linalg.norm(X[:mesh.nv] - [1.6, 0, 0], axis=1)
    assert np.all(pairs[:mesh.nv, 1][d < 0.95] == 1)
    assert np.all(pairs[:mesh.nv, 1][d > 1.05] == 0)
    w = res["winding"][:mesh.nv, 1]
    np.testing.assert_allclose(w[d < 0.9], 1.0, atol=1e-9)
    F = res["forces"][:mesh.nv]
    com0 = X[:mesh.nv].mean(0)
    for v in np.nonzero(pairs[:mesh.nv, 1])[0]:
        dv = com0 - X[v]
        np.testing.assert_allclose(F[v], abs(w[v]) * 0.5 * 2.0 * dv / np.linalg.norm(dv), rtol=1e-10)


def test_adhesion_is_plus_gradient_of_pair_spring():
    # Two cells with equal l0: F = +grad E, E = sum_pairs Kat*L0/2 (d/L0-1)^2
    # (the sign convention of Cell3D_Kernel.cl:353-360), away from cut-offs.
    mesh = W.icosphere(1)
    a0 = 0.2
    L0 = math.sqrt(4 * a0) / math.sqrt(3)
    X = np.concatenate([mesh.unit * 1.0, mesh.unit * 1.0 + [2.0 + 0.8 * L0, 0.0, 0.0]]).astype(np.float64)
    Kat = 0.7
    nv = mesh.nv

    def pairs_d(Y):
        d = Y[nv:][None, :, :] - Y[:nv][:, None, :]
        return np.linalg.norm(d, axis=2)

    D = pairs_d(X)
    sel = (D < 2 * L0) & (D > 1e-12)
    assert sel.sum() > 0 and np.abs(D[sel] - 2 * L0).min() > 1e-3

    def E(Y):
        d = pairs_d(Y)[sel]
        return np.sum(Kat * L0 / 2 * (d / L0 - 1.0) ** 2)

    F = oracle.forces3d(X, mesh.faces, nv, 1, 1, 1, 3.0, a0, 1000.0, 0.0, Kat=Kat, mask=16, prec=64)["forces"]
    np.testing.assert_allclose(F, fd_grad(E, X), rtol=1e-6, atol=1e-9)


def test_substrate_faces_and_zero_far_from_plane():
    mesh, X = ico(2, 1.0, (0.0, 0.0, 0.9))
    res = F3(X, mesh, 4)
    f = mesh.faces
    n = np.cross(X[f[:, 1]] - X[f[:, 0]], X[f[:, 2]] - X[f[:, 0]])
    nz = n[:, 2] / np.linalg.norm(n, axis=1)
    np.testing.assert_array_equal(res["down"][0], (nz < -0.1).astype(np.uint8))
    far = F3(X + [0, 0, 5.0], mesh, 4)
    assert np.all(far["forces"] == 0)
    # a vertex below the plane is pushed up by Ks|z| per facing-down face
    Y = X.copy()
    low = np.argmin(Y[:, 2])
    Y[:, 2] -= Y[low, 2] + 0.05   # lowest vertex at z = -0.05
    r = F3(Y, mesh, 4, Ks=0.5)
    ndown = sum(1 for k in range(len(f)) if low in f[k] and r["down"][0][k])
    assert ndown > 0
    a0 = 0.05
    l0 = math.sqrt(4 * a0) / math.sqrt(3)
    com = Y.mean(0)
    c2v = np.array([Y[low, 0] - com[0], Y[low, 1] - com[1], -com[2]])
    expect = ndown * (np.array([0, 0, 0.5 * 0.05]) + c2v / np.linalg.norm(c2v) * 0.5 * (1 - 0.05 / l0))
    np.testing.assert_allclose(r["forces"][low], expect, rtol=1e-9)


def test_ref32_close_to_ref64_near_origin(small3d):
    m = small3d
    X = m.X - np.float32(2.4)  # keep coordinates small: no cancellation in float
    a = oracle.forces3d(X, m.mesh.faces, m.mesh.nv, 1.0, 0.5, 0.2, m.v0, m.a0, m.L, 1.0, Kat=0.1, Kb=0.01,
                        mask=63, prec=32)["forces"]
    b = oracle.forces3d(X, m.mesh.faces, m.mesh.nv, 1.0, 0.5, 0.2, m.v0, m.a0, m.L, 1.0, Kat=0.1, Kb=0.01,
                        mask=63, prec=64)["forces"]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5


def test_euler_step_is_fused_update():
    mesh, X = ico(1, 1.0, (0.2, 0.1, 1.4), seed=9, noise=0.02)
    X = X.astype(np.float32)
    F = F3(X, mesh, 3, prec=32)["forces"]
    Xn, Fl = oracle.run3d(X, mesh.faces, mesh.nv, 1.3, 0.7, 0.5, 3.0, 0.05, 1000.0, 1.0, 1, 0.01, mask=3, prec=32)
    np.testing.assert_array_equal(Fl, F)
    x64 = X.astype(np.float64) + F.astype(np.float64) * np.float64(np.float32(0.01))
    np.testing.assert_array_equal(Xn, x64.astype(np.float32))  # one rounding: fma


def test_2d_euler_loop_is_fused_update_with_last_step_forces():
    # ora2d_run (Tissue2D.cpp:215-229 with EulerUpdate, Cell2D_kernel.cl:270-281):
    # each step evaluates the forces at the current positions and moves every
    # vertex by x + F dt rounded once (fused); the forces handed back are the
    # ones the last update used (read before EulerUpdate zeroes them).  A dense
    # packing with attraction so that every term contributes.
    p = W.tissue2d(6, 20, 1.4, 0.92, Kb=0.01, Kat=0.5, seed=11)
    P, stride = p.padded()
    P = P.astype(np.float32)
    # a large step, so that F dt is comparable to the coordinates and one
    # rounding (fused) and two roundings differ at many vertices
    dt = np.float32(0.37)
    args = (stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat)

    def forces(X):
        F, ov, at = oracle.forces2d(X, *args, prec=32)
        return F, ov, at

    def fused(X, F):
        return (X.astype(np.float64) + F.astype(np.float64) * np.float64(dt)).astype(np.float32)

    real = np.concatenate([np.arange(c * stride, c * stride + p.nv[c]) for c in range(p.C)])
    F0, ov, at = forces(P)
    assert ov[real].sum() > 0 and at[real].sum() > 0  # repulsion and attraction pairs present
    X1, Fl1 = oracle.run2d(P, *args, 1, dt, prec=32)
    np.testing.assert_array_equal(Fl1[real], F0[real])
    np.testing.assert_array_equal(X1[real], fused(P, F0)[real])
    F1, _, _ = forces(X1)
    assert not np.array_equal(F1[real], F0[real])
    X2, Fl2 = oracle.run2d(P, *args, 2, dt, prec=32)
    np.testing.assert_array_equal(Fl2[real], F1[real])  # forces at the pre-step positions of step 2
    np.testing.assert_array_equal(X2[real], fused(X1, F1)[real])
    # a two-rounding update (x + round(F dt)) differs somewhere: the check above has teeth
    two = (P + (F0 * dt).astype(np.float32)).astype(np.float32)
    assert not np.array_equal(two[real], fused(P, F0)[real])
    X0, Fl0 = oracle.run2d(P, *args, 0, dt, prec=32)
    np.testing.assert_array_equal(X0, P)
    assert not Fl0.any()


def fmix32(h):
    h = np.asarray(h, np.uint32).copy()
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x85EBCA6B)
    h ^= h >> np.uint32(13)
    h *= np.uint32(0xC2B2AE35)
    h ^= h >> np.uint32(16)
    return h


def keys3(P):
    b = np.ascontiguousarray(P, np.float32).view(np.uint32)
    return fmix32(b[:, 0] ^ fmix32(b[:, 1] ^ fmix32(b[:, 2])))


def keys2(P):
    b = np.ascontiguousarray(P, np.float32).view(np.uint32)
    return fmix32(b[:, 0] ^ fmix32(b[:, 1]))


def test_fmix32_is_murmur3_finaliser():
    # MurmurHash3's fmix32 (a bijection) on known values: fmix32(0) = 0 and
    # the published constants' effect on 1
    assert int(fmix32([0])[0]) == 0
    h = 1
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & 0xFFFFFFFF
    h ^= h >> 16
    assert int(fmix32([1])[0]) == h
    x = np.arange(1 << 16, dtype=np.uint32) * np.uint32(65537)
    assert len(np.unique(fmix32(x))) == len(x)


def test_adhesion_key_sums_match_brute_force_pairs():
    # the oracle's adhesion pair sets (own and reverse, as key sums) against an
    # independent float32 brute force of Cell3D_Kernel.cl:341-353: minimum
    # image per component, |d| < 2 l0 of the own cell, |d| > 1e-12; a small
    # periodic box, so that pairs across the boundary take an image
    m = W.monolayer3d(3, 1.75, seed=12)
    X = m.X.astype(np.float32)
    nv, L = m.mesh.nv, np.float32(m.L)
    r = oracle.forces3d(X, m.mesh.faces, nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=0.1, Kb=m.Kb,
                        mask=16, prec=32, pairs=True)
    l0 = (np.sqrt(np.float32(4) * np.asarray(m.a0, np.float32)) / np.sqrt(np.float32(3))).astype(np.float32)
    cell = np.repeat(np.arange(m.C), nv)
    key = keys3(X)
    own = np.zeros(len(X), np.uint32)
    rev = np.zeros(len(X), np.uint32)
    cnt = np.zeros(len(X), np.int32)
    crossing = 0
    for a in range(len(X)):
        d = X - X[a]
        img = np.round(d / L).astype(np.float32)
        crossing += int(np.any(img[cell != cell[a]] != 0))
        d = (d - L * img).astype(np.float32)
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        lim = np.float32(2) * l0[cell[a]]
        acc = (cell != cell[a]) & (dist < lim) & (dist > np.float32(1e-12))
        cnt[a] = int(acc.sum())
        own[a] = np.sum(key[acc], dtype=np.uint32)
        np.add.at(rev, np.nonzero(acc)[0], key[a])
    assert cnt.sum() > 0 and crossing > 0
    np.testing.assert_array_equal(r["adhesion"], cnt)
    np.testing.assert_array_equal(r["adhesion_hash"][:, 0], own)
    np.testing.assert_array_equal(r["adhesion_hash"][:, 1], rev)
    # a partner dropped from one vertex's set changes its key sum
    assert np.uint32(own[np.argmax(cnt)] - key[0]) != own[np.argmax(cnt)]


def test_attraction_key_sums_match_brute_force_pairs():
    # Cell2D_kernel.cl:250-266: minimum image, d < l0 of the own cell
    p = W.tissue2d(8, 20, 1.4, 0.92, Kb=0.01, Kat=0.5, seed=13)
    P, stride = p.padded()
    P = P.astype(np.float32)
    F, ov, at, ah = oracle.forces2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat,
                                    prec=32, attr_hash=True)
    real = np.concatenate([np.arange(c * stride, c * stride + p.nv[c]) for c in range(p.C)])
    X = P.reshape(-1, 2)[real]
    cell = np.repeat(np.arange(p.C), p.nv)
    L = np.float32(p.L)
    key = keys2(X)
    l0 = np.asarray(p.l0, np.float32)
    cnt = np.zeros(len(X), np.int32)
    own = np.zeros(len(X), np.uint32)
    for a in range(len(X)):
        d = X - X[a]
        d = (d - L * np.round(d / L).astype(np.float32)).astype(np.float32)
        dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
        acc = (cell != cell[a]) & (dist < l0[cell[a]])
        cnt[a] = int(acc.sum())
        own[a] = np.sum(key[acc], dtype=np.uint32)
    assert cnt.sum() > 0
    np.testing.assert_array_equal(at[real], cnt)
    np.testing.assert_array_equal(ah[real], own)
