"""Measurement studies of the 3D step on one GPU.

  python -m opencl_dpm_amd.studies precision [--steps 5000] [--every 500]
  python -m opencl_dpm_amd.studies batch [--n 512] [--fire-iters 200] [--jam]
  python -m opencl_dpm_amd.studies adjoint [--steps 200] [--every 20] [--fd-cells 8]

precision: the float production step (float positions and forces, double
per-cell sums) against the all-double step (dpm3d_create_f64) from the
same seeded start, on the 4096-cell x 642-vertex monolayer (icosphere
subdivision 3, radius 1 +- 15 %, 64 x 64 grid, spacing 2.4, L = 153.6,
Kv=1, Ka=0.5, Kb=0.01, Ks=0.2, Kre=1, Kat=0.1, dt=0.002, all force terms).
Reported every `--every` steps:
  * force error at the float run's current positions:
    ||F32 - F64||_2 / ||F64||_2, and per vertex
    e_v = |F32_v - F64_v| / max(|F64_v|, 1e-6 RMS(F64)) (its median, 99.99th
    percentile, maximum and the count above 1e-3), with F64 the double
    step's forces at the same positions;
  * the decisions of the reference's cut-offs that the two evaluations take
    differently at those positions: facing-down faces (n.z < -0.1) and
    vertices whose adhesion partner count (d < 2 l0) differs — these force
    laws are discontinuous at the cut-off, so one flipped decision moves a
    vertex's force by O(Ks) or O(Kat), far above rounding;
  * energy drift |U32 - U64| / |U64| of the two trajectories, U the elastic
    energy (volume, edge springs, bending) summed over cells;
  * the rate of each step in vertex-steps/s.
One JSON line at the end summarises the run.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import capi, hiptime
from . import workloads as W


def _tissue(m, precision, mask=capi.DPM3D_ALL):
    return capi.Tissues3D(m.mesh.faces, [0, m.C], m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, Kat=m.Kat,
                          Kb=m.Kb, ks_grad=m.ks_grad, mask=mask, nv=m.mesh.nv, precision=precision)


def _timed(t, n, dt):
    e0, e1 = hiptime.Event(), hiptime.Event()
    s = t.stream()
    e0.record(s)
    t.euler_async(n, dt)
    e1.record(s)
    ms = e0.elapsed_ms(e1)
    t.synchronize()
    return ms


def force_error(F32, F64):
    F32 = np.asarray(F32, np.float64)
    d = np.linalg.norm(F32 - F64, axis=1)
    n64 = np.linalg.norm(F64, axis=1)
    rms = float(np.sqrt(np.mean(n64 ** 2)))
    ev = d / np.maximum(n64, 1e-6 * rms)
    return {"force_rel_l2": float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64)),
            "vertex_median": float(np.median(ev)), "vertex_p9999": float(np.quantile(ev, 0.9999)),
            "vertex_max": float(ev.max()), "vertices_above_1e-3": int((ev > 1e-3).sum())}


MASKS = {"all": capi.DPM3D_ALL, "reference": capi.DPM3D_REFERENCE,
         "smooth": capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_BENDING}


def precision(steps: int, every: int, dt: float = 0.002, seed: int = 0, config: str = "cfg3",
              mask: str = "all"):
    m = W.monolayer_config(config, seed=seed)
    nvt = m.C * m.mesh.nv
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    log(f"[precision] {m.C} cells x {m.mesh.nv} vertices = {nvt} vertices, {steps} steps, dt={dt}, "
        f"terms {mask}")
    t32, t64, probe64, t32c = [_tissue(m, p, MASKS[mask]) for p in (32, 64, 64, 32)]
    t32c.set_compensated(True)
    t32.set_positions(m.X)
    t32c.set_positions(m.X)
    t64.set_positions(m.X.astype(np.float64))
    U0 = float(t64.energy().sum())
    rows, ms32, ms64, ms32c = [], 0.0, 0.0, 0.0
    done = 0
    while True:
        probe64.set_positions(t32.positions(np.float64))
        g32 = t32.probe(contact=False)
        g64 = probe64.probe(contact=False)
        r = force_error(g32["forces"], g64["forces"])
        r["flipped_down_faces"] = int((g32["down"] != g64["down"]).sum())
        r["flipped_adhesion_vertices"] = int((g32["adhesion"] != g64["adhesion"]).sum())
        U32, U64 = float(t32.energy().sum()), float(t64.energy().sum())
        U32c = float(t32c.energy().sum())
        X64 = t64.positions()
        r.update(step=done, U32=U32, U64=U64, drift=abs(U32 - U64) / abs(U64), U32c=U32c,
                 drift_compensated=abs(U32c - U64) / abs(U64),
                 dx_plain=float(np.abs(t32.positions(np.float64) - X64).max()),
                 dx_compensated=float(np.abs(t32c.positions(np.float64) - X64).max()))
        rows.append(r)
        log(f"[precision] step {done:5d}: force rel {r['force_rel_l2']:.3e}; per vertex median "
            f"{r['vertex_median']:.2e} p99.99 {r['vertex_p9999']:.2e} max {r['vertex_max']:.2e} "
            f"(>1e-3: {r['vertices_above_1e-3']}); flips: down {r['flipped_down_faces']} "
            f"adhesion {r['flipped_adhesion_vertices']};  U32 {U32:.10e} U64 {U64:.10e} "
            f"drift {r['drift']:.3e} (compensated {r['drift_compensated']:.3e}); max|x32 - x64| "
            f"{r['dx_plain']:.2e} (compensated {r['dx_compensated']:.2e})")
        if done >= steps:
            break
        n = min(every, steps - done)
        ms32 += _timed(t32, n, dt)
        ms32c += _timed(t32c, n, dt)
        ms64 += _timed(t64, n, dt)
        done += n
    out = {
        "study": "precision",
        "workload": f"3D monolayer {m.C} cells x {m.mesh.nv} vertices, terms {mask}, dt={dt}",
        "steps": steps,
        "U0": U0,
        "terms": mask,
        "force_rel_l2_max": max(r["force_rel_l2"] for r in rows),
        "force_rel_l2_median": float(np.median([r["force_rel_l2"] for r in rows])),
        "checkpoints_within_1e-5": sum(r["force_rel_l2"] <= 1e-5 for r in rows),
        "checkpoints": len(rows),
        "checkpoints_without_flips": sum(r["flipped_down_faces"] + r["flipped_adhesion_vertices"] == 0
                                         for r in rows),
        "energy_drift_final": rows[-1]["drift"],
        "energy_drift_max": max(r["drift"] for r in rows),
        "energy_drift_compensated_final": rows[-1]["drift_compensated"],
        "energy_drift_compensated_max": max(r["drift_compensated"] for r in rows),
        "f32_compensated_ms_per_step": ms32c / steps,
        "f32_vertex_steps_per_s": nvt * steps / (ms32 / 1e3),
        "f64_vertex_steps_per_s": nvt * steps / (ms64 / 1e3),
        "f32_ms_per_step": ms32 / steps,
        "f64_ms_per_step": ms64 / steps,
        "rows": rows,
    }
    print(json.dumps(out), flush=True)
    return out


def _handle2d(ps, mask, precision):
    cb = np.concatenate([[0], np.cumsum([p.C for p in ps])])
    cat = lambda f: np.concatenate([np.asarray(f(p)) for p in ps])
    t = capi.Tissues2D(cb, cat(lambda p: p.nv), cat(lambda p: p.Ka), cat(lambda p: p.Kl), cat(lambda p: p.Kb),
                       cat(lambda p: p.a0), cat(lambda p: p.l0), cat(lambda p: p.r0), [p.L for p in ps],
                       Kre=[p.Kre for p in ps], Kat=[p.Kat for p in ps], mask=mask, precision=precision)
    x = cat(lambda p: p.X)
    t.set_positions(x.astype(np.float64) if precision == 64 else x)
    return t, x


def _timed_call(t, fn):
    e0, e1 = hiptime.Event(), hiptime.Event()
    s = t.stream()
    e0.record(s)
    out = fn()
    e1.record(s)
    return out, e0.elapsed_ms(e1)


def batch(n: int, fire_iters: int, jam: bool, max_iter: int, seed: int = 0, jam_stages: int = 2):
    """Batching form: n 2D packings (cells U{8..128}, vertices per cell
    U{16..64}, bidisperse 1:1.4, Kb alternating 0 / 0.01, calA 1.02) in one
    handle: workgroup size autotuned; FIRE throughput (fixed iterations) in
    float and double; Euler throughput of the reference terms (float); then
    the jamming protocol of every packing, run as launches of at most
    `jam_stages` minimisations per packing (bit-identical to the one-launch
    dpm2d_jam) so that progress is reported between launches."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    ps = W.packings2d(n, seed=seed)
    nverts = sum(len(p.X) for p in ps)
    log(f"[batch] {n} packings, {sum(p.C for p in ps)} cells, {nverts} vertices "
        f"(largest {max(len(p.X) for p in ps)})")
    out = {"study": "batch", "packings": n, "vertices": nverts, "cells": int(sum(p.C for p in ps))}
    for prec in (32, 64):
        t, x = _handle2d(ps, capi.DPM2D_CONSERVATIVE, prec)
        best, times = t.autotune(iters=20)
        log(f"[batch] f{prec} autotune: {best} threads {times}")
        st, ms = _timed_call(t, lambda: t.fire(max_iter=fire_iters, ftol=0.0))
        rate = nverts * fire_iters / (ms / 1e3)
        log(f"[batch] f{prec} FIRE {fire_iters} iterations: {ms:.1f} ms, {rate:.4g} vertex-iterations/s")
        out[f"fire_f{prec}"] = {"threads": best, "autotune_ms": times, "iterations": fire_iters, "ms": ms,
                                "vertex_iterations_per_s": rate}
        del t
    te, _ = _handle2d(ps, capi.DPM2D_ALL, 32)
    _, ms = _timed_call(te, lambda: te.euler_async(200, 0.005))
    te.synchronize()
    out["euler_f32_reference_terms"] = {"steps": 200, "ms": ms, "vertex_steps_per_s": nverts * 200 / (ms / 1e3)}
    log(f"[batch] f32 Euler (reference terms) 200 steps: {ms:.1f} ms, {nverts * 200 / (ms / 1e3):.4g} vertex-steps/s")
    del te
    if jam:
        t, x = _handle2d(ps, capi.DPM2D_CONSERVATIVE, 64)
        best, _ = t.autotune(iters=20)
        log(f"[batch] jamming {n} packings (f64, {best} threads, FIRE max_iter {max_iter}), launches of at most "
            f"{jam_stages} minimisations per packing")
        kw = dict(fire=dict(max_iter=max_iter))
        total_ms, launches = 0.0, 0
        st = None
        while st is None or (st["status"] == 2).any():
            st, ms = _timed_call(t, lambda: t.jam_step(st is None, jam_stages, **kw))
            total_ms += ms
            launches += 1
            log(f"[batch] jamming launch {launches}: {ms / 1e3:.1f} s, {int((st['status'] == 2).sum())} packings "
                f"unfinished, {int(st['stages'].sum())} minimisations so far")
        phi = st["phi_hi"]
        ok = st["status"] == 0
        log(f"[batch] jamming: {total_ms / 1e3:.1f} s in {launches} launches; phi_J mean {phi[ok].mean():.4f} "
            f"sd {phi[ok].std():.4f} (Kb=0: {phi[ok][0::2].mean() if ok[0::2].any() else float('nan'):.4f}, "
            f"Kb=0.01: {phi[1::2][ok[1::2]].mean() if ok[1::2].any() else float('nan'):.4f}); status "
            f"{np.bincount(st['status'].astype(int), minlength=4).tolist()}; stages mean {st['stages'].mean():.1f}; "
            f"minimisations that hit max_iter {int(st['unconverged'].sum())} of {int(st['stages'].sum())}")
        out["jam"] = {"ms": total_ms, "launches": launches, "threads": best, "max_iter": max_iter,
                      "phi_j_mean": float(phi[ok].mean()) if ok.any() else None,
                      "phi_j_sd": float(phi[ok].std()) if ok.any() else None,
                      "status_counts": np.bincount(st["status"].astype(int), minlength=4).tolist(),
                      "stages_mean": float(st["stages"].mean()), "fire_iterations": float(st["iterations"].sum()),
                      "unconverged": int(st["unconverged"].sum()), "minimisations": int(st["stages"].sum()),
                      "phi_lo": st["phi_lo"].tolist(), "phi_hi": st["phi_hi"].tolist(),
                      "kb": [float(p.Kb[0]) for p in ps], "cells": [int(p.C) for p in ps]}
    print(json.dumps(out), flush=True)
    return out


ADJ_MASK = (capi.DPM3D_VOLUME | capi.DPM3D_SURFACE | capi.DPM3D_SUBSTRATE | capi.DPM3D_CONTACT |
            capi.DPM3D_BENDING)
ADJ_NAMES = ("Kv", "Ka", "Kb", "Ks")


def _adj_tissue(m, precision, th):
    return capi.Tissues3D(m.mesh.faces, [0, m.C], th["Kv"], th["Ka"], th["Ks"], m.v0, m.a0, m.L, m.Kre, Kat=0.0,
                          Kb=th["Kb"], ks_grad=m.ks_grad, mask=ADJ_MASK, nv=m.mesh.nv, precision=precision)


def _com_loss(m, XN, X0):
    d = XN.reshape(m.C, -1, 3).mean(1) - X0.reshape(m.C, -1, 3).mean(1)
    n = np.linalg.norm(d, axis=1)
    lam = np.repeat(d / n[:, None] / m.C / m.mesh.nv, m.mesh.nv, axis=0)
    return float(n.mean()), lam


def adjoint(nsteps: int, every: int, fd_cells: int, dt: float = 0.002, eps: float = 1e-6):
    """Differentiation form: 64 cells x 162 vertices on a substrate whose
    stiffness grows 10x across x; loss = mean |COM_N - COM_0| after nsteps;
    gradients with respect to per-cell (Kv, Ka, Kb, Ks) by the float and the
    double adjoint (checkpoints every `every` steps), against central
    differences of the double forward (step eps, between float-stored
    values) for the first fd_cells cells."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    m = W.monolayer_config("cfg5")
    C = m.C
    b = lambda v: np.broadcast_to(np.asarray(v, np.float32), (C,)).astype(np.float64)
    th = {"Kv": b(m.Kv), "Ka": b(m.Ka), "Kb": b(m.Kb), "Ks": b(m.Ks)}
    X0 = m.X.astype(np.float64)
    out = {"study": "adjoint", "cells": C, "verts_per_cell": m.mesh.nv, "steps": nsteps, "every": every}
    grads = {}
    for prec in (32, 64, "32c"):
        t = _adj_tissue(m, 64 if prec == 64 else 32, th)
        if prec == "32c":
            t.set_compensated(True)
        t.set_positions(X0 if prec == 64 else m.X)
        t.counters(reset=True)
        (_, ms_f) = _timed_call(t, lambda: t.trajectory(nsteps, dt, every))
        contacts = int(t.counters()[2])
        L, lam = _com_loss(m, t.positions(np.float64), X0)
        (res, ms_b) = _timed_call(t, lambda: t.adjoint(lam))
        grads[prec] = res[0]
        log(f"[adjoint] f{prec}: loss {L:.9e}; forward {ms_f:.1f} ms, adjoint {ms_b:.1f} ms; "
            f"contact pairs along the trajectory {contacts}")
        out[f"f{prec}"] = {"loss": L, "forward_ms": ms_f, "adjoint_ms": ms_b, "contact_pairs": contacts}
    # (1) central differences with step eps; (2) the median of the
    # difference quotients over consecutive float values theta_0 + k ulp,
    # k = -8..8, which ignores the isolated jumps that the force law's
    # cut-offs (substrate height and facing tests, contact onsets) put in the
    # loss; jumps are counted as quotients further than 20 median absolute
    # deviations from the median.
    def loss_at(c, name, value):
        tp = {q: a.copy() for q, a in th.items()}
        tp[name][c] = value
        t = _adj_tissue(m, 64, tp)
        t.set_positions(X0)
        t.euler(nsteps, dt)
        return _com_loss(m, t.positions(np.float64), X0)[0]

    fd = np.zeros((fd_cells, 4))
    med = np.zeros((fd_cells, 4))
    jumps = 0
    for c in range(fd_cells):
        for k, name in enumerate(ADJ_NAMES):
            x0 = np.float32(th[name][c])
            xp, xm = float(np.float32(x0 * (1 + eps))), float(np.float32(x0 * (1 - eps)))
            fd[c, k] = (loss_at(c, name, xp) - loss_at(c, name, xm)) / (xp - xm)
            xs = [float(x0)]
            up, dn = x0, x0
            for _ in range(8):
                up = np.nextafter(up, np.float32(np.inf))
                dn = np.nextafter(dn, np.float32(-np.inf))
                xs = [float(dn)] + xs + [float(up)]
            Ls = [loss_at(c, name, x) for x in xs]
            q = np.diff(Ls) / np.diff(xs)
            mq = float(np.median(q))
            mad = float(np.median(np.abs(q - mq))) + 1e-300
            jumps += int(np.sum(np.abs(q - mq) > 20 * mad))
            med[c, k] = mq
        log(f"[adjoint] differences: {c + 1}/{fd_cells} cells")
    out["jumps_in_ulp_scans"] = jumps
    log(f"[adjoint] quotients flagged as jumps in the 1-ulp scans: {jumps} of {fd.size * 16}")
    for prec in (32, "32c", 64):
        g = grads[prec][:fd_cells]
        for label, ref in (("central differences", fd), ("median 1-ulp quotients", med)):
            err = float(np.linalg.norm(g - ref) / np.linalg.norm(ref))
            per = [float(np.linalg.norm(g[:, k] - ref[:, k]) / np.linalg.norm(ref[:, k])) for k in range(4)]
            log(f"[adjoint] f{prec} adjoint vs FP64 {label}: relative error {err:.3e} (Kv {per[0]:.2e}, "
                f"Ka {per[1]:.2e}, Kb {per[2]:.2e}, Ks {per[3]:.2e}) over {ref.size} parameters")
            key = "central" if label.startswith("central") else "ulp_median"
            out[f"f{prec}"][f"rel_error_{key}"] = err
            out[f"f{prec}"][f"rel_error_{key}_per_kind"] = dict(zip(ADJ_NAMES, per))
    gd = float(np.linalg.norm(grads[32] - grads[64]) / np.linalg.norm(grads[64]))
    gdc = float(np.linalg.norm(grads["32c"] - grads[64]) / np.linalg.norm(grads[64]))
    log(f"[adjoint] f32 vs f64 adjoint: relative difference {gd:.3e}; compensated f32 vs f64 {gdc:.3e}")
    out["f32_vs_f64_adjoint"] = gd
    out["f32c_vs_f64_adjoint"] = gdc
    out["fd_cells"] = fd_cells
    print(json.dumps(out), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="study", required=True)
    p = sub.add_parser("precision")
    p.add_argument("--steps", type=int, default=5000)
    p.add_argument("--every", type=int, default=500)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--config", default="cfg3")
    p.add_argument("--terms", default="all", choices=sorted(MASKS))
    p = sub.add_parser("batch")
    p.add_argument("--n", type=int, default=512)
    p.add_argument("--fire-iters", type=int, default=200)
    p.add_argument("--jam", action="store_true")
    p.add_argument("--max-iter", type=int, default=50000)
    p.add_argument("--jam-stages", type=int, default=2)
    p = sub.add_parser("adjoint")
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--every", type=int, default=20)
    p.add_argument("--fd-cells", type=int, default=8)
    a = ap.parse_args(argv)
    if capi.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    if a.study == "precision":
        precision(a.steps, a.every, seed=a.seed, config=a.config, mask=a.terms)
    elif a.study == "adjoint":
        adjoint(a.steps, a.every, a.fd_cells)
    elif a.study == "batch":
        batch(a.n, a.fire_iters, a.jam, a.max_iter, jam_stages=a.jam_stages)
    return 0


if __name__ == "__main__":
    sys.exit(main())
