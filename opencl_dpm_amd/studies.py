"""Measurement studies of the 3D step on one GPU.

  python -m opencl_dpm_amd.studies precision [--steps 5000] [--every 500]
  python -m opencl_dpm_amd.studies batch [--n 512] [--fire-iters 200] [--jam]

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
    t32, t64, probe64 = [_tissue(m, p, MASKS[mask]) for p in (32, 64, 64)]
    t32.set_positions(m.X)
    t64.set_positions(m.X.astype(np.float64))
    U0 = float(t64.energy().sum())
    rows, ms32, ms64 = [], 0.0, 0.0
    done = 0
    while True:
        probe64.set_positions(t32.positions(np.float64))
        g32 = t32.probe(contact=False)
        g64 = probe64.probe(contact=False)
        r = force_error(g32["forces"], g64["forces"])
        r["flipped_down_faces"] = int((g32["down"] != g64["down"]).sum())
        r["flipped_adhesion_vertices"] = int((g32["adhesion"] != g64["adhesion"]).sum())
        U32, U64 = float(t32.energy().sum()), float(t64.energy().sum())
        r.update(step=done, U32=U32, U64=U64, drift=abs(U32 - U64) / abs(U64))
        rows.append(r)
        log(f"[precision] step {done:5d}: force rel {r['force_rel_l2']:.3e}; per vertex median "
            f"{r['vertex_median']:.2e} p99.99 {r['vertex_p9999']:.2e} max {r['vertex_max']:.2e} "
            f"(>1e-3: {r['vertices_above_1e-3']}); flips: down {r['flipped_down_faces']} "
            f"adhesion {r['flipped_adhesion_vertices']};  U32 {U32:.10e} U64 {U64:.10e} "
            f"drift {r['drift']:.3e}")
        if done >= steps:
            break
        n = min(every, steps - done)
        ms32 += _timed(t32, n, dt)
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
    a = ap.parse_args(argv)
    if capi.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    if a.study == "precision":
        precision(a.steps, a.every, seed=a.seed, config=a.config, mask=a.terms)
    elif a.study == "batch":
        batch(a.n, a.fire_iters, a.jam, a.max_iter, jam_stages=a.jam_stages)
    return 0


if __name__ == "__main__":
    sys.exit(main())
