#!/usr/bin/env python3
"""Throughput of the DPM timestep on one MI355X: the 3D monolayer.

Workload (one "step" = one overdamped Euler step of every vertex):
  3D monolayer of 256 cells x 162 vertices (icosphere subdivision 2,
  320 triangles) with radius U(0.9, 1.1) on a 16x16 jittered grid on the
  substrate plane, periodic box L = 38.4; volume, surface-area (edge spring)
  and bending springs Kv=1, Ka=0.5, Kb=0.01; winding-number contact Kre=1;
  vertex adhesion Kat=0.1; substrate Ks=0.2; dt=0.002.  Forces in float,
  per-cell volume / centre-of-mass sums in double.
  * plain   — one tissue per handle (the reference's Tissue3D);
  * batched — `--batch` independent tissues of that shape (seeds 0..B-1)
              advanced together by the same launches.
`value` is the batched whole-job rate in vertex-steps/s with all inputs
resident in HBM; the plain rate is reported beside it.

roofline: the only kernel of a batched step is k3d_wstep_f32_v162 (one wave
per cell; the wave-form step compiled for 162-vertex, 320-face cells; the
plain 256-cell tissue runs k3d_step_f32_v162, one workgroup per cell, which fills
the device better at that size).  Algorithmic
bytes per vertex-step = 24 (position read + write, float xyz) + 24/162
(per-cell state amortised) = 24.15 B; `achieved` = that x vertices per
launch / the average launch duration: a step is one launch of that kernel
and nothing else, launched back to back, so the timed region's per-step time
(HIP events on the library's own stream) is the launch duration with the
launch gaps included (an upper bound on it).  `traffic` = measured HBM bytes
per launch of the steps that do not write the forces (the last step of a
call also writes them) from the committed rocprofv3 summary, when that
summary was taken on the same form and cell-list choice.

cpu_baseline: the reference algorithm restated (oracle ref32: all-pairs,
serial float, the four terms the reference launches), OpenMP over cells on
the host cores, one step of the same 256-cell tissue.

reference_2d: the reference's CPU-runnable 2D case (16 polygons x 24
vertices, 1000 Euler steps) through the 2D step, with the oracle's serial
float run of the same case on one host core beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vertex-steps/s + HBM GB/s, 3D monolayer; batched FP32 vs plain at 1e-5 err"
B_ALG = 24.0 + 24.0 / 162.0   # bytes per vertex-step
HBM_PEAK = 8000.0             # GB/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_tissue(capi, ms):
    cb = np.concatenate([[0], np.cumsum([m.C for m in ms])])
    cat = lambda f: np.concatenate([np.broadcast_to(np.asarray(f(m), np.float32), (m.C,)) for m in ms])
    m0 = ms[0]
    t = capi.Tissues3D(m0.mesh.faces, cb, cat(lambda m: m.Kv), cat(lambda m: m.Ka), cat(lambda m: m.Ks),
                       cat(lambda m: m.v0), cat(lambda m: m.a0), [m.L for m in ms], [m.Kre for m in ms],
                       Kat=[m.Kat for m in ms], Kb=cat(lambda m: m.Kb), mask=capi.DPM3D_ALL, nv=m0.mesh.nv)
    t.set_positions(np.concatenate([m.X for m in ms]))
    return t


def time_steps(t, hiptime, steps, warmup, dt):
    t.euler_async(max(1, warmup), dt)  # also captures the step graph for this dt
    t.synchronize()
    t.counters(reset=True)
    e0, e1 = hiptime.Event(), hiptime.Event()
    s = t.stream()
    wall0 = time.perf_counter()
    e0.record(s)
    t.euler_async(steps, dt)
    e1.record(s)
    ms = e0.elapsed_ms(e1)
    wall = time.perf_counter() - wall0
    return ms, wall, t.counters()


def host_cores():
    """Physical cores among the CPUs this process may run on (one thread per
    core), capped by OMP_NUM_THREADS when it is set; and the CPU model."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            pkg = open(base + "physical_package_id").read().strip()
            core = open(base + "core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    n = max(1, len(cores))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, len(cpus), model


def cpu_baseline(m, steps):
    """The oracle's float restatement of the reference algorithm, OpenMP over
    cells with one thread per physical core, threads bound close
    (OMP_PROC_BIND=close), set before the OpenMP runtime loads."""
    threads, ncpus, model = host_cores()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    import oracle
    t0 = time.perf_counter()
    oracle.run3d(m.X, m.mesh.faces, m.mesh.nv, m.Kv, m.Ka, m.Ks, m.v0, m.a0, m.L, m.Kre, steps, m.dt,
                 mask=15, prec=32, lib="liboracle_omp.so")
    sec = time.perf_counter() - t0
    nvt = m.C * m.mesh.nv
    return {"value": nvt * steps / sec, "unit": "vertex-steps/s", "cores": threads, "kind": "port",
            "cpu": f"{model}; {ncpus} CPUs available, {threads} physical cores used, OMP_PROC_BIND=close",
            "sample": f"{steps} step(s) of the 256-cell x 162-vertex monolayer, the reference's four "
                      f"terms (volume, surface, substrate, all-pairs winding contact), float, "
                      f"OpenMP over cells: {sec:.2f} s"}


def plain2d_gpu(capi, hiptime, W, steps=1000, dt=0.005):
    """The reference's CPU-runnable 2D case (16 bidisperse 24-vertex polygons at packing
    fraction 0.6, Ka = Kl = 1, Kb = 0, vertex repulsion, dt = 0.005, 1000 Euler steps),
    through the 2D step (one launch for all steps), timed with HIP events."""
    p = W.tissue2d(16, 24, 1.4, 0.6, seed=0)
    t = capi.Tissues2D([0, p.C], p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, Kre=p.Kre, Kat=p.Kat,
                       mask=capi.DPM2D_ALL)
    t.set_positions(p.X)
    t.euler(steps, dt)  # warm-up
    t.set_positions(p.X)
    e0, e1 = hiptime.Event(), hiptime.Event()
    s = t.stream()
    e0.record(s)
    t.euler_async(steps, dt)
    e1.record(s)
    ms = e0.elapsed_ms(e1)
    nvt = int(sum(p.nv))
    return p, {"workload": "2D, 16 polygons x 24 vertices, packing fraction 0.6, 1000 Euler steps, dt=0.005",
               "steps": steps, "ms": ms, "value": nvt * steps / (ms / 1e3), "unit": "vertex-steps/s"}


def driver2d_gpu(capi, hiptime, steps=1000, dt=0.005):
    """The reference's test2D.py tissue (test2D.py:9-31): 200 pairs of
    Cell2D(0, 0, 1.2, 25, 1.0) and Cell2D(0, 0, 1.2, 22, 1.3), Ka 0.1, Kl 1,
    Kb 0.05, Tissue2D(cells, 0.9), Kre 1, Kat 0.5, Disperse (on the device,
    drand48 seeded as an unseeded reference run), then Euler steps of
    dt = 0.005: 400 cells, 9,400 vertices.  Both 2D step forms, timed with HIP
    events after a warm-up run from the same start."""
    from opencl_dpm_amd import clDPM
    c, c2 = clDPM.Cell2D(0, 0, 1.2, 25, 1.0), clDPM.Cell2D(0, 0, 1.2, 22, 1.3)
    for x in (c, c2):
        x.Ka, x.Kl, x.Kb = 0.1, 1.0, 0.05
    T = clDPM.Tissue2D([c, c2] * 200, 0.9)
    T.Kre, T.Kat = 1.0, 0.5
    clDPM.seed48(0x330E, 0xABCD, 0x1234)
    T.Disperse()
    X0 = np.concatenate([np.asarray(x.Verts, np.float32) for x in T.Cells])
    nvs = np.array([x.NV for x in T.Cells], np.int32)
    a0 = np.array([x.a0 for x in T.Cells], np.float32)
    l0 = np.array([x.l0 for x in T.Cells], np.float32)
    r0 = np.array([x.r0 for x in T.Cells], np.float32)
    out = {"workload": "test2D.py's tissue: 400 cells (25 and 22 vertices), packing 0.9, dispersed, "
                       f"{steps} Euler steps, dt={dt}", "steps": steps, "vertices": int(nvs.sum())}
    for kind, name in ((1, "one_workgroup"), (2, "wide")):
        t = capi.Tissues2D([0, len(nvs)], nvs, 0.1, 1.0, 0.05, a0, l0, r0, T.L, Kre=1.0, Kat=0.5)
        t.set_step_kernel(kind)
        t.set_positions(X0)
        t.euler(steps, dt)  # warm-up (captures the wide form's step graph)
        t.set_positions(X0)
        e0, e1 = hiptime.Event(), hiptime.Event()
        s = t.stream()
        e0.record(s)
        t.euler_async(steps, dt)
        e1.record(s)
        ms = e0.elapsed_ms(e1)
        out[name] = {"ms": ms, "us_per_step": ms * 1e3 / steps,
                     "value": float(nvs.sum()) * steps / (ms / 1e3), "unit": "vertex-steps/s"}
        del t
    return out


def plain2d_cpu(p, steps=1000, dt=0.005):
    """The same 2D run by the oracle's float restatement of the reference kernels,
    serial on one host core."""
    import oracle
    P, stride = p.padded()
    t0 = time.perf_counter()
    oracle.run2d(P, stride, p.nv, p.Ka, p.Kl, p.Kb, p.a0, p.l0, p.r0, p.L, p.Kre, p.Kat, steps, dt, prec=32)
    sec = time.perf_counter() - t0
    nvt = int(sum(p.nv))
    return {"value": nvt * steps / sec, "unit": "vertex-steps/s", "cores": 1, "kind": "port", "seconds": sec}


def measured_traffic_per_vertex_step(config: str):
    """HBM bytes per vertex-step of the batched step kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary, used only when that summary was taken on the same
    configuration (its `traffic_config:` line equals `config`); None otherwise."""
    path = os.path.join(ROOT, "profiles", "k3d_wstep_monolayer_64x256x162.md")
    value, cfg = None, None
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_vertex_step:"):
                value = float(line.split(":", 1)[1])
            elif line.startswith("traffic_config:"):
                cfg = line.split(":", 1)[1].strip()
    except OSError:
        return None
    return value if cfg == config else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="steps timed on the plain tissue (default: --steps, the batched window)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the batched handle's autotuner (dpm3d_autotune) and use --step-kernel")
    ap.add_argument("--step-kernel", type=int, default=None,
                    help="1 one workgroup per cell, 2 one wave per cell (default: the library's "
                         "automatic choice, which is 1 for the plain tissue and 2 for the batch)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 5 times (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copied 5 x 256 MiB")
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from opencl_dpm_amd import capi, hiptime
    from opencl_dpm_amd import workloads as W

    if capi.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    dt = 0.002
    m = W.monolayer_config("cfg2", seed=0)
    nvt = m.C * m.mesh.nv

    if args.plain_steps is None:
        args.plain_steps = args.steps  # the same window of the trajectory as the batched rate
    log(f"[bench] plain: {m.C} cells x {m.mesh.nv} vertices, {args.plain_steps} steps")
    plain = build_tissue(capi, [m])
    plain.set_step_kernel(args.step_kernel or 0)
    p_ms, p_wall, p_cnt = time_steps(plain, hiptime, args.plain_steps, args.warmup, dt)
    plain_rate = nvt * args.plain_steps / (p_ms / 1e3)
    log(f"[bench] plain {plain_rate:.4g} vertex-steps/s ({p_ms / args.plain_steps * 1e3:.1f} us/step)")
    del plain

    B = args.batch
    log(f"[bench] batched: {B} tissues, {args.steps} steps")
    ms = [W.monolayer_config("cfg2", seed=s) for s in range(B)]
    bt = build_tissue(capi, ms)
    b_kind = args.step_kernel or 2  # the automatic choice at 64 x 256 cells
    bt.set_step_kernel(b_kind)
    tuned = None
    best = {"form": b_kind, "cells_per_workgroup": 0, "cell_list": -1}
    if not args.no_autotune and args.step_kernel is None:
        # the autotuner on the start state (it restores it), then the
        # warm-up and the timed steps
        best, times = bt.autotune(steps=5, dt=dt)
        b_kind = best["form"]
        tuned = {"choice": best, "ms_per_step": times}
        log(f"[bench] autotune: {best}")
    b_ms, b_wall, b_cnt = time_steps(bt, hiptime, args.steps, args.warmup, dt)
    b_launch = b_ms / args.steps  # one launch per step (see the module docstring)
    nvb = nvt * B
    rate = nvb * args.steps / (b_ms / 1e3)
    log(f"[bench] batched {rate:.4g} vertex-steps/s ({b_launch:.3f} ms/step)")
    achieved = B_ALG * nvb / (b_launch / 1e3) / 1e9
    windings = float(b_cnt[1]) / max(1, int(b_cnt[0]))
    # the measured traffic per vertex-step does not depend on the cells per
    # workgroup (within 0.2 % in the profile); the key names what it does
    config_key = f"batch={B} form={best['form']} cell_list={best['cell_list']}"
    tpv = measured_traffic_per_vertex_step(config_key)
    traffic = None if tpv is None else tpv * nvb

    out = {
        "metric": METRIC,
        "value": rate,
        "unit": "vertex-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D monolayer, 256 cells x 162 vertices (icosphere s=2), batch of {B} "
                        f"independent tissues, all force terms, dt=0.002",
            "tissues": B, "cells_per_tissue": m.C, "verts_per_cell": m.mesh.nv,
            "vertices": nvb, "per_cell_sums": "f64",
        },
        "plain": {"value": plain_rate, "unit": "vertex-steps/s", "steps": args.plain_steps,
                  "ms_per_step": p_ms / args.plain_steps},
        "batched_vs_plain": rate / plain_rate,
        "autotune": tuned,
        "hbm_gbps_algorithmic": B_ALG * rate / 1e9,
        "winding_evals_per_step": windings,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k3d_wstep_f32_v162" if b_kind == 2 else "k3d_step_f32_v162", "launch_ms": b_launch,
                     "bytes_per_vertex_step": B_ALG, "config": config_key},
    }
    p2, out["reference_2d"] = plain2d_gpu(capi, hiptime, W)
    log(f"[bench] 2D reference case: {out['reference_2d']['ms']:.2f} ms for 1000 steps "
        f"({out['reference_2d']['value']:.4g} vertex-steps/s)")
    out["driver_2d"] = driver2d_gpu(capi, hiptime)
    log(f"[bench] test2D.py tissue (400 cells): one workgroup {out['driver_2d']['one_workgroup']['us_per_step']:.1f} "
        f"us/step, wide {out['driver_2d']['wide']['us_per_step']:.1f} us/step")
    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_steps} step(s)")
        out["cpu_baseline"] = cpu_baseline(m, args.cpu_steps)
        log(f"[bench] cpu {out['cpu_baseline']['value']:.4g} vertex-steps/s")
        out["reference_2d"]["cpu"] = plain2d_cpu(p2)
        log(f"[bench] 2D reference case on the cpu: {out['reference_2d']['cpu']['seconds']:.2f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
