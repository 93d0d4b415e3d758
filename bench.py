#!/usr/bin/env python3
"""Throughput of the per-particle SPH step on one MI355X.

Workload: Evrard collapse, N = 1,048,576, seed 0 (R = M = G = 1, u = 0.05,
N_ngb = 64, Monaghan AV alpha = 1, beta = 2, eps = h, direct-sum gravity),
fixed dt = 1e-4.  A step is one KDK step of the whole state resident in HBM:
half kick, drift, Morton sort, h iteration, density + EOS, gravity, div/curl,
hydro acceleration + du/dt, half kick, u update.

Prints one JSON line: mixed-precision particle-steps/s (value), the plain
FP64 GPU step's throughput on the same inputs and the speedup, the gravity
kernel's roofline (HIP events around its launches, 20 flop per pair), and
the FP64 CPU oracle timed on a sample of rows of the same step.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector (packed) peak
PEAK_HBM_GBS = 8000.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_048_576)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed steps of the plain FP64 GPU step (0: skip)")
    ap.add_argument("--cpu-rows", type=int, default=8192,
                    help="rows of the step timed on the FP64 CPU oracle (0: skip)")
    ap.add_argument("--grav", default=None, help="gravity config ipl,tj,js (default: autotuned)")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2

    import torch

    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 1

    n = a.n
    e = ic.evrard(n, seed=a.seed)
    print(f"[bench] IC Evrard N={n} seed={a.seed}", file=sys.stderr, flush=True)

    # ---------------- mixed precision, resident -----------------------------
    pl = SphPlan(n, mode="mixed")
    if a.grav:
        ipl, tj, js = (int(x) for x in a.grav.split(","))
        pl.set_gravity_config(ipl, tj, js)
    else:
        from sph_2_amd.autotune import choose_gravity_config
        cfg = choose_gravity_config(n)
        pl.set_gravity_config(*cfg)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(a.warmup, a.dt)
    torch.cuda.synchronize()
    print("[bench] mixed warmup done", file=sys.stderr, flush=True)
    pl.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl.step(a.steps, a.dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / a.steps
    value = n * a.steps / wall
    phases = pl.timing()
    counters = pl.counters()
    print(f"[bench] mixed {ms_per_step:.1f} ms/step", file=sys.stderr, flush=True)

    g_ms, g_cnt = phases["gravity"]
    grav_s = g_ms / max(g_cnt, 1) / 1e3
    flops = 20.0 * n * (n - 1)
    achieved = flops / grav_s / 1e12 if grav_s > 0 else None
    roofline = {"bound": "valu", "kernel": "k_grav (direct-sum gravity, FP32)",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
                "traffic": None,
                "algorithmic": "20 flop x N(N-1) pairs per launch",
                "avg_ms": 1e3 * grav_s}

    state = pl.state(("pos", "vel", "u", "h", "rho", "acc_g"))
    pl.close()
    del pl
    torch.cuda.empty_cache()

    # ---------------- plain FP64 GPU step on the same inputs ------------------
    fp64 = None
    if a.fp64_steps > 0:
        pf = SphPlan(n, mode="fp64")
        pf.load(e.pos, e.vel, e.u, e.m, e.h)
        pf.init()
        torch.cuda.synchronize()
        print("[bench] fp64 init done", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        pf.step(a.fp64_steps, a.dt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fp64 = {"value": n * a.fp64_steps / (t1 - t0),
                "ms_per_step": 1e3 * (t1 - t0) / a.fp64_steps, "steps": a.fp64_steps}
        pf.close()
        del pf
        torch.cuda.empty_cache()
        print(f"[bench] fp64 {fp64['ms_per_step']:.1f} ms/step", file=sys.stderr, flush=True)

    # ---------------- FP64 CPU oracle on a sample of rows --------------------
    cpu = None
    parity = {}
    if a.cpu_rows > 0:
        import oracle

        rows = np.arange(min(a.cpu_rows, n))
        p = oracle.Params()
        P, c = oracle.eos(state["rho"], state["u"], p.gamma)
        t0 = time.perf_counter()
        oracle.smoothing(state["pos"], state["h"], rows=rows, p=p)
        rho_o, _ = oracle.density(state["pos"], e.m, state["h"], rows=rows)
        g_o = oracle.gravity(state["pos"], e.m, state["h"], 1.0, rows=rows)
        dv, cv = oracle.divcurl(state["pos"], state["vel"], state["rho"], e.m, state["h"],
                                rows=rows)
        oracle.hydro(state["pos"], state["vel"], state["rho"], P, c, state["h"], e.m, p=p,
                     rows=rows)
        t1 = time.perf_counter()
        cores = len(os.sched_getaffinity(0))
        threads = int(os.environ.get("OMP_NUM_THREADS", cores))
        cpu = {"value": len(rows) / (t1 - t0), "unit": "particle-steps/s",
               "cores": min(threads, cores), "kind": "port",
               "sample": f"one step's forces (h iteration, density, gravity, div/curl, hydro) "
                         f"for {len(rows)} of {n} particles against all {n}, FP64 C/OpenMP "
                         f"oracle, {(t1 - t0):.1f} s on {cpu_model()}"}
        parity = {
            "rows": int(len(rows)),
            "density_rel_max": float(np.max(np.abs(state["rho"][rows] / rho_o - 1.0))),
            "gravity_rel_rms": float(np.sqrt(((state["acc_g"][rows] - g_o) ** 2).sum() /
                                             (g_o ** 2).sum())),
        }
        print(f"[bench] cpu oracle {cpu['value']:.1f} particle-steps/s", file=sys.stderr,
              flush=True)

    out = {
        "metric": "particle-steps/s at N=1M (mixed vs FP64 plain); p50/p99 latency batched runs",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair sums + f64 integrator (mixed)",
        "data": "synthetic",
        "config": {"workload": f"Evrard collapse N={n} seed {a.seed}, one KDK step at dt={a.dt}, "
                               f"mixed precision, direct-sum gravity", "n": n, "dt": a.dt,
                   "parallelism": "single GPU"},
        "fp64_plain": fp64,
        "speedup_vs_fp64_plain": (value / fp64["value"]) if fp64 else None,
        "phases_ms_per_step": {k: (v[0] / max(v[1], 1)) for k, v in phases.items()},
        "counters": counters,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_fp64_oracle": parity,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
