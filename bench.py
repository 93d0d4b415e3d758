#!/usr/bin/env python3
"""Throughput of the per-particle SPH step on one MI355X.

Workload: Evrard collapse, N = 1,048,576, seed 0 (R = M = G = 1, u = 0.05,
N_ngb = 64, Monaghan AV alpha = 1, beta = 2, eps = h, direct-sum gravity),
fixed dt = 1e-4.  A step is one KDK step of the whole state resident in HBM:
half kick, drift, Morton sort, h iteration, density + EOS, gravity, div/curl,
hydro acceleration + du/dt, half kick, u update.

Prints one JSON line: mixed-precision particle-steps/s (value), the plain
FP64 GPU step's throughput on the same inputs and the speedup, the gravity
kernel's roofline (HIP events around its launches, 20 flop per pair), the
FP64 CPU oracle timed on a sample of rows of the same step with the step's
parity against it on those rows, and the served stream (512 runs, N
log-uniform in [2048, 262144], 200 steps each, Poisson arrivals) with its
p50/p99 completion latency and aggregate particle-steps/s.
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
# HBM bytes per gravity-kernel launch at N = 1,048,576: 2 x FETCH_SIZE + WRITE_SIZE of the
# step's dispatch in two rocprofv3 --pmc runs of `bench.py --workload kernels` (FETCH_SIZE reads
# half of the bytes of that workload's 256 MiB calibration copy, hence the factor 2).
GRAV_TRAFFIC_BYTES_1M = {"k_grav": 11.982e9, "k_grav_sym": 52.163e9}
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


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_rms(a, b) -> float:
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="evrard",
                    choices=["evrard", "c1", "c2", "c3", "bh", "serve", "kernels", "aug", "gadget",
                             "block"])
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed steps of the plain FP64 GPU step (0: skip)")
    ap.add_argument("--cpu-rows", type=int, default=32768,
                    help="rows of the step timed on the FP64 CPU oracle (0: skip)")
    ap.add_argument("--grav", default=None, help="gravity config ipl,tj,js (default: autotuned)")
    ap.add_argument("--graph", type=int, default=0, help="c2/c3: replay steps from hipGraphs")
    ap.add_argument("--serve-graph", type=int, default=1,
                    help="served stream: replay steps from hipGraphs")
    ap.add_argument("--runs", type=int, default=512, help="serve: requests in the stream")
    ap.add_argument("--nmin", type=int, default=2048)
    ap.add_argument("--nmax", type=int, default=262144)
    ap.add_argument("--rate", type=float, default=4.0, help="serve: arrivals per second")
    ap.add_argument("--run-steps", type=int, default=200, help="serve: steps per request")
    ap.add_argument("--hbm-gb", type=float, default=64.0, help="serve: HBM budget")
    ap.add_argument("--eps", type=float, default=None,
                    help="gadget: fixed gravitational softening (default eps = h)")
    ap.add_argument("--bh-theta", type=float, default=0.5,
                    help="also time the step with Barnes-Hut gravity at this opening angle "
                         "(0: skip); reported beside the direct-sum headline")
    ap.add_argument("--serve-gravity", default="4,0.1",
                    help="serve: 'S,share' for the symmetric gravity sum (S block pairs per "
                         "wave, share of the budget for its partials) or 'tiled'")
    ap.add_argument("--serve-runs", type=int, default=512,
                    help="requests in the served stream of the default run (0: skip)")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2

    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 1
    if a.workload != "evrard":
        return {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "bh": bench_bh, "serve": bench_serve,
                "kernels": bench_kernels, "aug": bench_aug, "gadget": bench_gadget,
                "block": bench_block}[a.workload](a)
    if a.n is None:
        a.n = 1_048_576
    return bench_evrard(a)


def bench_evrard(a) -> int:
    import torch

    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan

    n = a.n
    e = ic.evrard(n, seed=a.seed)
    print(f"[bench] IC Evrard N={n} seed={a.seed}", file=sys.stderr, flush=True)

    # ---------------- mixed precision, resident -----------------------------
    pl = SphPlan(n, mode="mixed")
    if a.grav:
        cfg_used = tuple(int(x) for x in a.grav.split(","))
    else:
        from sph_2_amd.autotune import choose_gravity_config
        cfg_used = tuple(choose_gravity_config(n))
    pl.set_gravity_config(*cfg_used)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(a.warmup, a.dt)
    torch.cuda.synchronize()
    log("mixed warmup done")
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

    # the gravity kernel alone (HIP events on the plan's stream around its launch)
    k_ms, k_cnt = phases["gravity_kernel"]
    g_ms, g_cnt = phases["gravity"]
    kern_s = k_ms / max(k_cnt, 1) / 1e3
    flops = 20.0 * n * (n - 1)
    achieved = flops / kern_s / 1e12 if kern_s > 0 else None
    kname = "k_grav_sym" if cfg_used[0] == 8 else "k_grav"
    what = ("symmetric direct sum, near and far block pairs, each pair once for both particles"
            if kname == "k_grav_sym" else "tiled direct sum")
    roofline = {"bound": "valu", "kernel": f"{kname} ({what}, FP32)",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": (achieved / PEAK_FP32_TFLOPS) if achieved else None,
                "traffic": GRAV_TRAFFIC_BYTES_1M[kname] if n == 1 << 20 else None,
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                "algorithmic": "20 flop x N(N-1) pair interactions per launch",
                "avg_ms": 1e3 * kern_s,
                "gravity_phase_ms": g_ms / max(g_cnt, 1),
                "launch_config": list(cfg_used)}
    # the far-pair instruction mix on registers alone: the rate the symmetric
    # sum could reach if its memory, LDS rotation and near pairs cost nothing
    issue = None
    if kname == "k_grav_sym" and kern_s > 0:
        from sph_2_amd import libsx
        ceil = libsx.pair_issue_rate()
        rate = 0.5 * n * (n - 1) / kern_s
        issue = {"pairs_per_s": rate, "ceiling_pairs_per_s": ceil, "frac": rate / ceil,
                 "ceiling": "k_pair_probe: the far-pair instruction mix (8 i x 2 j per lane), "
                            "registers only, 8 waves per SIMD",
                 "pairs": "N(N-1)/2 evaluations per launch, each for both particles"}

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

    # ---------------- the same step with Barnes-Hut gravity ------------------
    bh_step = None
    if a.bh_theta > 0:
        from sph_2_amd import libsx

        pb = SphPlan(n, mode="mixed")
        pb.set_gravity_bh(a.bh_theta)
        pb.load(e.pos, e.vel, e.u, e.m, e.h)
        pb.init()
        pb.step(a.warmup, a.dt)
        torch.cuda.synchronize()
        pb.set_timing(True)
        t0 = time.perf_counter()
        pb.step(a.steps, a.dt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ph = pb.timing()
        sb = pb.state(("pos", "h", "acc_g"))
        pb.close()
        del pb
        torch.cuda.empty_cache()
        g_direct = libsx.getAcc_g_smth(sb["pos"], e.m, 1.0, sb["h"])
        bh_step = {"theta": a.bh_theta, "ms_per_step": 1e3 * (t1 - t0) / a.steps,
                   "particle_steps_per_s": n * a.steps / (t1 - t0),
                   "gravity_ms_per_step": ph["gravity"][0] / max(ph["gravity"][1], 1),
                   "gravity_rel_rms_vs_direct_sum": rel_rms(sb["acc_g"], g_direct)}
        log(f"bh step {bh_step['ms_per_step']:.1f} ms/step, force error "
            f"{bh_step['gravity_rel_rms_vs_direct_sum']:.2e}")

    # ---------------- FP64 CPU oracle on a sample of rows --------------------
    cpu = None
    parity = {}
    if a.cpu_rows > 0:
        import oracle
        from sph_2_amd import libsx

        rows = np.arange(min(a.cpu_rows, n))
        r1 = len(rows)
        p = oracle.Params()
        pos, vel, hh, rho = state["pos"], state["vel"], state["h"], state["rho"]
        P, c = oracle.eos(rho, state["u"], p.gamma)
        t0 = time.perf_counter()
        h_o, k_o, _, _ = oracle.smoothing(pos, hh, rows=rows, p=p)
        rho_o, _ = oracle.density(pos, e.m, hh, rows=rows)
        g_o = oracle.gravity(pos, e.m, hh, 1.0, rows=rows)
        dv_o, cv_o = oracle.divcurl(pos, vel, rho, e.m, hh, rows=rows)
        acc_o, du_o = oracle.hydro(pos, vel, rho, P, c, hh, e.m, p=p, rows=rows)
        t1 = time.perf_counter()
        # the same rows of the same final state through the GPU drop-ins: the h
        # iteration (from 2h), div/curl and the fused hydro pass (Monaghan AV)
        h_g, k_g, _ = libsx.smoothing_h(pos, hh, nbeg=0, nend=r1)
        dv_g, cv_g = libsx.div_curlVel_mpi(0, r1, pos, vel, rho, e.m, hh)
        acc_g, du_g = libsx.getAcc_sph_dU(pos, vel, rho, P, c, hh, e.m, nbeg=0, nend=r1)
        cores = len(os.sched_getaffinity(0))
        threads = int(os.environ.get("OMP_NUM_THREADS", cores))
        cpu = {"value": len(rows) / (t1 - t0), "unit": "particle-steps/s",
               "cores": min(threads, cores), "kind": "port",
               "sample": f"one step's forces (h iteration, density, gravity, div/curl, hydro) "
                         f"for {len(rows)} of {n} particles against all {n}, FP64 C/OpenMP "
                         f"oracle, {(t1 - t0):.1f} s on {cpu_model()}"}
        parity = {
            "rows": int(len(rows)),
            "what": "the step's final state (rho, acc_g from the step itself; h/k, div/curl "
                    "and hydro from the drop-ins on that state) against the FP64 oracle on "
                    "the same inputs",
            "h_k_bit_exact": bool(np.array_equal(h_g, h_o) and np.array_equal(k_g, k_o)),
            "k_mismatch": int(np.sum(k_g != k_o)),
            "h_rel_max": float(np.max(np.abs(h_g / h_o - 1.0))),
            "density_rel_max": float(np.max(np.abs(rho[rows] / rho_o - 1.0))),
            "gravity_rel_rms": rel_rms(state["acc_g"][rows], g_o),
            "divV_rel_rms": rel_rms(dv_g, dv_o),
            "curlV_rel_rms": rel_rms(cv_g, cv_o),
            "hydro_acc_rel_rms": rel_rms(acc_g, acc_o),
            "dudt_rel_rms": rel_rms(du_g, du_o),
        }
        print(f"[bench] cpu oracle {cpu['value']:.1f} particle-steps/s", file=sys.stderr,
              flush=True)

    # ---------------- the served stream: p50/p99 completion latency ---------
    serving = None
    if a.serve_runs > 0:
        serving = serve_stream(a, a.serve_runs)
        log(f"serving p50 {serving['p50_latency_s']:.2f}s p99 {serving['p99_latency_s']:.2f}s, "
            f"{serving['particle_steps_per_s']:.3e} particle-steps/s")

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
        "gravity_issue_ceiling": issue,
        "cpu_baseline": cpu,
        "parity_vs_fp64_oracle": parity,
        "serving": serving,
        "barnes_hut_step": bh_step,
    }
    print(json.dumps(out))
    return 0


def _plan_run(n_steps, e, mode, dt, cooling=None, record=(1,), graph=False):
    """Run a plan; returns ({step: state}, ms per step over the steps)."""
    import torch

    from sph_2_amd.plan import SphPlan

    pl = SphPlan(e.n, mode=mode, cooling=cooling)
    if graph:
        pl.set_graph(True)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    torch.cuda.synchronize()
    states = {}
    t = 0.0
    done = 0
    for target in sorted(set(record) | {n_steps}):
        t0 = time.perf_counter()
        pl.step(target - done, dt, cooling=cooling is not None)
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
        done = target
        if target in record:
            states[target] = pl.state()
    diag = pl.diagnostics(True)
    cnt = pl.counters()
    pl.close()
    return states, 1e3 * t / n_steps, diag, cnt


def bench_c1(a) -> int:
    """Evrard N=4096 seed 0: bootstrap + one KDK step timed in full on the FP64
    CPU oracle (all host threads) and on the GPU (mixed and plain FP64), with the
    GPU states compared against the oracle's after the bootstrap and the step.
    With --n 131072 this is step 1 of the C2 configuration timed on the CPU."""
    import torch

    import oracle
    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan

    n = a.n or 4096
    e = ic.evrard(n, seed=a.seed)
    p = oracle.Params()
    cores = len(os.sched_getaffinity(0))
    log(f"c1 Evrard N={n}: CPU oracle on {cores} threads")
    t0 = time.perf_counter()
    s0 = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    t1 = time.perf_counter()
    s1 = oracle.step(s0, a.dt, p)
    t2 = time.perf_counter()
    log(f"oracle bootstrap {t1 - t0:.3f} s, step {t2 - t1:.3f} s")

    def rel_max(x, y):
        return float(np.max(np.abs(x / y - 1)))

    gpu = {}
    for mode in ("mixed", "fp64"):
        pl = SphPlan(n, mode=mode)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        torch.cuda.synchronize()
        g0 = time.perf_counter()
        pl.init()
        torch.cuda.synchronize()
        g1 = time.perf_counter()
        st0 = pl.state()
        torch.cuda.synchronize()
        g2 = time.perf_counter()
        pl.step(1, a.dt)
        torch.cuda.synchronize()
        g3 = time.perf_counter()
        st1 = pl.state()
        pl.close()
        gpu[mode] = {
            "bootstrap_ms": 1e3 * (g1 - g0), "step_ms": 1e3 * (g3 - g2),
            "particle_steps_per_s": n / (g3 - g2),
            "bootstrap": {"h_identical": bool(np.array_equal(st0["h"], s0.h)),
                          "density_rel_max": rel_max(st0["rho"], s0.rho),
                          "acc_rel_rms": rel_rms(st0["acc"], s0.acc)},
            "step1": {"k_mismatch_frac": float(np.mean(st1["k"] != s1.k)),
                      "h_rel_max": rel_max(st1["h"], s1.h),
                      "density_rel_max": rel_max(st1["rho"], s1.rho),
                      "acc_rel_rms": rel_rms(st1["acc"], s1.acc),
                      "acc_grav_rel_rms": rel_rms(st1["acc_g"], s1.acc_g),
                      "dudt_rel_rms": rel_rms(st1["dudt"], s1.dudt),
                      "u_rel_max": rel_max(st1["u"], s1.u),
                      "pos_rel_rms": rel_rms(st1["pos"], s1.pos)},
        }
        log(f"{mode}: step {gpu[mode]['step_ms']:.2f} ms, "
            f"acc rel rms {gpu[mode]['step1']['acc_rel_rms']:.2e}")
    out = {"workload": f"c1: Evrard N={n} seed {a.seed}, bootstrap + one KDK step at dt={a.dt}",
           "cpu_oracle": {"bootstrap_s": t1 - t0, "step_s": t2 - t1, "cores": cores,
                          "cpu": cpu_model(), "particle_steps_per_s": n / (t2 - t1)},
           "gpu": gpu,
           "speedup_mixed_vs_cpu_step": (t2 - t1) * 1e3 / gpu["mixed"]["step_ms"]}
    print(json.dumps(out))
    return 0


def bench_c2(a) -> int:
    """Evrard N=131072, 100 steps at dt=1e-4, mixed vs plain FP64 on the GPU,
    errors at steps 1 and 100; the FP64 GPU forces at step 1 checked against
    the CPU oracle on sampled rows."""
    import oracle
    from sph_2_amd import ic

    n = a.n or 131072
    steps = 100
    e = ic.evrard(n, seed=a.seed)
    log(f"c2 Evrard N={n}")
    mx, ms_mx, d_mx, c_mx = _plan_run(steps, e, "mixed", a.dt, record=(1, steps), graph=bool(a.graph))
    log(f"mixed {ms_mx:.2f} ms/step")
    fp, ms_fp, d_fp, c_fp = _plan_run(steps, e, "fp64", a.dt, record=(1, steps))
    log(f"fp64 {ms_fp:.2f} ms/step")
    errs = {}
    for s in (1, steps):
        A, B = mx[s], fp[s]
        errs[str(s)] = {
            "density_rel_max": float(np.max(np.abs(A["rho"] / B["rho"] - 1))),
            "acc_rel_rms": rel_rms(A["acc"], B["acc"]),
            "acc_grav_rel_rms": rel_rms(A["acc_g"], B["acc_g"]),
            "h_rel_max": float(np.max(np.abs(A["h"] / B["h"] - 1))),
            "k_mismatch_frac": float(np.mean(A["k"] != B["k"])),
            "pos_rel_rms": rel_rms(A["pos"], B["pos"]),
            "u_rel_max": float(np.max(np.abs(A["u"] / B["u"] - 1))),
        }
    # chain of trust: FP64 GPU forces at step 1 vs the CPU oracle on rows.
    # Density and gravity depend only on the stored positions and h; the
    # hydro terms used the half-kicked velocity and the pre-update u, which the
    # stored state no longer holds (their parity is the C1 stateless tests').
    B = fp[1]
    rows = np.arange(0, n, max(1, n // 2048))
    t0 = time.perf_counter()
    rho_o, _ = oracle.density(B["pos"], e.m, B["h"], rows=rows)
    g_o = oracle.gravity(B["pos"], e.m, B["h"], 1.0, rows=rows)
    tcpu = time.perf_counter() - t0
    chain = {"rows": int(len(rows)),
             "density_rel_max": float(np.max(np.abs(B["rho"][rows] / rho_o - 1))),
             "gravity_rel_rms": rel_rms(B["acc_g"][rows], g_o),
             "cpu_s": tcpu}
    out = {"workload": f"c2: Evrard N={n}, {steps} KDK steps at dt={a.dt}",
           "mixed_ms_per_step": ms_mx, "fp64_ms_per_step": ms_fp,
           "mixed_particle_steps_per_s": n * 1e3 / ms_mx,
           "fp64_particle_steps_per_s": n * 1e3 / ms_fp,
           "speedup": ms_fp / ms_mx,
           "mixed_vs_fp64_gpu": errs,
           "fp64_gpu_vs_cpu_oracle_step1": chain,
           "energy_mixed": d_mx, "energy_fp64": d_fp,
           "counters_mixed": c_mx}
    print(json.dumps(out))
    return 0


def bench_c3(a) -> int:
    """Turbulent cooling cloud N=1,048,576, 50 mixed-precision steps with the
    implicit FP64 cooling solve."""
    import torch

    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan
    from sph_2_amd.serve import default_cooling

    n = a.n or 1_048_576
    steps = 50
    t0 = time.perf_counter()
    e = ic.turbulent_cloud(n, seed=a.seed)
    log(f"c3 cloud N={n} IC in {time.perf_counter() - t0:.1f}s")
    cool = default_cooling()
    pl = SphPlan(n, cooling=cool)
    if a.graph:
        pl.set_graph(True)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    d0 = pl.diagnostics(True)
    st0 = pl.state(("u",))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl.step(steps, a.dt, cooling=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    d1 = pl.diagnostics(True)
    st1 = pl.state(("u", "rho"))
    cnt = pl.counters()
    pl.close()
    units = e.units
    tfac = (5.0 / 3.0 - 1.0) * cool.kw["mu"] * ic.MH / ic.KB * units["unit_u_cgs"]
    out = {"workload": f"c3: turbulent cooling cloud N={n}, rms Mach 2, T0=1e4 K, {steps} steps "
                       f"at dt={a.dt} (code units: 1e4 Msun, 20 pc), 200-node Lambda(T) table",
           "ms_per_step": 1e3 * wall / steps, "particle_steps_per_s": n * steps / wall,
           "T_median_K_start": float(np.median(st0["u"]) * tfac),
           "T_median_K_end": float(np.median(st1["u"]) * tfac),
           "T_min_K_end": float(st1["u"].min() * tfac),
           "energy_start": d0, "energy_end": d1, "counters": cnt,
           "finite": bool(np.isfinite(st1["u"]).all() and np.isfinite(st1["rho"]).all())}
    print(json.dumps(out))
    return 0


def bench_bh(a) -> int:
    """Barnes-Hut on a Plummer sphere N=4,194,304, eps=0.01, theta 0.3/0.5/0.7;
    force RMS error against FP64 direct summation on 65,536 seeded rows."""
    import torch

    from sph_2_amd import ic, libsx

    n = a.n or 4_194_304
    e = ic.plummer(n, seed=a.seed, eps=0.01)
    nsub = min(65536, n)
    log(f"bh Plummer N={n}")
    t0 = time.perf_counter()
    ref = libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.eps, mode="fp64", nbeg=0, nend=nsub)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    log(f"fp64 direct rows {t_ref:.2f}s")
    res = {}
    for theta in (0.3, 0.5, 0.7):
        libsx.getAcc_g_bh(e.pos, e.m, 1.0, e.eps, theta)   # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc, st = libsx.getAcc_g_bh(e.pos, e.m, 1.0, e.eps, theta, return_stats=True)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        err = acc[:nsub] - ref
        rel = np.linalg.norm(err, axis=1) / np.linalg.norm(ref, axis=1)
        res[str(theta)] = {"s_per_eval_incl_transfer": t, "rel_rms": rel_rms(acc[:nsub], ref),
                           "rel_median": float(np.median(rel)), "rel_p99": float(np.percentile(rel, 99)),
                           "interactions_per_particle": st["interactions"] / n, **st}
        log(f"theta {theta}: {t:.3f}s (device build {st['build_us'] / 1e3:.1f} ms, walk "
            f"{st['walk_us'] / 1e3:.1f} ms) rel_rms {res[str(theta)]['rel_rms']:.2e}")
    out = {"workload": f"c4: Barnes-Hut, Plummer N={n}, a=1, eps=0.01, FP32; error vs FP64 "
                       f"direct sum on rows 0..{nsub} (random order IC)",
           "fp64_direct_rows_s": t_ref, "theta": res}
    print(json.dumps(out))
    return 0


def serve_stream(a, runs: int, verbose: bool = False) -> dict:
    """The served stream: `runs` Evrard/cloud requests, N log-uniform in
    [nmin, nmax], run_steps steps each, seeded Poisson arrivals at `rate`/s,
    admitted by the HBM budget, every live run advanced by one launch per
    pass (graph replay), retired as soon as it finishes."""
    from sph_2_amd.serve import Server, build_payload, make_stream

    reqs = make_stream(runs, n_min=a.nmin, n_max=a.nmax, rate=a.rate, steps=a.run_steps,
                       seed=a.seed)
    t0 = time.perf_counter()
    for r in reqs:
        r.payload = build_payload(r)
    log(f"serve: {len(reqs)} requests built in {time.perf_counter() - t0:.1f}s")
    sym = None if a.serve_gravity == "tiled" else (int(a.serve_gravity.split(",")[0]),
                                                    float(a.serve_gravity.split(",")[1]))
    srv = Server(int(a.hbm_gb * 2 ** 30), max_runs=512, graph=bool(a.serve_graph), symmetric=sym)
    done = [0]

    def progress(m):
        done[0] += 1
        if verbose or done[0] % 32 == 0:
            log(f"serve: {done[0]} runs done; {m}")

    rep = srv.run(reqs, log=progress)
    srv.batch.close()
    keep = ("runs", "wall_s", "particle_steps", "particle_steps_per_s", "p50_latency_s",
            "p99_latency_s", "p50_queue_s", "p99_queue_s", "capacity_particles", "host_s",
            "loop_steps", "max_steps_in_flight", "gpu_step_s", "gpu_step_frac")
    out = {k: rep[k] for k in keep if k in rep}
    out["config"] = {
        "workload": f"{runs} Evrard/cooling-cloud runs (half each), N = floor({a.nmin} * "
                    f"({a.nmax}/{a.nmin})^U) log-uniform, {a.run_steps} KDK steps each at "
                    f"dt=1e-4, seeded Poisson arrivals (seed {a.seed})",
        "runs": runs, "n_min": a.nmin, "n_max": a.nmax, "steps": a.run_steps,
        "arrival_rate_per_s": a.rate, "hbm_budget_gb": a.hbm_gb,
        "graph_replay": bool(a.serve_graph),
        "gravity": ("tiled" if sym is None else
                    f"symmetric ({sym[0]} block pairs per wave, {sym[1]:.0%} of the budget for "
                    f"its partial sums)")}
    return out


def bench_serve(a) -> int:
    """Served stream alone (the same as the default run's serving block)."""
    rep = serve_stream(a, a.runs, verbose=True)
    print(json.dumps(rep))
    return 0


def aug_run(n_steps, mode="mixed", log_every=0):
    """The reference's own Python Evrard run that wrote Evrard_collapse/Energy.pkl
    (augsphx6.5_Monagh_visc.py, the Monaghan-Gingold variant of augsphx6.5.py:406-567) on
    this engine: its N = 1472 IC (float32 positions), v = 0, u = 0.05, m = 1/N, softening
    0.05, kNN h each step (do_smoothingX), alpha = 1, beta = 2, eta = 0.1, dt = 1e-3.
    Returns per step (U, KE at the half-kicked velocity, PE), as Energy.pkl records them."""
    from sph_2_amd import ic as _ic, libsx
    from sph_2_amd.plan import SphPlan

    here = os.path.dirname(os.path.abspath(__file__))
    d = np.load(os.path.join(here, "tests", "golden", "evrard_1472_ic.npz"))
    pos = np.stack([d["x"], d["y"], d["z"]], 1).astype(np.float64)
    n = len(pos)
    m = np.full(n, 1.0 / n)
    u = np.full(n, 0.05)
    eps = np.full(n, 0.05)
    dt = 1e-3
    h0 = libsx.do_smoothingX((pos, pos))
    pl = SphPlan(n, mode=mode, h_mode=1, knn_kth=64, av_form=0, balsara=0, h_predictor=0,
                 eps_from_h=0, alpha=1.0, beta=2.0, eta=0.1)
    pl.load(pos, np.zeros((n, 3)), u, m, h0, eps)
    pl.init()
    out = []
    for k in range(n_steps):
        pl.step(1, dt)
        st = pl.state(("pos", "vel", "u", "acc"))
        vmid = st["vel"] - st["acc"] * (0.5 * dt)
        ke = float(0.5 * (m * (vmid ** 2).sum(1)).sum())
        pe = libsx.getPE(st["pos"], m, 1.0, eps)
        out.append((float((m * st["u"]).sum()), ke, pe))
        if log_every and (k + 1) % log_every == 0:
            log(f"aug step {k + 1}")
    pl.close()
    return np.array(out)


def bench_aug(a) -> int:
    """Energy history of the reference's N = 1472 Python run against its recorded Energy.pkl."""
    here = os.path.dirname(os.path.abspath(__file__))
    E = np.load(os.path.join(here, "tests", "golden", "evrard_1472_energy.npz"))
    steps = min(a.steps if a.steps > 5 else 3000, len(E["U"]) - 1)
    t0 = time.perf_counter()
    hist = aug_run(steps, log_every=250)
    wall = time.perf_counter() - t0
    ref = np.stack([E["U"][1:steps + 1], E["KE"][1:steps + 1], E["PE"][1:steps + 1]], 1)
    scale = np.abs(E["PE"][0])
    res = {"workload": f"reference Python Evrard run (Energy.pkl): N=1472, {steps} steps at "
                       f"dt=1e-3, kNN h, Monaghan-Gingold AV, eps=0.05, mixed precision",
           "wall_s": wall, "steps": steps}
    for tmax in (0.1, 0.5, 1.0, 2.0, 3.0):
        k = min(steps, int(round(tmax / 1e-3)))
        if k <= 0:
            continue
        dev = np.abs(hist[:k] - ref[:k])
        res[f"t<={tmax}"] = {"U_rel": float((dev[:, 0] / np.abs(ref[:k, 0])).max()),
                              "KE_over_|PE0|": float(dev[:, 1].max() / scale),
                              "PE_rel": float((dev[:, 2] / np.abs(ref[:k, 2])).max())}
    res["final"] = {"ours": hist[-1].tolist(), "reference": ref[-1].tolist()}
    print(json.dumps(res))
    return 0


def bench_gadget(a) -> int:
    """Evrard collapse to t = 0.8 with the canonical step and the device's adaptive dt, radial
    profiles of rho, P and v_r against the Gadget-4 profiles the reference keeps
    (Evrard_collapse/{rho,P,vr}Gadget.pkl, compared by its singlex.py:143-200)."""
    import torch

    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan

    here = os.path.dirname(os.path.abspath(__file__))
    G4 = np.load(os.path.join(here, "tests", "golden", "gadget4_profiles.npz"))
    n = a.n or 131072
    e = ic.evrard(n, seed=a.seed)
    fixed = a.eps is not None
    pl = SphPlan(n, eps_from_h=0) if fixed else SphPlan(n)
    pl.set_adaptive_dt(True, 0.25, 1e-5, 1e-3)
    pl.load(e.pos, e.vel, e.u, e.m, e.h, np.full(n, a.eps) if fixed else None)
    pl.init()
    t_end = 0.8
    steps = 0
    t0 = time.perf_counter()
    while True:
        t, dt, _ = pl.time()
        if t >= t_end - 1e-12:
            break
        pl.step(1, 1e-4)       # seeds the first step; later steps use the device's dt
        steps += 1
        if steps % 500 == 0:
            log(f"gadget t={t:.3f} dt={dt:.2e} step {steps}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t, _, _ = pl.time()
    st = pl.state(("pos", "vel", "u", "rho"))
    pl.close()
    r = np.linalg.norm(st["pos"], axis=1)
    vr = (st["vel"] * st["pos"]).sum(1) / np.maximum(r, 1e-300)
    P = (5.0 / 3.0 - 1.0) * st["rho"] * st["u"]
    rg = G4["rho_r"]
    edges = np.sqrt(rg[1:] * rg[:-1])
    edges = np.concatenate([[rg[0] ** 2 / edges[0]], edges, [rg[-1] ** 2 / edges[-1]]])
    k = np.digitize(r, edges) - 1
    soft = f"eps={a.eps}" if fixed else "eps=h"
    out = {"workload": f"Evrard N={n} to t={t:.4f} ({steps} adaptive steps), canonical step "
                       f"(band h, Monaghan-Gingold AV, {soft}), vs Gadget-4 profiles",
           "wall_s": wall, "t": t, "steps": steps}
    for name, ours_v, ref in (("rho", st["rho"], G4["rho_rho"]), ("P", P, G4["P_P"]),
                              ("vr", vr, G4["vr_vr"])):
        rows = []
        for b in range(len(rg)):
            sel = k == b
            if not np.isfinite(ref[b]) or sel.sum() < 8:
                continue
            rows.append((float(rg[b]), float(ours_v[sel].mean()), float(ref[b]), int(sel.sum())))
        rows = np.array(rows)
        if name == "vr":
            dev = np.abs(rows[:, 1] - rows[:, 2])
            out[name] = {"bins": len(rows), "mean_abs_diff": float(dev.mean()),
                         "max_abs_diff": float(dev.max())}
        else:
            lr = np.abs(np.log(rows[:, 1] / rows[:, 2]))
            out[name] = {"bins": len(rows), "median_abs_log_ratio": float(np.median(lr)),
                         "max_abs_log_ratio": float(lr.max())}
        out[name]["table_r_ours_gadget_count"] = rows.round(5).tolist()
    print(json.dumps(out))
    return 0


COPY_BYTES = 256 * 2 ** 20


def bench_block(a) -> int:
    """Block time steps at Evrard N (default 1,048,576): a probe substep finds
    the spread of the per-particle step criterion, which sets the longest step
    (the coarsest occupied level becomes level 0) and the number of levels;
    then one full cycle of the longest step with block steps is timed against
    the fixed-dt step at the shortest level over the same simulated time (what
    a single global step must use so that every particle meets its criterion)."""
    import torch

    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan

    n = a.n or 1_048_576
    e = ic.evrard(n, seed=a.seed)
    probe_dt, probe_levels = 0.4, 10
    pr = SphPlan(n)
    pr.set_block_steps(probe_levels)
    pr.load(e.pos, e.vel, e.u, e.m, e.h)
    pr.init()
    pr.step(1, probe_dt)
    b = pr.block_bins()
    pr.close()
    lo, hi = int(b.min()), int(b.max())
    levels = max(hi - lo, 1)
    dt_max = probe_dt * 2.0 ** -lo
    dt_min = dt_max * 2.0 ** -levels
    log(f"block: levels {lo}..{hi} of the probe -> dt_max {dt_max:.3e}, {levels} levels")
    res = {}
    for mode in ("block", "fixed"):
        pl = SphPlan(n)
        if mode == "block":
            pl.set_block_steps(levels)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        torch.cuda.synchronize()
        pl.set_timing(True)
        t0 = time.perf_counter()
        pl.step(2 ** levels, dt_max if mode == "block" else dt_min)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ph = pl.timing()
        entry = {"wall_s": wall, "substeps": 2 ** levels, "simulated_time": dt_max,
                 "gravity_ms_per_substep": ph["gravity"][0] / max(ph["gravity"][1], 1),
                 "ms_per_substep": 1e3 * wall / 2 ** levels}
        if mode == "block":
            bins = pl.block_bins()
            entry["levels_population"] = {int(k): int(v) for k, v in
                                          zip(*np.unique(bins, return_counts=True))}
            entry["counters"] = pl.counters()
        res[mode] = entry
        pl.close()
        log(f"{mode}: {wall:.2f} s for {dt_max:.3e} of simulated time")
    out = {"workload": f"block time steps: Evrard N={n} seed {a.seed}, {levels} levels below "
                       f"dt_max={dt_max:.4e}, one cycle ({2 ** levels} substeps) vs the fixed "
                       f"step dt_max/2^{levels}", "n": n, "levels": levels, "dt_max": dt_max,
           "block": res["block"], "fixed": res["fixed"],
           "speedup_simulated_time_per_wall": res["fixed"]["wall_s"] / res["block"]["wall_s"]}
    print(json.dumps(out))
    return 0


def bench_kernels(a) -> int:
    """Driver for hardware-counter runs (rocprofv3 --pmc): first a device copy
    of COPY_BYTES (a kernel of known traffic, to calibrate FETCH_SIZE and
    WRITE_SIZE), then the Evrard plan of N particles (default 1,048,576):
    bootstrap plus --steps mixed steps.  Prints what ran."""
    import torch

    from sph_2_amd import ic
    from sph_2_amd.plan import SphPlan

    n = a.n or 1_048_576
    x = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    e = ic.evrard(n, seed=a.seed)
    pl = SphPlan(n)
    if a.grav:
        pl.set_gravity_config(*(int(x) for x in a.grav.split(",")))
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(a.steps, a.dt)
    torch.cuda.synchronize()
    pl.close()
    print(json.dumps({"workload": f"counter driver: {COPY_BYTES}-byte device copy, then Evrard "
                                  f"N={n} bootstrap + {a.steps} mixed steps", "n": n,
                      "steps": a.steps, "copy_bytes": COPY_BYTES, "grav": a.grav}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
