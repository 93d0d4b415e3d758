"""Continuous batching of independent SPH runs on one GPU.

A ``Batch`` packs many runs into one device engine (``sphx_batch_*``); every
``step`` launch advances all live runs.  ``Server.run`` is the request loop:
requests arrive at their (seeded Poisson) arrival times, are admitted while
the HBM budget and run slots allow (first come, first served), advance one
step per loop turn together with every other live run, and retire as soon as
they finish; each result is copied out in the caller's particle order.
Reported: p50/p99 time to completion (arrival -> result on the host), the
queueing part of it, and aggregate particle-steps/s.
"""
from __future__ import annotations

import collections
import ctypes as C
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, ic
from ._lib import check, dev, ptr
from .libsx import CoolingTable


class Batch:
    def __init__(self, capacity: int, max_runs: int = 512, *, params: _lib.Params | None = None,
                 cooling: CoolingTable | None = None, device: int = 0, graph: bool = False,
                 **param_kw):
        _lib.require_gpu()
        self.params = params if params is not None else _lib.default_params(**param_kw)
        self._cool_c = cooling.c() if cooling is not None else None
        self._h = C.c_void_p()
        self.capacity = int(capacity)
        self.max_runs = int(max_runs)
        check(_lib.load().sphx_batch_create(device, _lib.MIXED, C.byref(self.params),
                                            C.byref(self._cool_c) if self._cool_c else None,
                                            self.capacity, self.max_runs, C.byref(self._h)),
              "sphx_batch_create")
        self.sizes: dict[int, int] = {}
        if graph:
            self.set_graph(True)

    @staticmethod
    def bytes_per_particle() -> int:
        return int(_lib.load().sphx_batch_bytes_per_particle())

    def close(self):
        if self._h:
            _lib.load().sphx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool = True):
        check(_lib.load().sphx_batch_set_graph(self._h, int(on)), "set_graph")

    @property
    def live_particles(self) -> int:
        return sum(self.sizes.values())

    def set_cooling_grid(self, grid):
        """Attach a libsx.CoolingGrid for runs admitted with cooling="table"."""
        g = grid.c()
        check(_lib.load().sphx_batch_set_cooling_grid(self._h, C.byref(g)), "set_cooling_grid")

    def set_cooling_network(self, network):
        """Attach a libsx.CoolingNetwork for runs admitted with cooling="network"."""
        c = network.c()
        check(_lib.load().sphx_batch_set_cooling_network(self._h, C.byref(c)),
              "set_cooling_network")

    def counters(self) -> dict:
        """Counters of the last step over all live runs (h-iteration caps,
        cooling caps, non-finite particles, ...)."""
        c = _lib.Counters()
        check(_lib.load().sphx_batch_query(self._h, C.byref(c)), "sphx_batch_query")
        return c.as_dict()

    def set_gravity_symmetric(self, block_pairs: int, budget_bytes: int):
        """Symmetric direct-sum gravity for every run (before the first admission),
        with budget_bytes reserved for its partial sums."""
        check(_lib.load().sphx_batch_set_gravity_symmetric(self._h, int(block_pairs),
                                                           int(budget_bytes)),
              "set_gravity_symmetric")
        self.sym = (int(block_pairs), int(budget_bytes))

    @staticmethod
    def symmetric_bytes(n: int, block_pairs: int) -> int:
        """All partial sums of one run of n particles (summed in row bands)."""
        return int(_lib.load().sphx_gravity_symmetric_bytes(int(n), int(block_pairs)))

    @staticmethod
    def symmetric_min_bytes(n: int, block_pairs: int) -> int:
        """The least reserved memory with which a run of n particles is admitted."""
        return int(_lib.load().sphx_gravity_symmetric_min_bytes(int(n), int(block_pairs)))

    def set_chimes_table(self, table):
        """Attach a libsx.ChimesTable for runs admitted with cooling="chimes"."""
        c = table.c()
        check(_lib.load().sphx_batch_set_chimes_table(self._h, C.byref(c)), "set_chimes_table")

    def admit(self, ic_: ic.IC, dt: float, cooling=False) -> int | None:
        """Admit one run; None when the budget or the run slots are full."""
        t = _lib.torch()
        _lib.check_run_arrays(ic_.pos, ic_.vel, ic_.u, ic_.m, ic_.h, ic_.eps)
        if ic_.n != len(ic_.m):
            raise ValueError(f"the run declares n = {ic_.n} but holds {len(ic_.m)} particles")
        bufs = [dev(np.asarray(a, np.float64), t.float64) for a in
                (ic_.pos, ic_.vel, ic_.u, ic_.m, ic_.h)]
        de = dev(np.asarray(ic_.eps, np.float64), t.float64) if ic_.eps is not None else None
        rid = C.c_int64()
        rc = check(_lib.load().sphx_batch_admit(self._h, ic_.n, *[ptr(b) for b in bufs], ptr(de),
                                                float(dt), _lib.cooling_mode(cooling), C.byref(rid),
                                                _lib.stream_handle()), "sphx_batch_admit")
        if rc == _lib.REFUSED:
            return None
        self.sizes[rid.value] = ic_.n
        return rid.value

    def set_adaptive_dt(self, run_id: int, on: bool = True, c_cfl: float = 0.25,
                        dt_min: float = 1e-5, dt_max: float = 1e-3):
        check(_lib.load().sphx_batch_set_adaptive_dt(self._h, int(run_id), int(on), c_cfl, dt_min,
                                                     dt_max, _lib.stream_handle()),
              "set_adaptive_dt")

    def time(self, run_id: int) -> tuple[float, float, tuple[float, float, float, float]]:
        t = C.c_double()
        dt = C.c_double()
        parts = (C.c_double * 4)()
        check(_lib.load().sphx_batch_time(self._h, int(run_id), C.byref(t), C.byref(dt), parts,
                                          _lib.stream_handle()), "time")
        return t.value, dt.value, tuple(parts)

    def save(self, run_id: int, path):
        """Exact checkpoint of one live run (it stays live)."""
        check(_lib.load().sphx_batch_save_run(self._h, int(run_id), os.fsencode(path),
                                              _lib.stream_handle()), "sphx_batch_save_run")

    def restore(self, path) -> int | None:
        """Re-admit a checkpointed run; None when the budget or run slots are full."""
        from .plan import checkpoint_info

        rid = C.c_int64()
        rc = check(_lib.load().sphx_batch_restore_run(self._h, os.fsencode(path), C.byref(rid),
                                                      _lib.stream_handle()),
                   "sphx_batch_restore_run")
        if rc == _lib.REFUSED:
            return None
        self.sizes[rid.value] = checkpoint_info(path)["n"]
        return rid.value

    def step(self, n_steps: int = 1):
        check(_lib.load().sphx_batch_step(self._h, int(n_steps), _lib.stream_handle()),
              "sphx_batch_step")

    def _out(self, run_id, fn):
        t = _lib.torch()
        n = self.sizes[run_id]
        o = {"pos": t.empty((n, 3), dtype=t.float64, device="cuda"),
             "vel": t.empty((n, 3), dtype=t.float64, device="cuda"),
             "u": t.empty(n, dtype=t.float64, device="cuda"),
             "h": t.empty(n, dtype=t.float64, device="cuda"),
             "rho": t.empty(n, dtype=t.float64, device="cuda"),
             "acc": t.empty((n, 3), dtype=t.float64, device="cuda"),
             "k": t.empty(n, dtype=t.int32, device="cuda")}
        check(fn(self._h, int(run_id), *[ptr(o[f]) for f in ("pos", "vel", "u", "h", "rho", "acc",
                                                           "k")], _lib.stream_handle()),
              "batch store")
        return {f: v.cpu().numpy() for f, v in o.items()}

    def store(self, run_id: int) -> dict:
        return self._out(run_id, _lib.load().sphx_batch_store)

    def retire(self, run_id: int) -> dict:
        out = self._out(run_id, _lib.load().sphx_batch_retire)
        del self.sizes[run_id]
        return out

    def runs(self):
        cap = max(1, len(self.sizes))
        ids = (C.c_int64 * cap)()
        n = (C.c_int64 * cap)()
        st = (C.c_int64 * cap)()
        cnt = C.c_int32()
        check(_lib.load().sphx_batch_runs(self._h, ids, n, st, cap, C.byref(cnt)), "runs")
        return [(ids[i], n[i], st[i]) for i in range(cnt.value)]


# ---------------------------------------------------------------------------
# Request stream and the serving loop
# ---------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    arrival: float        # seconds after the stream starts
    kind: str             # "evrard" or "cloud"
    n: int
    seed: int
    steps: int
    dt: float
    payload: ic.IC | None = field(default=None, repr=False)


def make_stream(n_runs: int, n_min: int = 2048, n_max: int = 262144, rate: float = 2.0,
                steps: int = 200, seed: int = 0, cloud_fraction: float = 0.5,
                dt: float = 1e-4) -> list[Request]:
    """Seeded Poisson arrivals (rate per second), N log-uniform in [n_min, n_max]."""
    u_gap = ic.uniform(seed, 7001, n_runs)
    u_n = ic.uniform(seed, 7002, n_runs)
    u_kind = ic.uniform(seed, 7003, n_runs)
    gaps = -np.log(np.maximum(1.0 - u_gap, 1e-300)) / rate
    arrivals = np.cumsum(gaps) - gaps[0]
    ns = np.floor(n_min * (n_max / n_min) ** u_n).astype(np.int64)
    return [Request(r, float(arrivals[r]), "cloud" if u_kind[r] < cloud_fraction else "evrard",
                    int(ns[r]), seed * 100003 + r, steps, dt) for r in range(n_runs)]


def build_payload(req: Request) -> ic.IC:
    if req.kind == "cloud":
        return ic.turbulent_cloud(req.n, seed=req.seed, ng=32)
    return ic.evrard(req.n, seed=req.seed)


def default_cooling() -> CoolingTable:
    units = ic.cloud_units()
    T, L = ic.cooling_table(seed=0)
    return CoolingTable(T, L, unit_u_cgs=units["unit_u_cgs"], unit_rho_cgs=units["unit_rho_cgs"],
                        unit_time_s=units["unit_time_s"], heat=2e-26)


class Server:
    """Admit by HBM budget, step every live run together, retire when done."""

    def __init__(self, hbm_budget_bytes: int, max_runs: int = 512, graph: bool = True,
                 cooling: CoolingTable | None = None, symmetric: tuple | None = (4, 0.1)):
        """symmetric: (block pairs per wave, share of the budget reserved for
        the symmetric gravity sum's partials, which it streams in row bands),
        or None for the tiled sum."""
        share = symmetric[1] if symmetric else 0.0
        self.capacity = int(hbm_budget_bytes * (1.0 - share) // Batch.bytes_per_particle())
        self.batch = Batch(self.capacity, max_runs, cooling=cooling or default_cooling(),
                           graph=graph)
        self.symmetric = symmetric
        if symmetric:
            self.sym_budget = int(hbm_budget_bytes * share)
            self.batch.set_gravity_symmetric(symmetric[0], self.sym_budget)

    def run(self, requests: list[Request], keep_results: bool = False, log=None,
            max_in_flight: int = 2) -> dict:
        """Serve `requests`; returns the latency and throughput report.

        The host does not wait for each step: up to `max_in_flight` batch steps
        are queued on the stream (each followed by an event) while the host
        admits arrivals and does its bookkeeping; it waits only for the oldest
        step's event when more are queued, and on the synchronous copy-out of a
        finished run, whose completion time is taken after that copy."""
        torch = _lib.torch()
        if max_in_flight < 1:
            raise ValueError("max_in_flight must be at least 1")
        for r in requests:
            if r.payload is None:
                r.payload = build_payload(r)
            if r.n > self.capacity or (self.symmetric and Batch.symmetric_min_bytes(
                    r.n, self.symmetric[0]) > self.sym_budget):
                raise ValueError(f"request {r.rid} (N={r.n}) exceeds the whole budget")
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        live: dict[int, tuple[Request, float, int]] = {}   # run id -> (req, admit time, steps)
        done = {}
        results = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        busy = 0.0
        loops = 0
        psteps = 0
        t_admit = t_step = t_retire = 0.0
        inflight = collections.deque()      # events of queued steps
        step_events = []                    # (start, end) of every batch step on the stream
        while pending or live:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                req = pending[0]
                ta0 = time.perf_counter()
                rid = self.batch.admit(req.payload, req.dt, cooling=req.kind == "cloud")
                t_admit += time.perf_counter() - ta0
                if rid is None:
                    break               # first come, first served: wait for room
                pending.popleft()
                live[rid] = (req, time.perf_counter() - t0, 0)
            if not live:
                nxt = pending[0].arrival if pending else now
                time.sleep(max(0.0, min(0.01, nxt - now)))
                continue
            tb = time.perf_counter()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            self.batch.step(1)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            inflight.append(ev)
            step_events.append((ev0, ev))
            while len(inflight) >= max_in_flight:
                inflight.popleft().synchronize()
            t_step += time.perf_counter() - tb
            loops += 1
            finished = []
            for rid, (req, ta, s) in list(live.items()):
                s += 1
                psteps += req.n
                live[rid] = (req, ta, s)
                if s >= req.steps:
                    finished.append(rid)
            if finished:
                inflight.clear()                  # the copy-out below waits for every step
            for rid in finished:
                req, ta, s = live.pop(rid)
                tr0 = time.perf_counter()
                out = self.batch.retire(rid)      # synchronous copy-out
                t_retire += time.perf_counter() - tr0
                tc = time.perf_counter() - t0
                done[req.rid] = {"arrival": req.arrival, "admitted": ta, "completed": tc,
                                 "n": req.n, "kind": req.kind}
                if keep_results:
                    results[req.rid] = out
                if log:
                    log(f"run {req.rid} N={req.n} {req.kind} done at {tc:.2f}s "
                        f"(latency {tc - req.arrival:.2f}s)")
            busy += time.perf_counter() - tb
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_step_s = 1e-3 * sum(a.elapsed_time(b) for a, b in step_events)
        lat = np.array([d["completed"] - d["arrival"] for d in done.values()])
        queue = np.array([d["admitted"] - d["arrival"] for d in done.values()])
        rep = {"runs": len(done), "wall_s": wall, "loop_steps": loops,
               "particle_steps": int(psteps), "particle_steps_per_s": psteps / wall,
               "p50_latency_s": float(np.percentile(lat, 50)),
               "p99_latency_s": float(np.percentile(lat, 99)),
               "p50_queue_s": float(np.percentile(queue, 50)),
               "p99_queue_s": float(np.percentile(queue, 99)),
               "capacity_particles": self.capacity,
               "max_steps_in_flight": max_in_flight,
               "gpu_step_s": gpu_step_s, "gpu_step_frac": gpu_step_s / wall,
               "host_s": {"admit": t_admit, "step": t_step, "retire": t_retire}}
        if keep_results:
            rep["results"] = results
            rep["done"] = done
        return rep
