"""Resident-state SPH run on one GPU (sphx_plan_* of include/sphx.h).

The whole step stays on the device between calls: half kick, drift, Morton
sort, smoothing-length iteration, density + EOS, gravity, div/curl, hydro
acceleration + du/dt, half kick, u trapezoid and optional implicit cooling
(the step order of 11_Dec_2022/GPU_sph/Important_Test_Simulations/
Evrard_collapse/hfv_sph_gpu_v2_evrard.cu:407-482).
"""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np

from . import _lib
from ._lib import check, dev, ptr
from .libsx import CoolingTable


def checkpoint_info(path) -> dict:
    """Header fields of a checkpoint written by SphPlan.save / Batch.save:
    magic "SPHXRUN1", int32 version, int32 mode, int64 n, int64 steps."""
    with open(path, "rb") as f:
        head = f.read(32)
    if len(head) < 32 or head[:8] != b"SPHXRUN1":
        raise ValueError(f"{path}: not an sphx checkpoint")
    version, mode, n, steps = struct.unpack("<iiqq", head[8:32])
    return {"version": version, "mode": "mixed" if mode == _lib.MIXED else "fp64", "n": n,
            "steps": steps}


class SphPlan:
    def __init__(self, capacity: int, *, mode: str = "mixed", params: _lib.Params | None = None,
                 cooling: CoolingTable | None = None, device: int = 0, **param_kw):
        _lib.require_gpu()
        self.mode = {"mixed": _lib.MIXED, "fp64": _lib.FP64}[mode]
        self.params = params if params is not None else _lib.default_params(**param_kw)
        self.cooling = cooling
        self._cool_c = cooling.c() if cooling is not None else None
        self._h = C.c_void_p()
        self.device = device
        self.n = 0
        check(_lib.load().sphx_plan_create(device, self.mode, C.byref(self.params),
                                           C.byref(self._cool_c) if self._cool_c else None,
                                           int(capacity), C.byref(self._h)), "sphx_plan_create")

    def close(self):
        if self._h:
            _lib.load().sphx_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_gravity_config(self, ipl: int, tj: int, js: int = 1):
        check(_lib.load().sphx_plan_set_gravity_config(self._h, ipl, tj, js), "gravity config")
        self._gcfg_set = True

    def set_cell_level(self, bias: int):
        """Shift of the bucket cut level of the neighbour passes (-3..3); results
        are the same at any level."""
        check(_lib.load().sphx_plan_set_cell_level(self._h, int(bias)), "set_cell_level")
        self._cell_set = True

    def set_neighbor_lists(self, on: bool = True):
        """Neighbour lists for the div/curl and hydro passes (default on); the
        candidate walks when off.  Results are the same bits either way."""
        check(_lib.load().sphx_plan_set_neighbor_lists(self._h, int(on)), "set_neighbor_lists")

    def load(self, pos, vel, u, m, h, eps=None):
        t = _lib.torch()
        _lib.check_run_arrays(pos, vel, u, m, h, eps)
        if self.mode == _lib.MIXED and not getattr(self, "_gcfg_set", False):
            from .autotune import choose_gravity_config
            self.set_gravity_config(*choose_gravity_config(len(m)))
        if self.mode == _lib.MIXED and not getattr(self, "_cell_set", False):
            from .autotune import choose_cell_level
            check(_lib.load().sphx_plan_set_cell_level(self._h, choose_cell_level(len(m))),
                  "set_cell_level")
        self._bufs = [dev(np.asarray(a, np.float64) if not isinstance(a, t.Tensor) else a,
                          t.float64) for a in (pos, vel, u, m, h)]
        de = dev(np.asarray(eps, np.float64), t.float64) if eps is not None else None
        self.n = int(self._bufs[3].shape[0])
        check(_lib.load().sphx_plan_load(self._h, self.n, *[ptr(b) for b in self._bufs], ptr(de),
                                         _lib.stream_handle()), "sphx_plan_load")
        self._eps = de

    def bind(self, pos, vel, u, m, h, eps=None):
        """Load from caller-owned torch tensors on the device (float64,
        contiguous; pos and vel (n, 3), u and h (n,)) and keep them as the
        run's state: every init / step reads them first and writes the state
        back into them (caller order) at its end."""
        t = _lib.torch()
        arrs = {"pos": pos, "vel": vel, "u": u, "m": m, "h": h}
        for k, a in arrs.items():
            if not (isinstance(a, t.Tensor) and a.is_cuda and a.dtype == t.float64
                    and a.is_contiguous()):
                raise TypeError(f"bind needs contiguous float64 device tensors ({k})")
        _lib.check_run_arrays(pos, vel, u, m, h, eps)
        if self.mode == _lib.MIXED and not getattr(self, "_gcfg_set", False):
            from .autotune import choose_gravity_config
            self.set_gravity_config(*choose_gravity_config(len(m)))
        if self.mode == _lib.MIXED and not getattr(self, "_cell_set", False):
            from .autotune import choose_cell_level
            check(_lib.load().sphx_plan_set_cell_level(self._h, choose_cell_level(len(m))),
                  "set_cell_level")
        de = dev(np.asarray(eps, np.float64), t.float64) if eps is not None else None
        self._bufs = [pos, vel, u, m, h]
        self._eps = de
        self.n = int(m.shape[0])
        check(_lib.load().sphx_plan_bind(self._h, self.n, *[ptr(b) for b in self._bufs], ptr(de),
                                         _lib.stream_handle()), "sphx_plan_bind")

    def init(self):
        check(_lib.load().sphx_plan_init(self._h, _lib.stream_handle()), "sphx_plan_init")

    def save(self, path):
        """Exact checkpoint of the resident run (restored runs continue bit-identically)."""
        check(_lib.load().sphx_plan_save(self._h, os.fsencode(path), _lib.stream_handle()),
              "sphx_plan_save")

    def restore(self, path):
        """Replace the plan's run by a checkpointed one (same mode, capacity >= N)."""
        check(_lib.load().sphx_plan_restore(self._h, os.fsencode(path), _lib.stream_handle()),
              "sphx_plan_restore")
        self.n = checkpoint_info(path)["n"]
        self._gcfg_set = True

    def set_cooling_grid(self, grid):
        """Attach a libsx.CoolingGrid; step(..., cooling="table") uses it."""
        g = grid.c()
        check(_lib.load().sphx_plan_set_cooling_grid(self._h, C.byref(g)), "set_cooling_grid")

    def set_chimes_table(self, table):
        """Attach a libsx.ChimesTable; step(..., cooling="chimes") uses it."""
        c = table.c()
        check(_lib.load().sphx_plan_set_chimes_table(self._h, C.byref(c)), "set_chimes_table")
        self._n_species = table.n_species

    def set_cooling_network(self, network):
        """Attach a libsx.CoolingNetwork; step(..., cooling="network") uses it."""
        c = network.c()
        check(_lib.load().sphx_plan_set_cooling_network(self._h, C.byref(c)),
              "set_cooling_network")

    def chimes_ion_fractions(self):
        """Species fractions of every particle at its last CHIMES lookup,
        (n, n_species) float32 in the original particle order (the ionFrac
        array of hfvGpuSphTyp_v7.cu)."""
        t = _lib.torch()
        out = t.empty((self.n, getattr(self, "_n_species", 0)), dtype=t.float32, device="cuda")
        check(_lib.load().sphx_plan_chimes_ions(self._h, ptr(out), _lib.stream_handle()),
              "chimes_ions")
        return out.cpu().numpy()

    def step(self, n_steps: int, dt: float, cooling=False) -> int:
        """cooling: False, True / "solve" (implicit Lambda(T) solve), "table"
        (2-D table), "chimes" (5-D CHIMES table) or "network" (the analytic H/He
        network)."""
        return check(_lib.load().sphx_plan_step(self._h, int(n_steps), float(dt),
                                                _lib.cooling_mode(cooling), _lib.stream_handle()),
                     "sphx_plan_step")

    def state(self, fields=("pos", "vel", "u", "h", "rho", "acc", "acc_g", "dudt", "k")) -> dict:
        """Copy the state out (original particle order) as numpy arrays."""
        t = _lib.torch()
        n = self.n
        shapes = {"pos": (n, 3), "vel": (n, 3), "u": (n,), "h": (n,), "rho": (n,),
                  "acc": (n, 3), "acc_g": (n, 3), "dudt": (n,), "k": (n,)}
        bufs = {}
        for f in ("pos", "vel", "u", "h", "rho", "acc", "acc_g", "dudt", "k"):
            if f in fields:
                bufs[f] = t.empty(shapes[f], dtype=t.int32 if f == "k" else t.float64,
                                  device="cuda")
        order = ["pos", "vel", "u", "h", "rho", "acc", "acc_g", "dudt", "k"]
        check(_lib.load().sphx_plan_store(self._h, *[ptr(bufs.get(f)) for f in order],
                                          _lib.stream_handle()), "sphx_plan_store")
        return {f: b.cpu().numpy() for f, b in bufs.items()}

    def counters(self) -> dict:
        c = _lib.Counters()
        check(_lib.load().sphx_plan_query(self._h, C.byref(c)), "sphx_plan_query")
        return c.as_dict()

    def set_graph(self, on: bool = True):
        """Replay steps from captured hipGraphs (one per pointer parity)."""
        check(_lib.load().sphx_plan_set_graph(self._h, int(on)), "set_graph")

    def set_gravity_bh(self, theta: float):
        """Barnes-Hut gravity in the step (theta > 0); 0 restores the direct sum."""
        check(_lib.load().sphx_plan_set_gravity_bh(self._h, float(theta)), "set_gravity_bh")

    def set_block_steps(self, levels: int, c_cfl: float = 0.25):
        """Block time steps with `levels` halvings of the longest step (0: off).
        Then step(n, dt) takes n substeps of dt / 2**levels; particle i steps
        with dt / 2**b_i chosen from its own criterion (before the first step
        after init)."""
        check(_lib.load().sphx_plan_set_block_steps(self._h, int(levels), float(c_cfl)),
              "set_block_steps")

    def block_bins(self):
        """Level b_i of every particle (original order)."""
        t = _lib.torch()
        out = t.empty(self.n, dtype=t.int32, device="cuda")
        check(_lib.load().sphx_plan_block_bins(self._h, ptr(out), _lib.stream_handle()),
              "block_bins")
        return out.cpu().numpy()

    def set_adaptive_dt(self, on: bool = True, c_cfl: float = 0.25, dt_min: float = 1e-5,
                        dt_max: float = 1e-3):
        """Global adaptive dt (hfv_sph_gpu_v2_evrard.cu:517-561) evaluated on the
        device after every step; the next ``step``'s dt seeds the first step."""
        check(_lib.load().sphx_plan_set_adaptive_dt(self._h, int(on), c_cfl, dt_min, dt_max),
              "set_adaptive_dt")

    def time(self) -> tuple[float, float, tuple[float, float, float, float]]:
        """(simulated time, next dt, (dt_f, dt_kin, dt_cfl, dt_dens))."""
        t = C.c_double()
        dt = C.c_double()
        parts = (C.c_double * 4)()
        check(_lib.load().sphx_plan_time(self._h, C.byref(t), C.byref(dt), parts,
                                         _lib.stream_handle()), "time")
        return t.value, dt.value, tuple(parts)

    def set_timing(self, on: bool = True):
        check(_lib.load().sphx_plan_set_timing(self._h, int(on)), "set_timing")

    def timing(self) -> dict:
        """Summed milliseconds and step counts per phase since set_timing."""
        ms = (C.c_double * 8)()
        cnt = (C.c_int64 * 8)()
        check(_lib.load().sphx_plan_timing(self._h, ms, cnt, 8), "timing")
        return {ph: (float(ms[k]), int(cnt[k])) for k, ph in enumerate(_lib.PHASES)}

    def diagnostics(self, with_pe: bool = True) -> dict:
        d = _lib.Diag()
        check(_lib.load().sphx_plan_diagnostics(self._h, int(with_pe), C.byref(d),
                                                _lib.stream_handle()), "diagnostics")
        return d.as_dict()
