"""ctypes binding of libsphx.so (the HIP library) and device-buffer helpers.

The library must have been built (``python -m sph_2_amd.build``); a missing
library or a missing GPU is an error, never a silent CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphx.so")

MIXED = 0
FP64 = 1
OK = 0
CAPPED = 1
REFUSED = 2

_lib = None
_lock = threading.Lock()


class SphxError(RuntimeError):
    pass


class Params(C.Structure):
    """sphx_params (include/sphx.h)."""
    _fields_ = [("gamma", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("eta", C.c_double), ("G", C.c_double), ("nngb", C.c_int32), ("ndown", C.c_int32),
                ("nup", C.c_int32), ("max_iter", C.c_int32), ("av_form", C.c_int32),
                ("balsara", C.c_int32), ("eps_from_h", C.c_int32), ("h_predictor", C.c_int32),
                ("h_mode", C.c_int32), ("knn_kth", C.c_int32), ("av_w_eps", C.c_double),
                ("h_min", C.c_double)]


class Cooling(C.Structure):
    """sphx_cooling (include/sphx.h)."""
    _fields_ = [("unit_u_cgs", C.c_double), ("unit_rho_cgs", C.c_double),
                ("unit_time_s", C.c_double), ("XH", C.c_double), ("mu", C.c_double),
                ("gamma", C.c_double), ("heat", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int32), ("max_bracket", C.c_int32), ("ntab", C.c_int32),
                ("pad", C.c_int32), ("T", C.POINTER(C.c_double)), ("L", C.POINTER(C.c_double))]


class Network(C.Structure):
    """sphx_network (include/sphx.h)."""
    _fields_ = [("XH", C.c_double), ("J21", C.c_double), ("unit_u_cgs", C.c_double),
                ("unit_rho_cgs", C.c_double), ("unit_time_s", C.c_double), ("tol", C.c_double),
                ("n_grid", C.c_int32), ("max_iter", C.c_int32), ("max_bracket", C.c_int32),
                ("pad", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [("n", C.c_int64), ("steps", C.c_int64), ("h_iters_max", C.c_int64),
                ("h_iters_sum", C.c_int64), ("h_capped", C.c_int64), ("cool_capped", C.c_int64),
                ("cand_buckets", C.c_int64), ("buckets", C.c_int64), ("bh_overflow", C.c_int64),
                ("nonfinite", C.c_int64), ("bin_floor", C.c_int64),
                ("cool_iters", C.c_int64), ("cool_bisected", C.c_int64),
                ("nbr_overflow", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class ParamsGPU(C.Structure):
    """The positional params.GPU file of the reference's GPU drivers."""
    _fields_ = [("N", C.c_int64)] + [(k, C.c_double) for k in (
        "c_0", "gamma", "Rcld_in_pc", "Rcld_in_cm", "Mcld_in_g", "mu", "Mach",
        "grav_const_in_cgs", "G")]

    def as_dict(self):
        return {k: (int(getattr(self, k)) if k == "N" else float(getattr(self, k)))
                for k, _ in self._fields_}


class CoolGrid(C.Structure):
    _fields_ = [("nu", C.c_int64), ("nrho", C.c_int64), ("u_nodes", C.c_void_p),
                ("rho_nodes", C.c_void_p), ("du", C.c_void_p), ("dt_ref_s", C.c_double),
                ("unit_u_cgs", C.c_double), ("unit_rho_cgs", C.c_double),
                ("unit_time_s", C.c_double)]


class ChimesTableC(C.Structure):
    _fields_ = [("n_kpc", C.c_int32), ("n_nh", C.c_int32), ("n_z", C.c_int32),
                ("n_t", C.c_int32), ("n_m", C.c_int32), ("n_time", C.c_int32),
                ("z_index", C.c_int32), ("pad", C.c_int32),
                ("kpc", C.c_void_p), ("nh", C.c_void_p), ("time_s", C.c_void_p),
                ("u_evol", C.c_void_p), ("mu", C.c_void_p), ("metalz", C.c_void_p),
                ("xh", C.c_double), ("m_h", C.c_double), ("u_max_cgs", C.c_double),
                ("kpc_cm", C.c_double), ("unit_length_cm", C.c_double),
                ("unit_rho_cgs", C.c_double), ("unit_u_cgs", C.c_double),
                ("unit_time_s", C.c_double)]


class Diag(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("ke", "ut", "pe", "px", "py", "pz", "lx", "ly", "lz",
                                          "mass")]

    def as_dict(self):
        return {k: float(getattr(self, k)) for k, _ in self._fields_}


# Every exported entry point with its argument types (pointers are void*).
_V = C.c_void_p
_S = C.c_char_p
_I64 = C.c_int64
_I32 = C.c_int32
_D = C.c_double
SIGNATURES = {
    "sphx_last_error": (C.c_char_p, []),
    "sphx_version": (C.c_char_p, []),
    "sphx_default_params": (None, [_V]),
    "sphx_ctx_create": (C.c_int, [C.c_int, _V]),
    "sphx_ctx_destroy": (C.c_int, [_V]),
    "sphx_getDensity": (C.c_int, [_V, C.c_int, _I64, _V, _V, _V, _I64, _I64, _V, _V, _V]),
    "sphx_eos": (C.c_int, [_V, _I64, _V, _V, _D, _V, _V, _V]),
    "sphx_smoothing": (C.c_int, [_V, C.c_int, _I64, _V, _V, _V, _V, _D, _V, _I64, _I64, _V, _V,
                                 _V, _V]),
    "sphx_div_curlVel": (C.c_int, [_V, C.c_int, _I64, _V, _V, _V, _V, _V, _I64, _I64, _V, _V,
                                   _V]),
    "sphx_getAcc_sph": (C.c_int, [_V, C.c_int, _I64, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V,
                                  _I64, _I64, _V, _V, _V]),
    "sphx_getAcc_g_smth": (C.c_int, [_V, C.c_int, _I64, _V, _V, _V, _D, _I64, _I64, _V, _V,
                                     _V]),
    "sphx_pair_issue_rate": (C.c_int, [_V, _I32, _V, _V]),
    "sphx_bh_gravity": (C.c_int, [_V, _I64, _V, _V, _V, _D, _D, _I64, _I64, _V, _V, _V, _V]),
    "sphx_neighbors": (C.c_int, [_V, _I64, _V, _V, _I64, _I64, _V, _V, _I64, _V, _V]),
    "sphx_cool": (C.c_int, [_V, _I64, _V, _V, _D, _V, _V, _V, _V]),
    "sphx_network_rates": (C.c_int, [_V, _V]),
    "sphx_cool_network": (C.c_int, [_V, _I64, _V, _V, _D, _V, _V, _V, _V]),
    "sphx_plan_set_cooling_network": (C.c_int, [_V, _V]),
    "sphx_batch_set_cooling_network": (C.c_int, [_V, _V]),
    "sphx_diagnostics": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _D, _V, _V]),
    "sphx_global_dt": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _D, _D, _D, _V, _V]),
    "sphx_global_dt_noalloc": (C.c_int, [C.c_int, _I64, _V, _V, _V, _V, _V, _D, _D, _D, _V,
                                         _V]),
    "sphx_knn_h": (C.c_int, [_V, _I64, _V, _V, _I32, _I64, _I64, _V, _V]),
    "sphx_morton_sort": (C.c_int, [_V, _I64, _V, _V, _V, _V]),
    "sphx_plan_create": (C.c_int, [C.c_int, C.c_int, _V, _V, _I64, _V]),
    "sphx_plan_load": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_plan_bind": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_plan_init": (C.c_int, [_V, _V]),
    "sphx_plan_step": (C.c_int, [_V, _I32, _D, _I32, _V]),
    "sphx_plan_store": (C.c_int, [_V, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_plan_query": (C.c_int, [_V, _V]),
    "sphx_plan_diagnostics": (C.c_int, [_V, _I32, _V, _V]),
    "sphx_plan_destroy": (C.c_int, [_V]),
    "sphx_plan_set_gravity_config": (C.c_int, [_V, _I32, _I32, _I32]),
    "sphx_ctx_set_gravity_config": (C.c_int, [_V, _I32, _I32, _I32]),
    "sphx_plan_set_timing": (C.c_int, [_V, _I32]),
    "sphx_plan_set_neighbor_lists": (C.c_int, [_V, _I32]),
    "sphx_plan_set_adaptive_dt": (C.c_int, [_V, _I32, _D, _D, _D]),
    "sphx_plan_set_gravity_bh": (C.c_int, [_V, _D]),
    "sphx_plan_set_cooling_grid": (C.c_int, [_V, _V]),
    "sphx_batch_set_cooling_grid": (C.c_int, [_V, _V]),
    "sphx_cool_grid_apply": (C.c_int, [_V, _I64, _V, _V, _D, _V, _V, _V]),
    "sphx_chimes_apply": (C.c_int, [_V, _I64, _V, _V, _V, _D, _V, _V, _V, _V, _V]),
    "sphx_plan_set_chimes_table": (C.c_int, [_V, _V]),
    "sphx_plan_chimes_ions": (C.c_int, [_V, _V, _V]),
    "sphx_batch_set_chimes_table": (C.c_int, [_V, _V]),
    "sphx_batch_set_gravity_symmetric": (C.c_int, [_V, _I32, _I64]),
    "sphx_gravity_symmetric_bytes": (C.c_int64, [_I64, _I32]),
    "sphx_gravity_symmetric_min_bytes": (C.c_int64, [_I64, _I32]),
    "sphx_plan_set_cell_level": (C.c_int, [_V, _I32]),
    "sphx_plan_set_block_steps": (C.c_int, [_V, _I32, _D]),
    "sphx_plan_block_bins": (C.c_int, [_V, _V, _V]),
    "sphx_batch_set_cell_level": (C.c_int, [_V, _I32]),
    "sphx_read_chimes": (C.c_int, [_S, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_write_chimes": (C.c_int, [_S, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_cool_grid_build": (C.c_int, [_V, _V, _V, _V]),
    "sphx_plan_time": (C.c_int, [_V, _V, _V, _V, _V]),
    "sphx_plan_timing": (C.c_int, [_V, _V, _V, _I32]),
    "sphx_plan_set_graph": (C.c_int, [_V, _I32]),
    "sphx_batch_create": (C.c_int, [C.c_int, C.c_int, _V, _V, _I64, _I32, _V]),
    "sphx_batch_admit": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _V, _D, _I32, _V, _V]),
    "sphx_batch_retire": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_batch_store": (C.c_int, [_V, _I64, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_batch_step": (C.c_int, [_V, _I32, _V]),
    "sphx_batch_runs": (C.c_int, [_V, _V, _V, _V, _I32, _V]),
    "sphx_batch_set_adaptive_dt": (C.c_int, [_V, _I64, _I32, _D, _D, _D, _V]),
    "sphx_read_params_gpu": (C.c_int, [_S, _V]),
    "sphx_write_params_gpu": (C.c_int, [_S, _V]),
    "sphx_read_ic_csv": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_write_restart_csv": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_write_snapshot_csv": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V]),
    "sphx_write_ic_bin": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_read_ic_bin": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_write_snapshot_bin": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_read_snapshot_bin": (C.c_int, [_S, _I64, _V, _V, _V, _V, _V, _V, _V, _V]),
    "sphx_plan_save": (C.c_int, [_V, _S, _V]),
    "sphx_plan_restore": (C.c_int, [_V, _S, _V]),
    "sphx_batch_save_run": (C.c_int, [_V, _I64, _S, _V]),
    "sphx_batch_restore_run": (C.c_int, [_V, _S, _V, _V]),
    "sphx_batch_time": (C.c_int, [_V, _I64, _V, _V, _V, _V]),
    "sphx_batch_set_graph": (C.c_int, [_V, _I32]),
    "sphx_batch_query": (C.c_int, [_V, _V]),
    "sphx_batch_bytes_per_particle": (_I64, []),
    "sphx_batch_destroy": (C.c_int, [_V]),
}

PHASES = ("sort", "smoothing", "density", "gravity", "divcurl", "hydro", "finish",
          "gravity_kernel")


def load(path: str = LIB_PATH):
    """Load libsphx.so and declare every entry point; raises if it is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise SphxError(f"{path} is missing: build it with `python -m sph_2_amd.build`")
        # torch first: its HIP runtime is then the one libsphx binds to (by
        # soname).  Loaded the other way round, the process would hold two HIP
        # runtimes, and the library's would see no device.
        torch()
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = load().sphx_last_error().decode()
        raise SphxError(f"{what} failed ({rc}): {msg}")
    return rc


def cooling_mode(cooling) -> int:
    """False/0: none; True/1/"solve": implicit solve; 2/"table": 2-D table;
    3/"chimes": 5-D CHIMES table; 4/"network": analytic H/He network."""
    if cooling in ("network", 4):
        return 4
    if cooling in ("chimes", 3):
        return 3
    if cooling in ("table", 2):
        return 2
    if cooling in ("solve", True, 1):
        return 1
    if cooling in (False, None, 0):
        return 0
    raise ValueError(f"unknown cooling mode {cooling!r}")


def default_params(**kw) -> Params:
    p = Params()
    load().sphx_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------------------
# Device buffers (torch-ROCm tensors used only as memory)
# ---------------------------------------------------------------------------
def torch():
    import torch as _t
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise SphxError("no HIP device visible: libsphx runs on the GPU only")
    return t


def check_run_arrays(pos, vel, u, m, h, eps=None) -> int:
    """Shapes of one run's particle arrays (numpy or torch): pos and vel
    (n, 3), u, m, h and eps (n,).  The library reads n rows of each, so a
    shorter array would be read past its end; returns n."""
    n = len(m)
    for name, a, shape in (("pos", pos, (n, 3)), ("vel", vel, (n, 3)), ("u", u, (n,)),
                           ("m", m, (n,)), ("h", h, (n,)), ("eps", eps, (n,))):
        if a is not None and tuple(np.shape(a) if not hasattr(a, "shape") else a.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}, got {tuple(a.shape)}")
    return n


def dev(a, dtype=None):
    """numpy / torch -> contiguous device tensor."""
    t = require_gpu()
    if isinstance(a, t.Tensor):
        x = a
        if dtype is not None and x.dtype != dtype:
            x = x.to(dtype)
        if not x.is_cuda:
            x = x.cuda()
        return x.contiguous()
    arr = np.ascontiguousarray(a)
    if not arr.flags.writeable:
        arr = arr.copy()
    x = t.from_numpy(arr)
    if dtype is not None:
        x = x.to(dtype)
    return x.cuda().contiguous()


def ptr(x) -> int | None:
    return None if x is None else x.data_ptr()


def stream_handle():
    t = torch()
    return t.cuda.current_stream().cuda_stream
