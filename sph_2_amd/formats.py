"""The reference's on-disk formats, read and written by the library's C++ I/O.

Names follow the reference's own readers and writers (paths relative to
11_Dec_2022/ of hassanfv/SPH_2):

* ``read_params_gpu`` / ``write_params_gpu``: the positional ``params.GPU``
  (GPU_sph/Important_Test_Simulations/Evrard_collapse/hfv_sph_gpu_v2_evrard.cu:38-60,
  IC/3.py:66-73);
* ``read_ic_csv``: IC or restart CSV rows x,y,z,vx,vy,vz,m,h,eps,u[,dudt]
  (hfv_sph_gpu_v2_evrard.cu:75-131; Check_Restart_version/hfv_sph_gpu_v5.cu:348-396);
* ``write_ic_csv`` / ``write_restart_csv``: the same rows (IC/3.py:51-60;
  hfv_sph_gpu_v5.cu:745-753);
* ``write_snapshot_csv``: header ``x,y,z,vx,vy,vz,h,rho,u`` (hfv_sph_gpu_v2_evrard.cu:495-514);
* ``read_ic_binary`` / ``write_ic_binary``: int32 Typ then float32 arrays
  (GPU_sph_Type/Evrard_collapse/IC_evrard_binary.py:60-82, hfvCppLibs.h:8-63);
* ``read_snapshot_binary`` / ``write_snapshot_binary``: int32 N, Typ, float32
  x,y,z,vx,vy,vz,rho,h,u,mass (hfvCppLibs.h:68-96);
* ``read_chimes_table`` / ``write_chimes_table``: coolHeatGridZ.bin, the 5-D
  CHIMES table (GPU_sph_Type_cooling_II/cooling_grids_use_EC2/step_2x.py:101-140
  writes it, hfvGpuSphTyp_v7.cu:35-79 reads it).

All arrays are float64 numpy arrays in and out (binary files hold float32, as
the reference's do).  Errors raise ``SphxError`` with the file and the reason.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _path(path) -> bytes:
    return os.fsencode(path)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def read_params_gpu(path) -> dict:
    out = _lib.ParamsGPU()
    check(_lib.load().sphx_read_params_gpu(_path(path), C.byref(out)), "read_params_gpu")
    return out.as_dict()


def write_params_gpu(path, **values):
    p = _lib.ParamsGPU()
    for k, _ in p._fields_:
        setattr(p, k, values.get(k, 0))
    check(_lib.load().sphx_write_params_gpu(_path(path), C.byref(p)), "write_params_gpu")


def _ic_arrays(pos, vel, m, h, eps, u):
    pos = _f64(pos)
    n = len(pos)
    return (n, pos, _f64(vel, (n, 3)), _f64(m, (n,)), _f64(h, (n,)),
            _f64(np.broadcast_to(eps, (n,))), _f64(u, (n,)))


def read_ic_csv(path, N: int | None = None, restart: bool = False) -> dict:
    """IC (10 columns) or restart (11 columns, ``restart=True`` returns dudt)."""
    L = _lib.load()
    rows = C.c_int64()
    if N is None:
        check(L.sphx_read_ic_csv(_path(path), 0, None, None, None, None, None, None, None,
                                 C.byref(rows)), "read_ic_csv")
        N = rows.value
    o = {"pos": np.zeros((N, 3)), "vel": np.zeros((N, 3))}
    for k in ("m", "h", "eps", "u") + (("dudt",) if restart else ()):
        o[k] = np.zeros(N)
    check(L.sphx_read_ic_csv(_path(path), N, _p(o["pos"]), _p(o["vel"]), _p(o["m"]), _p(o["h"]),
                             _p(o["eps"]), _p(o["u"]), _p(o.get("dudt")), C.byref(rows)),
          "read_ic_csv")
    return o


def write_ic_csv(path, pos, vel, m, h, eps, u):
    n, *a = _ic_arrays(pos, vel, m, h, eps, u)
    check(_lib.load().sphx_write_restart_csv(_path(path), n, *[_p(x) for x in a], None),
          "write_ic_csv")


def write_restart_csv(path, pos, vel, m, h, eps, u, dudt):
    n, *a = _ic_arrays(pos, vel, m, h, eps, u)
    check(_lib.load().sphx_write_restart_csv(_path(path), n, *[_p(x) for x in a],
                                             _p(_f64(dudt, (n,)))), "write_restart_csv")


def write_snapshot_csv(path, pos, vel, h, rho, u):
    pos = _f64(pos)
    n = len(pos)
    check(_lib.load().sphx_write_snapshot_csv(_path(path), n, _p(pos), _p(_f64(vel, (n, 3))),
                                              _p(_f64(h, (n,))), _p(_f64(rho, (n,))),
                                              _p(_f64(u, (n,)))), "write_snapshot_csv")


def write_ic_binary(path, pos, vel, m, h, eps, u, typ=None):
    n, *a = _ic_arrays(pos, vel, m, h, eps, u)
    t = None if typ is None else np.ascontiguousarray(typ, dtype=np.int32)
    check(_lib.load().sphx_write_ic_bin(_path(path), n, _p(t), *[_p(x) for x in a]),
          "write_ic_binary")


def read_ic_binary(path, N: int) -> dict:
    o = {"typ": np.zeros(N, np.int32), "pos": np.zeros((N, 3)), "vel": np.zeros((N, 3))}
    for k in ("m", "h", "eps", "u"):
        o[k] = np.zeros(N)
    check(_lib.load().sphx_read_ic_bin(_path(path), N, _p(o["typ"]), _p(o["pos"]), _p(o["vel"]),
                                       _p(o["m"]), _p(o["h"]), _p(o["eps"]), _p(o["u"])),
          "read_ic_binary")
    return o


def write_snapshot_binary(path, pos, vel, rho, h, u, m, typ=None):
    pos = _f64(pos)
    n = len(pos)
    t = None if typ is None else np.ascontiguousarray(typ, dtype=np.int32)
    check(_lib.load().sphx_write_snapshot_bin(
        _path(path), n, _p(t), _p(pos), _p(_f64(vel, (n, 3))), _p(_f64(rho, (n,))),
        _p(_f64(h, (n,))), _p(_f64(u, (n,))), _p(_f64(m, (n,)))), "write_snapshot_binary")


def read_snapshot_binary(path) -> dict:
    L = _lib.load()
    n = C.c_int64()
    check(L.sphx_read_snapshot_bin(_path(path), 0, C.byref(n), None, None, None, None, None, None,
                                   None), "read_snapshot_binary")
    N = n.value
    o = {"typ": np.zeros(N, np.int32), "pos": np.zeros((N, 3)), "vel": np.zeros((N, 3))}
    for k in ("rho", "h", "u", "m"):
        o[k] = np.zeros(N)
    if N:
        check(L.sphx_read_snapshot_bin(_path(path), N, C.byref(n), _p(o["typ"]), _p(o["pos"]),
                                       _p(o["vel"]), _p(o["rho"]), _p(o["h"]), _p(o["u"]),
                                       _p(o["m"])), "read_snapshot_binary")
    return o


_CHIMES_KEYS = ("kpc", "nh", "z", "temp", "time_s", "u_evol", "mu", "metalz")


def read_chimes_table(path) -> dict:
    """coolHeatGridZ.bin: float32 axes kpc, nh (densities), z (metallicities),
    temp (temperatures), time_s and arrays u_evol, mu of shape
    (kpc, T, nH, Z, time), metalz of shape (kpc, T, nH, Z, species, time)."""
    L = _lib.load()
    dims = (C.c_int32 * 6)()
    check(L.sphx_read_chimes(_path(path), dims, None, None, None, None, None, None, None, None),
          "read_chimes_table")
    nk, nn, nz, nt, nm, ntime = (int(v) for v in dims)
    out = {"kpc": np.empty(nk, np.float32), "nh": np.empty(nn, np.float32),
           "z": np.empty(nz, np.float32), "temp": np.empty(nt, np.float32),
           "time_s": np.empty(ntime, np.float32),
           "u_evol": np.empty((nk, nt, nn, nz, ntime), np.float32),
           "mu": np.empty((nk, nt, nn, nz, ntime), np.float32),
           "metalz": np.empty((nk, nt, nn, nz, nm, ntime), np.float32)}
    check(L.sphx_read_chimes(_path(path), dims, *[_p(out[k]) for k in _CHIMES_KEYS]),
          "read_chimes_table")
    return out


def write_chimes_table(path, kpc, nh, z, temp, time_s, u_evol, mu, metalz):
    """Write coolHeatGridZ.bin from arrays shaped as read_chimes_table returns them."""
    a = {k: np.ascontiguousarray(v, dtype=np.float32)
         for k, v in zip(_CHIMES_KEYS, (kpc, nh, z, temp, time_s, u_evol, mu, metalz))}
    nk, nn, nz, nt, ntime = (len(a[k]) for k in ("kpc", "nh", "z", "temp", "time_s"))
    if a["u_evol"].shape != (nk, nt, nn, nz, ntime) or a["mu"].shape != a["u_evol"].shape:
        raise ValueError("u_evol and mu must have shape (kpc, T, nH, Z, time)")
    if a["metalz"].ndim != 6 or a["metalz"].shape[:4] != (nk, nt, nn, nz) or \
            a["metalz"].shape[5] != ntime:
        raise ValueError("metalz must have shape (kpc, T, nH, Z, species, time)")
    dims = (C.c_int32 * 6)(nk, nn, nz, nt, a["metalz"].shape[4], ntime)
    check(_lib.load().sphx_write_chimes(_path(path), dims, *[_p(a[k]) for k in _CHIMES_KEYS]),
          "write_chimes_table")
