"""ctypes bindings of the FP64 C oracle and the oracle's composition of a full step.

Each wrapper names the reference function it restates (paths relative to
11_Dec_2022/ of hassanfv/SPH_2; aliases as in sph_oracle.c).  All arrays are
float64 numpy arrays; ``rows`` selects a subset of target particles (the
reference's ``*_mpi(nbeg, nend, ...)`` row ranges generalised to index lists).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=_HERE, check=True)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        lp = C.POINTER(C.c_int64)
        i64 = C.c_int64
        i32 = C.c_int32
        L.orc_h_factor.restype = C.c_double
        L.orc_h_factor.argtypes = [i32, i32]
        L.orc_count_within.argtypes = [i64, dp, dp, lp, i64, ip]
        L.orc_smoothing.argtypes = [i64, dp, dp, ip, dp, C.c_double, i32, i32, i32, i32, lp, i64,
                                    dp, ip, ip, ip]
        L.orc_smoothing_steps.argtypes = [i64, dp, dp, ip, dp, C.c_double, i32, i32, i32, i32,
                                          i32, C.c_double, lp, i64, dp, ip, ip, ip]
        L.orc_neighbors.argtypes = [i64, dp, dp, lp, i64, lp, lp]
        L.orc_density.argtypes = [i64, dp, dp, dp, lp, i64, dp, ip]
        L.orc_eos.argtypes = [i64, dp, dp, C.c_double, dp, dp]
        L.orc_divcurl.argtypes = [i64, dp, dp, dp, dp, dp, lp, i64, dp, dp]
        L.orc_hydro.argtypes = [i64, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_void_p, lp, i64, dp, dp]
        L.orc_gravity.argtypes = [i64, dp, dp, dp, C.c_double, lp, i64, dp, dp]
        L.orc_lambda.restype = C.c_double
        L.orc_lambda.argtypes = [C.c_double, dp, dp, i32]
        L.orc_cool_one.restype = C.c_double
        L.orc_cool_one.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p, dp, dp, ip]
        L.orc_cool.argtypes = [i64, dp, dp, C.c_double, C.c_void_p, dp, dp, dp, ip]
        L.orc_ke.restype = C.c_double
        L.orc_ke.argtypes = [i64, dp, dp]
        L.orc_pe.restype = C.c_double
        L.orc_pe.argtypes = [i64, dp, dp, dp, C.c_double]
        L.orc_cool_grid_apply.argtypes = [i64, dp, dp, C.c_double, i64, dp, i64, dp, dp,
                                          C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.orc_cool_grid_build.argtypes = [i64, dp, i64, dp, C.c_double, C.c_void_p, dp, dp, dp]
        L.orc_knn_h.argtypes = [i64, dp, i64, lp, i64, dp]
        fp = C.POINTER(C.c_float)
        L.orc_chimes_apply.argtypes = [i64, dp, dp, dp, C.c_double, C.c_void_p, fp, fp, fp, fp,
                                       fp, fp, dp, fp, dp]
        L.orc_net_rates.argtypes = [C.c_void_p, dp]
        L.orc_net_ci.argtypes = [C.c_double, dp]
        L.orc_net_abundance.argtypes = [C.c_double, C.c_double, dp, dp, ip]
        L.orc_net_heat_cool.argtypes = [C.c_double, C.c_double, dp, dp]
        L.orc_net_u_to_temp.restype = C.c_double
        L.orc_net_u_to_temp.argtypes = [C.c_double, C.c_double, C.c_double, dp, ip]
        L.orc_net_temp_to_u.restype = C.c_double
        L.orc_net_temp_to_u.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_net_rate.restype = C.c_double
        L.orc_net_rate.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.orc_net_cool_one.restype = C.c_double
        L.orc_net_cool_one.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p, dp, ip, ip]
        L.orc_net_cool.argtypes = [i64, dp, dp, C.c_double, C.c_void_p, dp, dp, ip, ip]
        L.orc_net_sigma.restype = C.c_double
        L.orc_net_sigma.argtypes = [C.c_double] * 8
        L.orc_global_dt.restype = C.c_double
        L.orc_global_dt.argtypes = [i64, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, dp]
        _LIB = L
    return _LIB


class _Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("eta", C.c_double), ("G", C.c_double), ("nngb", C.c_int32), ("ndown", C.c_int32),
                ("nup", C.c_int32), ("max_iter", C.c_int32), ("av_form", C.c_int32),
                ("balsara", C.c_int32), ("av_w_eps", C.c_double)]


class _CoolParams(C.Structure):
    _fields_ = [("unit_u_cgs", C.c_double), ("unit_rho_cgs", C.c_double),
                ("unit_time_s", C.c_double), ("XH", C.c_double), ("mu", C.c_double),
                ("gamma", C.c_double), ("heat", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int32), ("max_bracket", C.c_int32), ("ntab", C.c_int32),
                ("pad", C.c_int32)]


@dataclass
class Params:
    """Physics constants of the step (AUGX:472-484, hfv_sph_gpu_v2_evrard.cu:283-300)."""
    gamma: float = 5.0 / 3.0
    alpha: float = 1.0
    beta: float = 2.0
    eta: float = 0.1
    G: float = 1.0
    nngb: int = 64
    ndown: int = 59
    nup: int = 69
    max_iter: int = 100
    av_form: int = 0          # 0 Monaghan-Gingold 1983, 1 Gadget signal velocity
    balsara: int = 0
    eps_from_h: bool = True   # softening eps_i = h_i
    h_predictor: bool = True  # Gadget-2 eq. 31 start for the h iteration
    h_mode: int = 0           # 0 count-band iteration, 1 kNN (do_smoothingX), 2 the
                              # reference's fixed-step rule (LIBSX3:947-1006), 3 the
                              # same with ten times the step (do_smoothingX4), 4 the
                              # step-halving rule of the LIBV7 driver, 5 the relative
                              # steps of h_smooth_fast (LIBSX:1133-1184)
    knn_kth: int = 64
    av_w_eps: float = 1e-5    # Gadget AV: w = v.r / (r + av_w_eps)
    h_min: float = 0.0        # h_mode 5: floor on h after eleven steps (LIBSX:1192-1245)

    def c(self) -> _Params:
        return _Params(self.gamma, self.alpha, self.beta, self.eta, self.G, self.nngb, self.ndown,
                       self.nup, self.max_iter, self.av_form, self.balsara, self.av_w_eps)


@dataclass
class CoolParams:
    unit_u_cgs: float
    unit_rho_cgs: float
    unit_time_s: float
    XH: float = 0.76
    mu: float = 0.62
    gamma: float = 5.0 / 3.0
    heat: float = 0.0
    tol: float = 1e-10
    max_iter: int = 100
    max_bracket: int = 200
    T: np.ndarray = field(default=None)
    L: np.ndarray = field(default=None)

    def c(self) -> _CoolParams:
        return _CoolParams(self.unit_u_cgs, self.unit_rho_cgs, self.unit_time_s, self.XH, self.mu,
                           self.gamma, self.heat, self.tol, self.max_iter, self.max_bracket,
                           len(self.T), 0)


class _NetParams(C.Structure):
    _fields_ = [("XH", C.c_double), ("J21", C.c_double), ("unit_u_cgs", C.c_double),
                ("unit_rho_cgs", C.c_double), ("unit_time_s", C.c_double), ("tol", C.c_double),
                ("n_grid", C.c_int32), ("max_iter", C.c_int32), ("max_bracket", C.c_int32),
                ("pad", C.c_int32)]


@dataclass
class NetParams:
    """The analytic H/He network of photolibs3.py (GPU_sph/Cooling_v1/photolibs3.py:140-543):
    hydrogen mass fraction, radiation-field strength (its Jcoeff; 1e-5 there, 50 in
    myPhotoLibsGPU.h), frequency grid (0: np.arange with step 0.02e16 Hz as photolibs3.py;
    n > 0: n points per edge as myPhotoLibsGPU.h), code -> cgs unit factors, and the solve's
    relative bracket width (1e-4) and caps."""
    XH: float = 0.76
    J21: float = 1e-5
    unit_u_cgs: float = 1.0
    unit_rho_cgs: float = 1.0
    unit_time_s: float = 1.0
    tol: float = 1e-4
    n_grid: int = 0
    max_iter: int = 100
    max_bracket: int = 400

    def c(self) -> _NetParams:
        return _NetParams(self.XH, self.J21, self.unit_u_cgs, self.unit_rho_cgs,
                          self.unit_time_s, self.tol, self.n_grid, self.max_iter,
                          self.max_bracket, 0)


def net_rates(p: NetParams):
    """RadiationField (photolibs3.py:230-329): (gJH0, gJHe0, gJHep, HRate_H0, HRate_He0,
    HRate_Hep)."""
    out = np.zeros(6)
    c = p.c()
    lib().orc_net_rates(C.byref(c), _d(out)[1])
    return out


def net_ci(T):
    """RandCIRates (photolibs3.py:140-166)."""
    out = np.zeros(7)
    lib().orc_net_ci(float(T), _d(out)[1])
    return out


def net_abundance(T, nH, rf):
    """Abundance_hX (photolibs3.py:169-211): (nH0, nHe0, nHp, ne, nHep, nHepp), iterations."""
    out = np.zeros(6)
    it = np.zeros(1, np.int32)
    rf_, rfp = _d(rf)
    lib().orc_net_abundance(float(T), float(nH), rfp, _d(out)[1], _i(it)[1])
    return out, int(it[0])


def net_heat_cool(T, nH, rf):
    """coolingHeatingRates (photolibs3.py:332-376): (heating, Lambda) per nH^2."""
    out = np.zeros(2)
    rf_, rfp = _d(rf)
    lib().orc_net_heat_cool(float(T), float(nH), rfp, _d(out)[1])
    return out


def net_u_to_temp(u, nH, XH, rf):
    """convert_u_to_temp_h (photolibs3.py:379-422)."""
    rf_, rfp = _d(rf)
    return lib().orc_net_u_to_temp(float(u), float(nH), float(XH), rfp, C.POINTER(C.c_int32)())


def net_temp_to_u(T, nH, XH, rf):
    """convert_Temp_to_u (photolibs3.py:425-467)."""
    rf_, rfp = _d(rf)
    return lib().orc_net_temp_to_u(float(T), float(nH), float(XH), rfp)


def net_rate(u, nH, XH, rf):
    """coolingRateFromU_h (photolibs3.py:470-486): heating - Lambda per nH^2 at T(u)
    clamped to [1e4, 1e8] K."""
    rf_, rfp = _d(rf)
    return lib().orc_net_rate(float(u), float(nH), float(XH), rfp)


def net_cool(rho, u, dt, p: NetParams, rf=None):
    """DoCooling_h (photolibs3.py:489-543) on arrays in the code units of p:
    returns (u_new, bisection iterations, status 0 ok / 1 bisection cap / 2 bracket cap)."""
    rho, rp = _d(np.atleast_1d(rho))
    u, up = _d(np.atleast_1d(u))
    rf = net_rates(p) if rf is None else rf
    rf_, rfp = _d(rf)
    n = len(u)
    out = np.zeros(n)
    it = np.zeros(n, np.int32)
    st = np.zeros(n, np.int32)
    c = p.c()
    lib().orc_net_cool(n, rp, up, float(dt), C.byref(c), rfp, _d(out)[1], _i(it)[1], _i(st)[1])
    return out, it, st


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _rows(rows, n):
    if rows is None:
        return None, None, n
    r = np.ascontiguousarray(rows, dtype=np.int64)
    return r, r.ctypes.data_as(C.POINTER(C.c_int64)), len(r)


def h_factor(nngb: int, k: int) -> float:
    return lib().orc_h_factor(nngb, k)


def count_within(pos, radius, rows=None):
    """k_i = #{j : |x_i - x_j|^2 <= s_i^2}, self included (LIBSX:906-915)."""
    n = len(pos)
    pos, pp = _d(pos)
    rr, rp, nr = _rows(rows, n)
    rad, radp = _d(np.broadcast_to(radius, (nr,)))
    k = np.zeros(nr, np.int32)
    lib().orc_count_within(n, pp, radp, rp, nr, _i(k)[1])
    return k


def smoothing(pos, h_prev, k_prev=None, divV=None, dt=0.0, p: Params = Params(), rows=None):
    """Smoothing-length iteration (LIBSX3:946-1010, LIBV7:255-330): the
    count-band iteration, or with p.h_mode == 2 (3) the reference's own rule
    of fixed steps of 0.002 (0.02) h_prev (LIBSX3:947-1006, LIBV2:7-64;
    do_smoothingX4), or (4) the LIBV7 driver's steps of 0.01 h_prev halved on
    each jump across the band (LIBV7:255-328), or (5) h_smooth_fast's steps of
    0.003 h on the current h with a strict count (LIBSX:1133-1184).

    Returns (h, k, iterations, status) per row; status 1 = not converged (the
    step rule: the pass limit reached)."""
    n = len(pos)
    pos, pp = _d(pos)
    hp, hpp = _d(h_prev)
    rr, rp, nr = _rows(rows, n)
    if k_prev is not None and divV is not None:
        kp, kpp = _i(k_prev)
        dv, dvp = _d(divV)
    else:
        kp = dv = None
        kpp = C.POINTER(C.c_int32)()
        dvp = C.POINTER(C.c_double)()
    h = np.zeros(nr)
    k = np.zeros(nr, np.int32)
    it = np.zeros(nr, np.int32)
    st = np.zeros(nr, np.int32)
    if p.h_mode in (2, 3, 4, 5):
        lib().orc_smoothing_steps(n, pp, hpp, kpp, dvp, dt, p.nngb, p.ndown, p.nup, p.max_iter,
                                  p.h_mode, p.h_min, rp, nr, _d(h)[1], _i(k)[1], _i(it)[1], _i(st)[1])
    else:
        lib().orc_smoothing(n, pp, hpp, kpp, dvp, dt, p.nngb, p.ndown, p.nup, p.max_iter, rp, nr,
                            _d(h)[1], _i(k)[1], _i(it)[1], _i(st)[1])
    return h, k, it, st


def neighbors(pos, h, rows=None):
    """Neighbour lists under r_ij^2 <= (h_i + h_j)^2 (gate of LIBV2:89); CSR (offsets, idx)."""
    n = len(pos)
    pos, pp = _d(pos)
    hh, hp = _d(h)
    rr, rp, nr = _rows(rows, n)
    off = np.zeros(nr + 1, np.int64)
    lp = C.POINTER(C.c_int64)
    lib().orc_neighbors(n, pp, hp, rp, nr, off.ctypes.data_as(lp), lp())
    idx = np.zeros(max(int(off[-1]), 1), np.int64)
    lib().orc_neighbors(n, pp, hp, rp, nr, off.ctypes.data_as(lp), idx.ctypes.data_as(lp))
    return off, idx[: off[-1]]


def density(pos, m, h, rows=None):
    """getDensity (LIBSX:17-51); returns (rho, neighbour count)."""
    n = len(pos)
    pos, pp = _d(pos)
    m, mp = _d(m)
    h, hp = _d(h)
    rr, rp, nr = _rows(rows, n)
    rho = np.zeros(nr)
    k = np.zeros(nr, np.int32)
    lib().orc_density(n, pp, mp, hp, rp, nr, _d(rho)[1], _i(k)[1])
    return rho, k


def eos(rho, u, gamma):
    """getPressure (LIBSX:100-105) and the adiabatic sound speed (LIBV2:783-789)."""
    n = len(rho)
    rho, rp = _d(rho)
    u, up = _d(u)
    P = np.zeros(n)
    c = np.zeros(n)
    lib().orc_eos(n, rp, up, gamma, _d(P)[1], _d(c)[1])
    return P, c


def divcurl(pos, v, rho, m, h, rows=None):
    """div_curlVel (SHEAR:8-69): signed divergence and |curl|."""
    n = len(pos)
    pos, pp = _d(pos)
    v, vp = _d(v)
    rho, rhop = _d(rho)
    m, mp = _d(m)
    h, hp = _d(h)
    rr, rp, nr = _rows(rows, n)
    dv = np.zeros(nr)
    cv = np.zeros(nr)
    lib().orc_divcurl(n, pp, vp, rhop, mp, hp, rp, nr, _d(dv)[1], _d(cv)[1])
    return dv, cv


def hydro(pos, v, rho, P, c, h, m, divV=None, curlV=None, p: Params = Params(), rows=None):
    """getAcc_sph (LIBSX:110-186) and get_dU (LIBSX:363-430) in one pair loop."""
    n = len(pos)
    pos, pp = _d(pos)
    v, vp = _d(v)
    rho, rhop = _d(rho)
    P, Pp = _d(P)
    c, cp = _d(c)
    h, hp = _d(h)
    m, mp = _d(m)
    dv, dvp = _d(divV if divV is not None else np.zeros(n))
    cv, cvp = _d(curlV if curlV is not None else np.zeros(n))
    rr, rp, nr = _rows(rows, n)
    acc = np.zeros((nr, 3))
    du = np.zeros(nr)
    pc = p.c()
    lib().orc_hydro(n, pp, vp, rhop, Pp, cp, hp, mp, dvp, cvp, C.byref(pc), rp, nr, _d(acc)[1],
                    _d(du)[1])
    return acc, du


def gravity(pos, m, eps, G=1.0, rows=None, potential=False):
    """getAcc_g_smth (LIBSX:633-671); optionally the pair potential of getPE (LIBSX:593-629)."""
    n = len(pos)
    pos, pp = _d(pos)
    m, mp = _d(m)
    e, ep = _d(np.broadcast_to(eps, (n,)))
    rr, rp, nr = _rows(rows, n)
    acc = np.zeros((nr, 3))
    pot = np.zeros(nr) if potential else None
    lib().orc_gravity(n, pp, mp, ep, G, rp, nr, _d(acc)[1],
                      _d(pot)[1] if potential else C.POINTER(C.c_double)())
    return (acc, pot) if potential else acc


def lam(T, cool: CoolParams):
    Tt, Tp = _d(cool.T)
    Lt, Lp = _d(cool.L)
    return lib().orc_lambda(float(T), Tp, Lp, len(Tt))


def cool(rho, u, dt, cp: CoolParams):
    """Implicit per-particle cooling solve (PHOTOPY:489-543 structure) on a tabulated Lambda(T)."""
    n = len(rho)
    rho, rp = _d(rho)
    u, up = _d(u)
    Tt, Tp = _d(cp.T)
    Lt, Lp = _d(cp.L)
    out = np.zeros(n)
    it = np.zeros(n, np.int32)
    c = cp.c()
    lib().orc_cool(n, rp, up, dt, C.byref(c), Tp, Lp, _d(out)[1], _i(it)[1])
    return out, it


def ke(v, m):
    """getKE (LIBSX:579-588)."""
    v, vp = _d(v)
    m, mp = _d(m)
    return lib().orc_ke(len(m), vp, mp)


def pe(pos, m, eps, G=1.0):
    """getPE (LIBSX:593-629)."""
    n = len(pos)
    pos, pp = _d(pos)
    m, mp = _d(m)
    e, ep = _d(np.broadcast_to(eps, (n,)))
    return lib().orc_pe(n, pp, mp, ep, G)


def knn_h(pos, rows=None, kth=64):
    """do_smoothingX((pos, pos[rows])) (LIBSX:762-783): 0.5 x the kth smallest
    distance (0-based, self included) to all particles."""
    n = len(pos)
    pos, pp = _d(pos)
    r, rp, nr = _rows(rows, n)
    out = np.zeros(nr)
    out, op = _d(out)
    lib().orc_knn_h(n, pp, kth, rp, nr, op)
    return out


def cool_grid_apply(rho, u, dt, u_nodes, rho_nodes, du, dt_ref_s, unit_u_cgs, unit_rho_cgs,
                    unit_time_s):
    """applyCooling of the 2-D (u, rho) table (myCppSPHLibs_v3.h:794-859)."""
    n = len(u)
    rho, rp = _d(rho)
    u, up = _d(u)
    un, unp = _d(u_nodes)
    rn, rnp = _d(rho_nodes)
    du, dup = _d(du)
    out = np.zeros(n)
    out, op = _d(out)
    lib().orc_cool_grid_apply(n, rp, up, dt, len(un), unp, len(rn), rnp, dup, dt_ref_s,
                              unit_u_cgs, unit_rho_cgs, unit_time_s, op)
    return out


def cool_grid_build(u_nodes, rho_nodes, dt_ref_s, cp: CoolParams):
    """The table over the grid nodes from the oracle's implicit solve
    (createCoolingGridX2.cu:9-35): du[j, i] = u_j - u_new(u_j, rho_i), cgs."""
    un, unp = _d(u_nodes)
    rn, rnp = _d(rho_nodes)
    du = np.zeros((len(un), len(rn)))
    du, dup = _d(du)
    c = cp.c()
    Tt, Tp = _d(cp.T)
    Lt, Lp = _d(cp.L)
    lib().orc_cool_grid_build(len(un), unp, len(rn), rnp, dt_ref_s, C.byref(c), Tp, Lp, dup)
    return du


class _ChimesParams(C.Structure):
    _fields_ = [("n_kpc", C.c_int32), ("n_nh", C.c_int32), ("n_z", C.c_int32),
                ("n_t", C.c_int32), ("n_m", C.c_int32), ("n_time", C.c_int32),
                ("z_index", C.c_int32), ("pad", C.c_int32), ("xh", C.c_double),
                ("m_h", C.c_double), ("u_max_cgs", C.c_double), ("unit_length_cm", C.c_double),
                ("kpc_cm", C.c_double), ("unit_rho_cgs", C.c_double), ("unit_u_cgs", C.c_double),
                ("unit_time_s", C.c_double)]


@dataclass
class ChimesTable:
    """The 5-D CHIMES table of the v7 GPU drivers (GPU_sph_Type_cooling_II/
    hfvGpuSphTyp_v7.cu:35-79): float32 axes and arrays u_evol, mu of shape
    (kpc, T, nH, Z, time) and metalz of shape (kpc, T, nH, Z, species, time),
    with the lookup constants of hfvCLibs_v7.h:1077-1240 and the driver."""
    kpc: np.ndarray
    nh: np.ndarray
    z: np.ndarray
    temp: np.ndarray
    time_s: np.ndarray
    u_evol: np.ndarray
    mu: np.ndarray
    metalz: np.ndarray
    unit_length_cm: float
    unit_rho_cgs: float
    unit_u_cgs: float
    unit_time_s: float
    z_index: int = 1
    xh: float = 0.70
    m_h: float = 1.6726e-24
    u_max_cgs: float = 2.0287e18
    kpc_cm: float = 3.086e21

    def c(self) -> _ChimesParams:
        nk, nt, nn, nz, nm, ntime = self.metalz.shape
        return _ChimesParams(nk, nn, nz, nt, nm, ntime, self.z_index, 0, self.xh, self.m_h,
                             self.u_max_cgs, self.unit_length_cm, self.kpc_cm, self.unit_rho_cgs,
                             self.unit_u_cgs, self.unit_time_s)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def chimes_apply(pos, rho, u, dt, tab: ChimesTable):
    """hcoolingx (hfvCLibs_v7.h:1077-1240): returns (u_new, ion fractions
    (n, n_species) float32, mu at the chosen node)."""
    n = len(u)
    pos, pp = _d(np.asarray(pos).reshape(n, 3))
    rho, rp = _d(rho)
    u, up = _d(u)
    arrs = [_f(a) for a in (tab.kpc, tab.nh, tab.time_s, tab.u_evol, tab.mu, tab.metalz)]
    nm = tab.metalz.shape[4]
    out, op = _d(np.zeros(n))
    ion, ip_ = _f(np.zeros((n, nm), np.float32))
    mu, mp = _d(np.zeros(n))
    c = tab.c()
    lib().orc_chimes_apply(n, pp, rp, up, dt, C.byref(c), *[a[1] for a in arrs], op, ip_, mp)
    return out, ion, mu


def global_dt(h, c, acc, acc_g, divV, c_cfl=0.25, dt_min=1e-5, dt_max=1e-3):
    """Adaptive global dt of the GPU driver (myCppSPHLibs_v2.h:491-522 dt_array +
    hfv_sph_gpu_v2_evrard.cu:517-561).  Returns (dt, (dt_f, dt_kin, dt_cfl, dt_dens))."""
    n = len(h)
    h, hp = _d(h)
    c, cp = _d(c)
    acc, ap = _d(acc)
    acc_g, gp = _d(acc_g)
    dv, dvp = _d(divV)
    parts = np.zeros(4)
    parts, pp = _d(parts)
    dt = lib().orc_global_dt(n, hp, cp, ap, gp, dvp, c_cfl, dt_min, dt_max, pp)
    return dt, tuple(float(v) for v in parts)


# ---------------------------------------------------------------------------
# Full step, composed from the functions above in the order of the GPU driver
# (hfv_sph_gpu_v2_evrard.cu:407-482), with du/dt evaluated at the same
# velocity as the hydro acceleration (augsphx6.5.py:514-528).
# ---------------------------------------------------------------------------
@dataclass
class State:
    pos: np.ndarray
    vel: np.ndarray
    u: np.ndarray
    m: np.ndarray
    h: np.ndarray
    eps: np.ndarray | None = None
    k: np.ndarray | None = None
    rho: np.ndarray | None = None
    P: np.ndarray | None = None
    c: np.ndarray | None = None
    divV: np.ndarray | None = None
    curlV: np.ndarray | None = None
    acc_g: np.ndarray | None = None
    acc_sph: np.ndarray | None = None
    acc: np.ndarray | None = None
    dudt: np.ndarray | None = None
    ut_prev: np.ndarray | None = None
    h_iters: np.ndarray | None = None
    h_status: np.ndarray | None = None
    ion: np.ndarray | None = None
    t: float = 0.0

    def copy(self):
        return State(**{k: (v if v is None or isinstance(v, float) else np.array(v, copy=True))
                        for k, v in self.__dict__.items()})


def forces(s: State, p: Params, dt: float, use_pred: bool, cool_p: CoolParams | None = None):
    if p.h_mode == 1:
        # kNN h each step, as augsphx6.5.py:510 (no iteration state)
        s.h = knn_h(s.pos, kth=p.knn_kth)
        s.k = np.full(len(s.h), p.knn_kth + 1, np.int32)
        s.h_iters = np.zeros(len(s.h), np.int32)
        s.h_status = np.zeros(len(s.h), np.int32)
    elif use_pred and p.h_predictor and s.k is not None and s.divV is not None:
        s.h, s.k, s.h_iters, s.h_status = smoothing(s.pos, s.h, s.k, s.divV, dt, p)
    else:
        s.h, s.k, s.h_iters, s.h_status = smoothing(s.pos, s.h, p=p)
    s.rho, _ = density(s.pos, s.m, s.h)
    s.P, s.c = eos(s.rho, s.u, p.gamma)
    eps = s.h if p.eps_from_h else s.eps
    s.acc_g = gravity(s.pos, s.m, eps, p.G)
    s.divV, s.curlV = divcurl(s.pos, s.vel, s.rho, s.m, s.h)
    s.acc_sph, s.dudt = hydro(s.pos, s.vel, s.rho, s.P, s.c, s.h, s.m, s.divV, s.curlV, p)
    s.acc = s.acc_g + s.acc_sph


def bootstrap(s: State, p: Params) -> State:
    """Forces at t = 0 (hfv_sph_gpu_v2_evrard.cu:311-392, without its Euler u update)."""
    s = s.copy()
    forces(s, p, 0.0, use_pred=False)
    s.ut_prev = s.dudt.copy()
    return s


def step(s: State, dt: float, p: Params, cool_p: CoolParams | None = None,
         grid: dict | None = None, chimes: "ChimesTable | None" = None,
         network: NetParams | None = None) -> State:
    """One KDK step (hfv_sph_gpu_v2_evrard.cu:410-482).  grid (dict of
    cool_grid_apply's table arguments) cools by the 2-D table after the u
    update, as hfv_sph_gpu_v3.cu:600-609 does; chimes looks the energy up in
    the 5-D table after the u update, as hfvGpuSphTyp_v7.cu:665-673 does
    (s.ion receives the species fractions); network solves the analytic H/He
    network's implicit step (DoCooling_h) after the u update."""
    s = s.copy()
    s.vel = s.vel + s.acc * (0.5 * dt)
    s.pos = s.pos + s.vel * dt
    forces(s, p, dt, use_pred=True)
    s.vel = s.vel + s.acc * (0.5 * dt)
    s.u = s.u + (0.5 * dt) * (s.dudt + s.ut_prev)
    s.ut_prev = s.dudt.copy()
    if cool_p is not None:
        s.u, _ = cool(s.rho, s.u, dt, cool_p)
    elif grid is not None:
        s.u = cool_grid_apply(s.rho, s.u, dt, **grid)
    elif chimes is not None:
        s.u, s.ion, _ = chimes_apply(s.pos, s.rho, s.u, dt, chimes)
    elif network is not None:
        s.u, _, _ = net_cool(s.rho, s.u, dt, network)
    s.t = s.t + dt
    return s


def particle_dt(h, c, acc, acc_g, divV, c_cfl=0.25):
    """Per-particle form of the global criterion (hfv_sph_gpu_v2_evrard.cu:
    555-561 limits, each from the particle's own h, c, |a|, |a_g|, div v):
    0.25 min(sqrt(h/|a_g|), sqrt(h/|a|), c_cfl h/c, c_cfl h/|h div v/3|)."""
    g, a = acc_g, acc
    ag = np.sqrt((g[:, 0] * g[:, 0] + g[:, 1] * g[:, 1]) + g[:, 2] * g[:, 2])
    at = np.sqrt((a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]) + a[:, 2] * a[:, 2])
    with np.errstate(divide="ignore", invalid="ignore"):
        dd = h / np.abs((1.0 / 3.0) * h * divV)
        dt_f = np.sqrt(h / ag)
        dt_kin = np.sqrt(h / at)
        dt_cfl = c_cfl * h / c
    dt_dens = c_cfl * dd
    m = dt_f
    m = np.where(dt_kin < m, dt_kin, m)
    m = np.where(dt_cfl < m, dt_cfl, m)
    m = np.where(dt_dens < m, dt_dens, m)
    return 0.25 * m


def block_level(dt_i, dt_max, levels):
    """The smallest b in 0..levels with dt_max / 2^b <= dt_i (the longest step
    not above dt_i); `levels` and a floor flag when there is none."""
    b = np.full(len(dt_i), levels, np.int64)
    floor = np.ones(len(dt_i), bool)
    for k in range(levels, -1, -1):
        ok = (dt_max * 2.0 ** -k) <= dt_i
        b[ok] = k
        floor[ok] = False
    return b, floor


def block_run(s: State, dt_max: float, levels: int, n_sub: int, p: Params, c_cfl=0.25,
              cool_p: CoolParams | None = None):
    """Block time steps (the plan's sphx_plan_set_block_steps): s is a
    bootstrapped state; n_sub substeps of dt_max / 2^levels.  Particle i steps
    with dt_max / 2^b_i; every particle drifts every substep, forces come for
    all (only the active particles' are used), the active particles take the
    closing half kick, the u trapezoid and cooling over their step, a new level
    (shorter at once, longer one level at a time where the substep count is a
    multiple of the longer step), then the opening half kick of the next step.
    Returns (state, levels, floor counts per substep)."""
    s = s.copy()
    d = dt_max * 2.0 ** -levels
    bins, fl = block_level(particle_dt(s.h, s.c, s.acc, s.acc_g, s.divV, c_cfl), dt_max, levels)
    floors = [int(fl.sum())]
    half = 0.5 * (dt_max * 2.0 ** -bins)
    s.vel = s.vel + s.acc * half[:, None]
    tick = 0
    for _ in range(n_sub):
        s.pos = s.pos + s.vel * d
        forces(s, p, d, use_pred=True)
        t1 = tick + 1
        act = (t1 & ((1 << (levels - bins)) - 1)) == 0
        dto = dt_max * 2.0 ** -bins
        half = 0.5 * dto
        v = s.vel + s.acc * half[:, None]
        u = s.u + half * (s.dudt + s.ut_prev)
        if cool_p is not None:
            for b in np.unique(bins[act]):
                sel = act & (bins == b)
                u[sel], _ = cool(s.rho[sel], u[sel], dt_max * 2.0 ** -int(b), cool_p)
        bt, fl = block_level(particle_dt(s.h, s.c, s.acc, s.acc_g, s.divV, c_cfl), dt_max,
                             levels)
        up = (bins > 0) & ((t1 & ((1 << np.maximum(levels - (bins - 1), 0)) - 1)) == 0)
        bn = np.where(bt >= bins, bt, np.where(up, bins - 1, bins))
        h2 = 0.5 * (dt_max * 2.0 ** -bn)
        v2 = v + s.acc * h2[:, None]
        s.vel = np.where(act[:, None], v2, s.vel)
        s.u = np.where(act, u, s.u)
        s.ut_prev = np.where(act, s.dudt, s.ut_prev)
        bins = np.where(act, bn, bins)
        floors.append(int((fl & act).sum()))
        s.t = s.t + d
        tick = t1 & ((1 << levels) - 1)
    return s, bins, floors


def next_dt(s: State, c_cfl=0.25, dt_min=1e-5, dt_max=1e-3) -> float:
    """dt of the next step from this step's forces (hfv_sph_gpu_v2_evrard.cu:517-561)."""
    return global_dt(s.h, s.c, s.acc, s.acc_g, s.divV, c_cfl, dt_min, dt_max)[0]
