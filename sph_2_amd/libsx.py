"""Drop-in replacements for the reference's FP64 SPH library functions.

Same names, argument order and return shapes as hassanfv/SPH_2
(11_Dec_2022/GPU_sph/IC/libsx2_2t.py = LIBSX, GPU_sph/work_on_smoothing/
libsx2_2t.py = LIBSX3, GPU_sph/work_on_smoothing/shear.py = SHEAR,
GPU_sph/Cooling_v1/photolibs3.py = PHOTOPY): numpy float64 arrays in, fresh
numpy arrays out.  Each call uploads to the GPU, runs the gfx950 kernels of
libsphx.so through its C ABI and copies the result back.

Keyword-only extras select the precision (``mode="mixed"`` or ``"fp64"``),
the artificial-viscosity form and the device.  Where the reference only
printed (iteration caps), these functions warn.
"""
from __future__ import annotations

import ctypes as C
import threading
import warnings

import numpy as np

from . import _lib
from ._lib import check, dev, ptr

_MODES = {"mixed": _lib.MIXED, "fp64": _lib.FP64}
_ctx = {}
_ctx_lock = threading.Lock()


def _context(device: int = 0):
    """The per-device context of the stateless calls (created once; calls
    that share it must not overlap, as the library's contexts are not
    re-entrant)."""
    with _ctx_lock:
        c = _ctx.get(device)
        if c is None:
            _lib.require_gpu()
            h = C.c_void_p()
            check(_lib.load().sphx_ctx_create(device, C.byref(h)), "sphx_ctx_create")
            c = _ctx[device] = h
    return c


_gcfg_fixed = {}  # device -> gravity launch shape fixed by set_gravity_config


def set_gravity_config(ipl=None, tj=512, js=1, *, device=0):
    """Fix the launch shape of the stateless gravity calls on a device
    ((ipl, tj, js) as sphx_ctx_set_gravity_config: ipl 1/2/4 the tiled sum,
    8 the symmetric one), or, called with no shape, return to the tuned shape
    for each call's particle count (sph_2_amd/autotune.py), the default."""
    if ipl is None:
        _gcfg_fixed.pop(device, None)
        return
    cfg = (int(ipl), int(tj), int(js))
    check(_lib.load().sphx_ctx_set_gravity_config(_context(device), *cfg), "set_gravity_config")
    _gcfg_fixed[device] = cfg


def _gravity_shape(device: int, n: int):
    """Give the device's context the tuned shape for n particles unless one
    is fixed."""
    if device not in _gcfg_fixed:
        from .autotune import choose_gravity_config

        check(_lib.load().sphx_ctx_set_gravity_config(_context(device),
                                                      *choose_gravity_config(n)),
              "set_gravity_config")


def _mode(mode: str) -> int:
    try:
        return _MODES[mode]
    except KeyError:
        raise ValueError(f"mode must be one of {sorted(_MODES)}") from None


def _f64(a, shape_tail=()):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def _xyz(a, n=None, name="pos"):
    """Contiguous float64 (n, 3) array; n is taken from it when None.  The
    kernels read n rows of every array they are given, so a shorter one
    would be read past its end: refused here instead."""
    a = _f64(a)
    if a.ndim != 2 or a.shape[1] != 3 or (n is not None and a.shape[0] != n):
        want = "(N, 3)" if n is None else f"({n}, 3)"
        raise ValueError(f"{name} must have shape {want}, got {a.shape}")
    return a


def _vec(a, n, name):
    """Contiguous float64 (n,) array (see _xyz)."""
    a = _f64(a)
    if a.shape != (n,):
        raise ValueError(f"{name} must have shape ({n},), got {a.shape}")
    return a


def _rho_u(rho, u):
    """rho and u as device arrays of one length n (1-D float64)."""
    rho = _f64(rho)
    if rho.ndim != 1:
        raise ValueError(f"rho must be 1-D, got shape {rho.shape}")
    n = len(rho)
    u = _vec(u, n, "u")
    return dev(rho), dev(u)


def _rows(n, nbeg, nend):
    if nbeg is None:
        return 0, n
    if not (0 <= nbeg <= nend <= n):
        raise ValueError("row range out of bounds")
    return int(nbeg), int(nend)


def _params(gama=5.0 / 3.0, eta=0.1, alpha=1.0, beta=2.0, av_form=1, balsara=0, **kw):
    return _lib.default_params(gamma=gama, eta=eta, alpha=alpha, beta=beta, av_form=av_form,
                               balsara=balsara, **kw)


def _t():
    return _lib.torch()


# ---------------------------------------------------------------------------
# Density / EOS
# ---------------------------------------------------------------------------
def getDensity(pos, m, h, *, mode="mixed", device=0, return_count=False, nbeg=None, nend=None):
    """rho_i = sum_j m_j W(r_ij, h_ij), self included (LIBSX:17-51)."""
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    m, h = _vec(m, n, "m"), _vec(h, n, "h")
    dp, dm, dh = dev(pos), dev(m), dev(h)
    rho = t.empty(r1 - r0, dtype=t.float64, device="cuda")
    cnt = t.empty(r1 - r0, dtype=t.int32, device="cuda") if return_count else None
    check(_lib.load().sphx_getDensity(_context(device), _mode(mode), n, ptr(dp), ptr(dm), ptr(dh),
                                      r0, r1, ptr(rho), ptr(cnt), _lib.stream_handle()),
          "getDensity")
    out = rho.cpu().numpy()
    return (out, cnt.cpu().numpy()) if return_count else out


def getDensity_mpi(nbeg, nend, pos, m, h, **kw):
    """Rows [nbeg, nend) of getDensity (LIBSX:58-95)."""
    return getDensity(pos, m, h, nbeg=nbeg, nend=nend, **kw)


def getPressure(rho, u, gama):
    """P = (gama - 1) rho u (LIBSX:100-105)."""
    return getPressure_csound(rho, u, gama)[0]


def getPressure_csound(rho, u, gama, *, device=0):
    """P and the adiabatic sound speed c = sqrt(gama (gama - 1) u) (LIBV2:766-789)."""
    t = _t()
    dr, du = _rho_u(rho, u)
    n = len(dr)
    P = t.empty(n, dtype=t.float64, device="cuda")
    c = t.empty(n, dtype=t.float64, device="cuda")
    check(_lib.load().sphx_eos(_context(device), n, ptr(dr), ptr(du), float(gama), ptr(P), ptr(c),
                               _lib.stream_handle()), "eos")
    return P.cpu().numpy(), c.cpu().numpy()


# ---------------------------------------------------------------------------
# Hydro acceleration and du/dt
# ---------------------------------------------------------------------------
def _hydro(pos, v, rho, P, c, h, m, divV, curlV, p, mode, device, nbeg, nend, want_acc, want_du):
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    host = [pos, _xyz(v, n, "v")] + [
        _vec(a, n, name) for a, name in ((rho, "rho"), (P, "P"), (c, "c"), (h, "h"), (m, "m"))]
    divV = _vec(divV, n, "divV") if divV is not None else None
    curlV = _vec(curlV, n, "curlV") if curlV is not None else None
    arrs = [dev(a) for a in host]
    dv = dev(divV) if divV is not None else None
    cv = dev(curlV) if curlV is not None else None
    acc = t.empty((r1 - r0, 3), dtype=t.float64, device="cuda") if want_acc else None
    du = t.empty(r1 - r0, dtype=t.float64, device="cuda") if want_du else None
    check(_lib.load().sphx_getAcc_sph(_context(device), _mode(mode), n, *[ptr(a) for a in arrs],
                                      ptr(dv), ptr(cv), C.byref(p), r0, r1, ptr(acc), ptr(du),
                                      _lib.stream_handle()), "getAcc_sph")
    return (acc.cpu().numpy() if want_acc else None), (du.cpu().numpy() if want_du else None)


def getAcc_sph(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, *, av_form=1, mode="mixed",
               device=0, nbeg=None, nend=None):
    """Hydro acceleration (LIBSX:110-186).  av_form=1 is the Gadget
    signal-velocity viscosity of that function; av_form=0 selects the
    Monaghan-Gingold form of augsphx6.5x.py:103-139."""
    p = _params(gama, eta, alpha, beta, av_form=av_form)
    return _hydro(pos, v, rho, P, c, h, m, None, None, p, mode, device, nbeg, nend, True, False)[0]


def getAcc_sph_mpi(nbeg, nend, pos, v, rho, P, c, h, m, gama, eta, alpha, beta, **kw):
    """Rows [nbeg, nend) of getAcc_sph (LIBSX:276-355)."""
    return getAcc_sph(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, nbeg=nbeg, nend=nend, **kw)


def get_dU(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, *, av_form=1, mode="mixed", device=0,
           nbeg=None, nend=None):
    """du/dt = sum_j m_j (P_i/rho_i^2 + Pi_ij/2) v_ij . grad_i W (LIBSX:363-430)."""
    p = _params(gama, eta, alpha, beta, av_form=av_form)
    return _hydro(pos, v, rho, P, c, h, m, None, None, p, mode, device, nbeg, nend, False, True)[1]


def get_dU_mpi(nbeg, nend, pos, v, rho, P, c, h, m, gama, eta, alpha, beta, **kw):
    return get_dU(pos, v, rho, P, c, h, m, gama, eta, alpha, beta, nbeg=nbeg, nend=nend, **kw)


def getAcc_sph_dU(pos, v, rho, P, c, h, m, gama=5.0 / 3.0, eta=0.1, alpha=1.0, beta=2.0, *,
                  av_form=0, balsara=0, divV=None, curlV=None, mode="mixed", device=0, nbeg=None,
                  nend=None):
    """Acceleration and du/dt from one fused pair pass (both at the same v)."""
    p = _params(gama, eta, alpha, beta, av_form=av_form, balsara=balsara)
    return _hydro(pos, v, rho, P, c, h, m, divV, curlV, p, mode, device, nbeg, nend, True, True)


def getAcc_sph_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, *, mode="mixed", device=0,
                     nbeg=None, nend=None):
    """Gadget-form viscosity with the Balsara switch (SHEAR:151-235); f_j uses
    divV[j] as in myCppSPHLibs_v2.h:400-401."""
    p = _params(alpha=alpha, av_form=1, balsara=1)
    return _hydro(pos, v, rho, P, c, h, m, divV, curlV, p, mode, device, nbeg, nend, True, False)[0]


def get_dU_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, *, mode="mixed", device=0,
                 nbeg=None, nend=None):
    """du/dt with the Balsara-switched viscosity (SHEAR:334-406)."""
    p = _params(alpha=alpha, av_form=1, balsara=1)
    return _hydro(pos, v, rho, P, c, h, m, divV, curlV, p, mode, device, nbeg, nend, False, True)[1]


def getAcc_sph_shear_mpi(nbeg, nend, pos, v, rho, P, c, h, m, divV, curlV, alpha, **kw):
    """Rows [nbeg, nend) of getAcc_sph_shear (SHEAR:243-330)."""
    return getAcc_sph_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, nbeg=nbeg, nend=nend, **kw)


def get_dU_shear_mpi(nbeg, nend, pos, v, rho, P, c, h, m, divV, curlV, alpha, **kw):
    """Rows [nbeg, nend) of get_dU_shear (SHEAR:413-485)."""
    return get_dU_shear(pos, v, rho, P, c, h, m, divV, curlV, alpha, nbeg=nbeg, nend=nend, **kw)


# ---------------------------------------------------------------------------
# Velocity divergence and curl
# ---------------------------------------------------------------------------
def div_curlVel(pos, v, rho, m, h, *, mode="mixed", device=0, nbeg=None, nend=None):
    """Signed div v and |curl v| (SHEAR:8-69)."""
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    host = [pos, _xyz(v, n, "v")] + [
        _vec(a, n, name) for a, name in ((rho, "rho"), (m, "m"), (h, "h"))]
    arrs = [dev(a) for a in host]
    dv = t.empty(r1 - r0, dtype=t.float64, device="cuda")
    cv = t.empty(r1 - r0, dtype=t.float64, device="cuda")
    check(_lib.load().sphx_div_curlVel(_context(device), _mode(mode), n, *[ptr(a) for a in arrs],
                                       r0, r1, ptr(dv), ptr(cv), _lib.stream_handle()),
          "div_curlVel")
    return dv.cpu().numpy(), cv.cpu().numpy()


def div_curlVel_mpi(nbeg, nend, pos, v, rho, m, h, **kw):
    return div_curlVel(pos, v, rho, m, h, nbeg=nbeg, nend=nend, **kw)


# ---------------------------------------------------------------------------
# Gravity and energies
# ---------------------------------------------------------------------------
def getAcc_g_smth(pos, mass, G, epsilon, *, mode="mixed", device=0, nbeg=None, nend=None,
                  return_potential=False):
    """Spline-softened gravity, eps_ij = (eps_i + eps_j)/2 (LIBSX:633-671)."""
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    eps = np.broadcast_to(np.asarray(epsilon, dtype=np.float64), (n,))
    mass = _vec(mass, n, "mass")
    dp, dm, de = dev(pos), dev(mass), dev(np.ascontiguousarray(eps))
    acc = t.empty((r1 - r0, 3), dtype=t.float64, device="cuda")
    pot = t.empty(r1 - r0, dtype=t.float64, device="cuda") if return_potential else None
    _gravity_shape(device, n)
    check(_lib.load().sphx_getAcc_g_smth(_context(device), _mode(mode), n, ptr(dp), ptr(dm),
                                         ptr(de), float(G), r0, r1, ptr(acc), ptr(pot),
                                         _lib.stream_handle()), "getAcc_g_smth")
    a = acc.cpu().numpy()
    return (a, pot.cpu().numpy()) if return_potential else a


def getAcc_g_smth_mpi(nbeg, nend, pos, mass, G, epsilon, **kw):
    """Rows [nbeg, nend) of the gravity (LIBSX:678-715)."""
    return getAcc_g_smth(pos, mass, G, epsilon, nbeg=nbeg, nend=nend, **kw)


def getAcc_g_smth_mimj_mpi(nbeg, nend, pos, mass, G, epsilon, **kw):
    """(N, 3) partial field of one rank (LIBSX:1377-1414).

    The reference returns the pair terms of rows [nbeg, nend) against j > i,
    applied to both particles, and the driver sums the fields of all ranks
    (sph_mpi4_v4.py:154-160).  Here the rank's field holds the complete
    accelerations of rows [nbeg, nend) and zeros elsewhere, so the sum over a
    partition of the rows is the same total field."""
    n = len(pos)
    out = np.zeros((n, 3))
    out[nbeg:nend] = getAcc_g_smth(pos, mass, G, epsilon, nbeg=nbeg, nend=nend, **kw)
    return out


def getAcc_g_bh(pos, mass, G, epsilon, theta=0.5, *, device=0, nbeg=None, nend=None,
                return_potential=False, return_stats=False):
    """Barnes-Hut spline-softened gravity (Morton-built octree, monopoles,
    opening angle theta); the direct-sum counterpart is getAcc_g_smth."""
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    eps = np.ascontiguousarray(np.broadcast_to(np.asarray(epsilon, dtype=np.float64), (n,)))
    mass = _vec(mass, n, "mass")
    dp, dm, de = dev(pos), dev(mass), dev(eps)
    acc = t.empty((r1 - r0, 3), dtype=t.float64, device="cuda")
    pot = t.empty(r1 - r0, dtype=t.float64, device="cuda") if return_potential else None
    st = (C.c_int64 * 8)()
    check(_lib.load().sphx_bh_gravity(_context(device), n, ptr(dp), ptr(dm), ptr(de), float(G),
                                      float(theta), r0, r1, ptr(acc), ptr(pot), st,
                                      _lib.stream_handle()), "bh_gravity")
    out = [acc.cpu().numpy()]
    if return_potential:
        out.append(pot.cpu().numpy())
    if return_stats:
        out.append(dict(zip(("cells", "depth", "interactions", "overflow", "max_fill",
                             "build_us", "walk_us", "groups"), [int(v) for v in st])))
    return out[0] if len(out) == 1 else tuple(out)


def _diag(pos, v, m, u=None, eps=None, G=1.0, device=0):
    d = _lib.Diag()
    m = _f64(m)
    n = len(m)
    v = _xyz(v, n, "v")
    pos = _xyz(pos, n, "pos") if pos is not None else None
    u = _vec(u, n, "u") if u is not None else None
    dv, dm = dev(v), dev(m)
    dp = dev(pos) if pos is not None else None
    du = dev(u) if u is not None else None
    de = dev(np.ascontiguousarray(np.broadcast_to(np.asarray(eps, np.float64), (len(dm),)))) \
        if eps is not None else None
    check(_lib.load().sphx_diagnostics(_context(device), len(dm), ptr(dp), ptr(dv), ptr(dm),
                                       ptr(du), ptr(de), float(G), C.byref(d),
                                       _lib.stream_handle()), "diagnostics")
    return d.as_dict()


def getKE(v, m, *, device=0):
    """sum 0.5 m v^2 (LIBSX:579-588)."""
    return _diag(None, v, m, device=device)["ke"]


def getPE(pos, m, G, epsilon, *, device=0):
    """sum_{i<j} -G m_i m_j phi(r_ij, eps_ij) (LIBSX:593-629)."""
    return _diag(pos, np.zeros_like(np.asarray(pos, np.float64)), m, eps=epsilon, G=G,
                 device=device)["pe"]


def global_dt(h, csnd, acc_tot, acc_g, divV, C_CFL=0.25, dt_min=1e-5, dt_max=1e-3, *,
              device=0):
    """Adaptive global dt of the GPU driver: dt_array (myCppSPHLibs_v2.h:491-522)
    and the host reduction after it (hfv_sph_gpu_v2_evrard.cu:517-561), on the
    device.  Returns (dt, (dt_f, dt_kin, dt_cfl, dt_dens))."""
    t = _t()
    h = _f64(h)
    n = len(h)
    host = [_vec(csnd, n, "csnd"), _xyz(acc_tot, n, "acc_tot"), _xyz(acc_g, n, "acc_g"),
            _vec(divV, n, "divV")]
    dh = dev(h)
    bufs = [dev(a) for a in host]
    out = t.empty(5, dtype=t.float64, device="cuda")
    check(_lib.load().sphx_global_dt(_context(device), n, ptr(dh), *[ptr(b) for b in bufs],
                                     float(C_CFL), float(dt_min), float(dt_max), ptr(out),
                                     _lib.stream_handle()), "global_dt")
    o = out.cpu().numpy()
    return float(o[0]), tuple(float(v) for v in o[1:])


def pair_issue_rate(iters=2048, *, device=0):
    """Pair evaluations per second that this device sustains on the far-pair
    instruction mix of the symmetric direct sum with registers only (no
    memory traffic): the ceiling the direct sum's pair rate is compared with.
    No counterpart in the reference."""
    import ctypes as C

    out = C.c_double(0.0)
    check(_lib.load().sphx_pair_issue_rate(_context(device), int(iters), C.byref(out),
                                           _lib.stream_handle()), "pair_issue_rate")
    return out.value


def diagnostics(pos, v, m, u=None, epsilon=None, G=1.0, *, device=0):
    """Kinetic, thermal, potential energy, momentum and angular momentum."""
    return _diag(pos, v, m, u=u, eps=epsilon, G=G, device=device)


# ---------------------------------------------------------------------------
# Smoothing length
# ---------------------------------------------------------------------------
def _smoothing(pos, h, k_prev, divV, dt, nbeg, nend, mode, device, **pkw):
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    p = _lib.default_params(**pkw)
    h = _vec(h, n, "h")
    if k_prev is not None:
        k_prev = np.ascontiguousarray(k_prev, dtype=np.int32)
        if k_prev.shape != (n,):
            raise ValueError(f"Nngb_previous must have shape ({n},), got {k_prev.shape}")
    divV = _vec(divV, n, "divV") if divV is not None else None
    dp, dh = dev(pos), dev(h)
    dk = dev(k_prev) if k_prev is not None else None
    ddv = dev(divV) if divV is not None else None
    hout = t.empty(r1 - r0, dtype=t.float64, device="cuda")
    kout = t.empty(r1 - r0, dtype=t.int32, device="cuda")
    it = t.empty(r1 - r0, dtype=t.int32, device="cuda")
    rc = check(_lib.load().sphx_smoothing(_context(device), _mode(mode), n, ptr(dp), ptr(dh),
                                          ptr(dk), ptr(ddv), float(dt), C.byref(p), r0, r1,
                                          ptr(hout), ptr(kout), ptr(it), _lib.stream_handle()),
               "smoothing")
    if rc == _lib.CAPPED:
        warnings.warn("smoothing-length iteration capped for some particles", RuntimeWarning)
    return hout.cpu().numpy(), kout.cpu().numpy(), it.cpu().numpy()


def smoothing_fast_new_mpi3(nbeg, nend, pos, h, Nngb_previous, divV, dt, *, mode="mixed",
                            device=0, **pkw):
    """Smoothing lengths and neighbour counts for rows [nbeg, nend)
    (LIBSX3:947-1010) by the reference's own rule: the Gadget-2 eq. 31 start,
    then steps of 0.002 h until 59..69 neighbours lie within 2h, at most 101
    counts (the largest radius then).  Returns (h, Nngb) like the reference
    (Nngb: the last count).  h_mode=0 gives the count-band iteration instead."""
    pkw.setdefault("h_mode", 2)
    h_out, k, _ = _smoothing(pos, h, np.asarray(Nngb_previous).astype(np.int32), divV, dt, nbeg,
                             nend, mode, device, **pkw)
    return h_out, k.astype(np.float64)


def smoothing_fast_new_mpi(nbeg, nend, pos, h, *, mode="mixed", device=0, **pkw):
    """smoothing_fast_new_mpi (LIBSX:1006-1061): the rule of
    smoothing_fast_new_mpi2 on rows [nbeg, nend)."""
    return smoothing_fast_new_mpi2(nbeg, nend, pos, h, mode=mode, device=device, **pkw)


def smoothing_fast_new_mpi2(nbeg, nend, pos, h, *, mode="mixed", device=0, **pkw):
    """Rows [nbeg, nend) by the reference's own rule (LIBSX:940-990): steps of
    0.002 h from 2h until 59..69 neighbours lie within 2h, at most 101
    counts.  h_mode=0 gives the count-band iteration instead."""
    pkw.setdefault("h_mode", 2)
    return _smoothing(pos, h, None, None, 0.0, nbeg, nend, mode, device, **pkw)[0]


def smoothing_h(pos, h, *, mode="mixed", device=0, nbeg=None, nend=None, **pkw):
    """Rows [nbeg, nend) (default all), iteration started at 2h; returns (h, k,
    iterations)."""
    return _smoothing(pos, h, None, None, 0.0, nbeg, nend, mode, device, **pkw)


def do_smoothingX(poz, *, kth=64, device=0):
    """do_smoothingX((pos, subpos)) (LIBSX:762-783; AUG:305-326): h = 0.5 x
    np.sort(distances to all of pos)[kth] for each row of subpos, exact.
    subpos must be pos itself or a block of its rows (every call in the
    reference passes (pos, pos))."""
    pos_in, sub = poz
    pos = _xyz(pos_in)
    sub = np.asarray(sub)
    n = len(pos)
    if sub is pos_in or (sub.shape == pos.shape and np.array_equal(sub, pos)):
        r0, r1 = 0, n
    else:
        off = _row_offset(np.asarray(pos_in), sub)
        if off is None:
            raise ValueError("do_smoothingX: subpos must be pos or a contiguous block of its rows")
        r0, r1 = off, off + len(sub)
    return _knn_rows(pos, r0, r1, kth, device, "do_smoothingX")


def smoothing_length_mpi(nbeg, nend, pos, *, kth=64, device=0):
    """smoothing_length_mpi(nbeg, nend, pos) (LIBSX:1070-1092): rows [nbeg,
    nend) of do_smoothingX((pos, pos)), 0.5 x np.sort(distances)[64], exact."""
    pos = _xyz(pos)
    r0, r1 = _rows(len(pos), nbeg, nend)
    return _knn_rows(pos, r0, r1, kth, device, "smoothing_length_mpi")


def smoothing_length_mpix(nbeg, nend, pos, h, *, kth=64, device=0):
    """smoothing_length_mpix(nbeg, nend, pos, h) (LIBSX:1096-1126): the same
    value, which the reference takes from the particles within 2.5 h only
    (equal whenever at least 65 lie there; with fewer its loop never ends or
    its index fails).  h is not needed for the exact search."""
    return smoothing_length_mpi(nbeg, nend, pos, kth=kth, device=device)


def _knn_rows(pos, r0, r1, kth, device, what):
    """0.5 x the distance of rank kth (self at rank 0) for rows [r0, r1)."""
    n = len(pos)
    t = _t()
    if r1 == r0:
        return np.zeros(0)
    h0 = _knn_start(pos, kth)
    dp = dev(pos)
    dh = dev(np.full(n, h0))
    out = t.empty(r1 - r0, dtype=t.float64, device="cuda")
    rc = check(_lib.load().sphx_knn_h(_context(device), n, ptr(dp), ptr(dh), int(kth), r0, r1,
                                      ptr(out), _lib.stream_handle()), what)
    if rc == _lib.CAPPED:
        warnings.warn("kNN search did not resolve some rows (NaN)", RuntimeWarning)
    return out.cpu().numpy()


def _knn_start(pos, kth):
    """First search radius / 2 of the kNN walk: the radius that holds kth + 1
    particles at the mean density over the extents that are not degenerate
    (volume for 3-D input, area for planar, length for collinear input), so a
    flat or collinear set does not start at a vanishing radius.  Only a seed:
    the walk brackets and selects the exact value from any positive start."""
    n = len(pos)
    ext = pos.max(0) - pos.min(0)
    big = float(ext.max())
    if not big > 0.0:
        return 1.0                      # every particle on one point: h = 0 is found exactly
    e = ext[ext > 1e-12 * big]
    q = (kth + 1) / n
    if len(e) == 3:
        r = (float(np.prod(e)) * q / 4.18879020478639) ** (1.0 / 3.0)
    elif len(e) == 2:
        r = (float(np.prod(e)) * q / np.pi) ** 0.5
    else:
        r = 0.5 * float(e[0]) * q
    return 0.5 * r


def _row_offset(pos, sub):
    """Row offset of sub inside pos when sub is a contiguous block of pos's rows."""
    if sub.ndim != 2 or sub.shape[1] != 3 or len(sub) > len(pos):
        return None
    if np.shares_memory(pos, sub) and sub.strides == pos.strides:
        d = sub.__array_interface__["data"][0] - pos.__array_interface__["data"][0]
        if d % pos.strides[0] == 0:
            off = d // pos.strides[0]
            if 0 <= off <= len(pos) - len(sub):
                return int(off)
    return None


def do_smoothingX2(pos, h, *, kth=64, device=0):
    """do_smoothingX2(pos, h) (LIBSX:788-820): 0.5 x np.sort(distances within
    3h)[64], the exact kNN h whenever at least 65 particles lie within 3h
    (with fewer the reference's index fails); h is not needed here."""
    return smoothing_length_mpi(None, None, pos, kth=kth, device=device)


def do_smoothingX3(pos, h, *, mode="mixed", device=0, **pkw):
    """do_smoothingX3(pos, h) (LIBSX:825-877): the rule of do_smoothingX4
    (steps of 0.02 h from 2h, at most 101 counts); returns (h, N_iter) like the
    reference, N_iter being the steps taken per row (101 where capped)."""
    pkw.setdefault("h_mode", 3)
    p = _lib.default_params(**pkw)
    h_out, k, it = _smoothing(pos, h, None, None, 0.0, None, None, mode, device, **pkw)
    inband = (k >= p.ndown) & (k <= p.nup)
    return h_out, [int(x) for x in np.where(inband, it - 1, it)]


def do_smoothingX4(pos, h, *, mode="mixed", device=0, **pkw):
    """do_smoothingX4(pos, h) (LIBSX:887-939) by the reference's own rule:
    steps of 0.02 h from 2h until 59..69 neighbours lie within 2h, at most 101
    counts.  h_mode=0 gives the count-band iteration instead."""
    pkw.setdefault("h_mode", 3)
    return _smoothing(pos, h, None, None, 0.0, None, None, mode, device, **pkw)[0]


def h_smooth_fast(pos, h, *, mode="mixed", device=0, nbeg=None, nend=None, **pkw):
    """h_smooth_fast(pos, h) (LIBSX:1133-1184) by the reference's rule: from h,
    h -+= 0.003 h until 59..69 neighbours lie strictly within 2h (nngb=50,
    ndown=40, nup=60 give augsphx6.5x.py's band).  The reference loops without limit;
    a row still outside the band after 10000 counts keeps its last h with a
    RuntimeWarning.  h_mode=0 gives the count-band iteration instead."""
    pkw.setdefault("h_mode", 5)
    return _smoothing(pos, h, None, None, 0.0, nbeg, nend, mode, device, **pkw)[0]


def h_smooth_fast_mpi(nbeg, nend, pos, h, *, mode="mixed", device=0, **pkw):
    """h_smooth_fast_mpi(nbeg, nend, pos, h): rows [nbeg, nend) of h_smooth_fast."""
    return h_smooth_fast(pos, h, mode=mode, device=device, nbeg=nbeg, nend=nend, **pkw)


def h_smooth_fast_h_minimum_set(pos, h, minimum_h, *, mode="mixed", device=0, **pkw):
    """h_smooth_fast_h_minimum_set(pos, h, minimum_h) (LIBSX:1192-1245): the
    rule of h_smooth_fast, except that once eleven steps are taken, a count
    made at h < minimum_h sets h = minimum_h and ends the row."""
    return h_smooth_fast(pos, h, mode=mode, device=device, h_min=float(minimum_h), **pkw)


def h_smooth_fast_mpi_min_h_set(nbeg, nend, pos, h, minimum_h, *, mode="mixed", device=0, **pkw):
    """h_smooth_fast_mpi_min_h_set(nbeg, nend, pos, h, minimum_h)
    (LIBSX:1314-1370): rows [nbeg, nend) of h_smooth_fast_h_minimum_set."""
    return h_smooth_fast(pos, h, mode=mode, device=device, nbeg=nbeg, nend=nend,
                         h_min=float(minimum_h), **pkw)


def getAcc_g_smthx(pos, mass, G, epsilon, **kw):
    """getAcc_g_smthx (LIBSX:720-756): the same softened sum as getAcc_g_smth."""
    return getAcc_g_smth(pos, mass, G, epsilon, **kw)


def neighbors(pos, h, *, device=0, nbeg=None, nend=None):
    """Neighbour lists under r_ij <= h_i + h_j (CSR offsets, sorted indices per row)."""
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    r0, r1 = _rows(n, nbeg, nend)
    h = _vec(h, n, "h")
    dp, dh = dev(pos), dev(h)
    off = t.zeros(r1 - r0 + 1, dtype=t.int64, device="cuda")
    total = C.c_int64()
    L = _lib.load()
    check(L.sphx_neighbors(_context(device), n, ptr(dp), ptr(dh), r0, r1, ptr(off), None, 0,
                           C.byref(total), _lib.stream_handle()), "neighbors")
    idx = t.empty(max(total.value, 1), dtype=t.int32, device="cuda")
    check(L.sphx_neighbors(_context(device), n, ptr(dp), ptr(dh), r0, r1, ptr(off), ptr(idx),
                           total.value, C.byref(total), _lib.stream_handle()), "neighbors")
    o = off.cpu().numpy()
    ix = idx.cpu().numpy()[: total.value].astype(np.int64)
    for r in range(r1 - r0):
        ix[o[r]:o[r + 1]].sort()
    return o, ix


def morton_sort(pos, *, device=0):
    """63-bit Morton keys of pos over its bounding box and the stable sort
    permutation (keys_sorted, perm)."""
    t = _t()
    pos = _xyz(pos)
    n = len(pos)
    dp = dev(pos)
    keys = t.empty(n, dtype=t.int64, device="cuda")
    perm = t.empty(n, dtype=t.int32, device="cuda")
    check(_lib.load().sphx_morton_sort(_context(device), n, ptr(dp), ptr(keys), ptr(perm),
                                       _lib.stream_handle()), "morton_sort")
    return keys.cpu().numpy().view(np.uint64), perm.cpu().numpy()


# ---------------------------------------------------------------------------
# Cooling
# ---------------------------------------------------------------------------
class CoolingTable:
    """Tabulated Lambda(T) plus the unit system of the implicit solve."""

    def __init__(self, T, L, unit_u_cgs=1.0, unit_rho_cgs=1.0, unit_time_s=1.0, XH=0.76, mu=1.22,
                 gamma=5.0 / 3.0, heat=0.0, tol=1e-10, max_iter=100, max_bracket=200):
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        self.L = np.ascontiguousarray(L, dtype=np.float64)
        self.kw = dict(unit_u_cgs=unit_u_cgs, unit_rho_cgs=unit_rho_cgs, unit_time_s=unit_time_s,
                       XH=XH, mu=mu, gamma=gamma, heat=heat, tol=tol)
        self.max_iter = max_iter
        self.max_bracket = max_bracket

    def c(self) -> _lib.Cooling:
        dp = C.POINTER(C.c_double)
        return _lib.Cooling(**self.kw, max_iter=self.max_iter, max_bracket=self.max_bracket,
                            ntab=len(self.T), pad=0, T=self.T.ctypes.data_as(dp),
                            L=self.L.ctypes.data_as(dp))


def DoCooling(rho, u, dt, table: CoolingTable, *, device=0, return_iters=False):
    """Implicit per-particle cooling update in code units (PHOTOPY:489-543 structure)."""
    t = _t()
    dr, du = _rho_u(rho, u)
    n = len(dr)
    out = t.empty(n, dtype=t.float64, device="cuda")
    it = t.empty(n, dtype=t.int32, device="cuda")
    c = table.c()
    rc = check(_lib.load().sphx_cool(_context(device), n, ptr(dr), ptr(du), float(dt), C.byref(c),
                                     ptr(out), ptr(it), _lib.stream_handle()), "cool")
    if rc == _lib.CAPPED:
        warnings.warn("cooling solve hit its iteration cap for some particles", RuntimeWarning)
    o = out.cpu().numpy()
    return (o, it.cpu().numpy()) if return_iters else o


MH_CGS = 1.6726e-24
KB_CGS = 1.380649e-16
YEAR_S = 365.24 * 24.0 * 3600.0


class CoolingGrid:
    """2-D (u, rho) cooling table of the reference's v3 driver
    (GPU_sph/With_cooling/Model_A/hfv_sph_gpu_v3.cu:115-204): strictly
    increasing cgs nodes u_nodes[nu], rho_nodes[nrho] and du[nu, nrho], the cgs
    energy lost over dt_ref_s at each node.  Arrays live on the device."""

    def __init__(self, u_nodes, rho_nodes, du, dt_ref_s, unit_u_cgs, unit_rho_cgs, unit_time_s):
        t = _t()
        self.u_nodes = dev(_f64(u_nodes))
        self.rho_nodes = dev(_f64(rho_nodes))
        du = _f64(du).reshape(len(self.u_nodes), len(self.rho_nodes))
        self.du = dev(du.ravel())
        self.dt_ref_s = float(dt_ref_s)
        self.units = (float(unit_u_cgs), float(unit_rho_cgs), float(unit_time_s))
        self._t = t

    def c(self) -> _lib.CoolGrid:
        return _lib.CoolGrid(len(self.u_nodes), len(self.rho_nodes), ptr(self.u_nodes),
                             ptr(self.rho_nodes), ptr(self.du), self.dt_ref_s, *self.units)

    @staticmethod
    def nodes(xmin, xmax, n):
        """The reference's log-uniform grid: 10^(log10 xmin + i (log10 xmax - log10 xmin)/n),
        i < n (hfv_sph_gpu_v3.cu:158-166, 183-191)."""
        step = (np.log10(xmax) - np.log10(xmin)) / n
        return 10.0 ** (np.log10(xmin) + np.arange(n) * step)

    @classmethod
    def from_solve(cls, table: CoolingTable, T_min=1e4, T_max=1e6, n_u=1000, nH_min=1e-4,
                   nH_max=1e2, n_rho=1000, dt_ref_s=200.0 * YEAR_S, *, device=0):
        """Fill the table from the implicit solve over dt_ref_s (the role of
        GPU_sph/Cooling_v1/createCoolingGridX2.cu); u nodes from T nodes with the
        solve's T(u) = (gamma-1) mu m_H u / k_B, rho nodes = n_H m_H."""
        kw = table.kw
        tfac = (kw["gamma"] - 1.0) * kw["mu"] * MH_CGS / KB_CGS
        un = cls.nodes(T_min, T_max, n_u) / tfac
        rn = cls.nodes(nH_min * MH_CGS, nH_max * MH_CGS, n_rho)
        g = cls(un, rn, np.zeros((n_u, n_rho)), dt_ref_s, kw["unit_u_cgs"], kw["unit_rho_cgs"],
                kw["unit_time_s"])
        c = table.c()
        gc = g.c()
        check(_lib.load().sphx_cool_grid_build(_context(device), C.byref(c), C.byref(gc),
                                               _lib.stream_handle()), "cool_grid_build")
        return g

    def du_host(self):
        return self.du.cpu().numpy().reshape(len(self.u_nodes), len(self.rho_nodes))


def applyCooling(rho, u, dt, grid: CoolingGrid, *, device=0):
    """applyCooling (myCppSPHLibs_v3.h:794-859) in code units: returns the cooled u."""
    t = _t()
    dr, du = _rho_u(rho, u)
    out = t.empty(len(dr), dtype=t.float64, device="cuda")
    g = grid.c()
    check(_lib.load().sphx_cool_grid_apply(_context(device), len(dr), ptr(dr), ptr(du), float(dt),
                                           C.byref(g), ptr(out), _lib.stream_handle()),
          "applyCooling")
    return out.cpu().numpy()


class ChimesTable:
    """5-D CHIMES table of the reference's v7 drivers (GPU_sph_Type_cooling_II/
    hfvGpuSphTyp_v7.cu:35-79, looked up by hcoolingx in hfvCLibs_v7.h:1077-1240):
    float32 axes kpc, nh, z, temp, time_s and arrays u_evol, mu of shape
    (kpc, T, nH, Z, time), metalz of shape (kpc, T, nH, Z, species, time), all
    held on the device; code -> cgs unit factors and the lookup constants
    (hydrogen fraction xh, hydrogen mass m_h, energy cap u_max_cgs, kpc in cm,
    metallicity node z_index) with the reference's values as defaults."""

    def __init__(self, kpc, nh, z, temp, time_s, u_evol, mu, metalz, *, unit_length_cm,
                 unit_rho_cgs, unit_u_cgs, unit_time_s, z_index=1, xh=0.70, m_h=1.6726e-24,
                 u_max_cgs=2.0287e18, kpc_cm=3.086e21):
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
        self.host = {"kpc": f32(kpc), "nh": f32(nh), "z": f32(z), "temp": f32(temp),
                     "time_s": f32(time_s), "u_evol": f32(u_evol), "mu": f32(mu),
                     "metalz": f32(metalz)}
        h = self.host
        shape = (len(h["kpc"]), len(h["temp"]), len(h["nh"]), len(h["z"]), len(h["time_s"]))
        if h["u_evol"].shape != shape or h["mu"].shape != shape:
            raise ValueError(f"u_evol and mu must have shape (kpc, T, nH, Z, time) = {shape}")
        if h["metalz"].ndim != 6 or h["metalz"].shape[:4] != shape[:4] or \
                h["metalz"].shape[5] != shape[4]:
            raise ValueError("metalz must have shape (kpc, T, nH, Z, species, time)")
        self.n_species = int(h["metalz"].shape[4])
        self.shape = shape
        self.d = {k: dev(v.ravel()) for k, v in h.items() if k not in ("z", "temp")}
        self.kw = dict(z_index=int(z_index), xh=float(xh), m_h=float(m_h),
                       u_max_cgs=float(u_max_cgs), kpc_cm=float(kpc_cm),
                       unit_length_cm=float(unit_length_cm), unit_rho_cgs=float(unit_rho_cgs),
                       unit_u_cgs=float(unit_u_cgs), unit_time_s=float(unit_time_s))

    def c(self) -> _lib.ChimesTableC:
        nk, nt, nn, nz, ntime = self.shape
        d = self.d
        return _lib.ChimesTableC(nk, nn, nz, nt, self.n_species, ntime, self.kw["z_index"], 0,
                                 ptr(d["kpc"]), ptr(d["nh"]), ptr(d["time_s"]), ptr(d["u_evol"]),
                                 ptr(d["mu"]), ptr(d["metalz"]) if self.n_species else None,
                                 self.kw["xh"], self.kw["m_h"], self.kw["u_max_cgs"],
                                 self.kw["kpc_cm"], self.kw["unit_length_cm"],
                                 self.kw["unit_rho_cgs"], self.kw["unit_u_cgs"],
                                 self.kw["unit_time_s"])

    @classmethod
    def read(cls, path, **kw):
        """Load coolHeatGridZ.bin; kw: the unit factors and lookup constants."""
        from .formats import read_chimes_table
        return cls(**read_chimes_table(path), **kw)

    def write(self, path):
        from .formats import write_chimes_table
        write_chimes_table(path, **self.host)

    @classmethod
    def from_solve(cls, table: CoolingTable, temp, nh, time_s, kpc=(0.0,), z=(-2.0, -1.0, 0.0),
                   z_index=1, *, unit_length_cm, device=0):
        """A table filled from the library's own implicit solve over Lambda(T)
        (the role the CHIMES runs of cooling_grids_use_EC2/step_1x.py play for
        the reference): node energies u = k_B T / ((gamma-1) mu m_H) at
        time_s[0] = 0, the solve's energy after each later time, repeated over
        the kpc and Z axes (the solve depends on neither); mu constant, no
        species.  xh and m_h follow the solve, so that n_H agrees."""
        kw = table.kw
        temp = np.asarray(temp, np.float64)
        nh = np.asarray(nh, np.float64)
        ts = np.asarray(time_s, np.float64)
        nk, nz = len(kpc), len(z)
        tfac = (kw["gamma"] - 1.0) * kw["mu"] * MH_CGS / KB_CGS
        u0 = np.float32(temp / tfac).astype(np.float64)      # the stored node energies
        nh32 = np.float32(nh).astype(np.float64)
        U, NH = np.meshgrid(u0, nh32, indexing="ij")          # (T, nH)
        rho_code = NH * MH_CGS / kw["XH"] / kw["unit_rho_cgs"]
        evol = np.empty(U.shape + (len(ts),))
        for j, tj in enumerate(ts):
            if tj == 0.0:
                evol[..., j] = U
            else:
                evol[..., j] = DoCooling(rho_code.ravel(), (U / kw["unit_u_cgs"]).ravel(),
                                         tj / kw["unit_time_s"], table,
                                         device=device).reshape(U.shape) * kw["unit_u_cgs"]
        u_evol = np.broadcast_to(evol[None, :, :, None, :], (nk, len(temp), len(nh), nz, len(ts)))
        mu = np.full(u_evol.shape, kw["mu"], np.float32)
        metalz = np.zeros(u_evol.shape[:4] + (0, len(ts)), np.float32)
        return cls(kpc, nh, z, temp, ts, u_evol, mu, metalz, unit_length_cm=unit_length_cm,
                   unit_rho_cgs=kw["unit_rho_cgs"], unit_u_cgs=kw["unit_u_cgs"],
                   unit_time_s=kw["unit_time_s"], z_index=z_index, xh=kw["XH"], m_h=MH_CGS)


def hcoolingx(pos, rho, u, dt, table: ChimesTable, *, device=0):
    """hcoolingx (hfvCLibs_v7.h:1077-1240) on numpy arrays in code units:
    returns (u after the lookup, species fractions (n, n_species) float32, mu at
    the chosen node)."""
    t = _t()
    pos = _xyz(pos, len(np.asarray(rho)), "pos")
    dr, du = _rho_u(rho, u)
    n = len(du)
    dp = dev(pos)
    out = t.empty(n, dtype=t.float64, device="cuda")
    mu = t.empty(n, dtype=t.float64, device="cuda")
    ion = t.empty((n, max(table.n_species, 1)), dtype=t.float32, device="cuda")
    c = table.c()
    check(_lib.load().sphx_chimes_apply(_context(device), n, ptr(dp), ptr(dr), ptr(du), float(dt),
                                        C.byref(c), ptr(out), ptr(ion) if table.n_species else None,
                                        ptr(mu), _lib.stream_handle()), "hcoolingx")
    return (out.cpu().numpy(), ion[:, :table.n_species].cpu().numpy(), mu.cpu().numpy())


def hcooler(r_p, u_p, nH_p, dt_sec, table: ChimesTable, *, device=0):
    """Scalar cgs form of the reference's CPU twin hcooler (GPU_sph_Type_cooling_II/
    cooling_grids_use_EC2/checker1.py:53-146): radius in kpc, u in erg/g, n_H in
    cm^-3, dt in s; returns (u after the lookup in erg/g, mu)."""
    kw = table.kw
    pos = np.array([[r_p * kw["kpc_cm"] / kw["unit_length_cm"], 0.0, 0.0]])
    rho = np.array([nH_p * kw["m_h"] / kw["xh"] / kw["unit_rho_cgs"]])
    u = np.array([u_p / kw["unit_u_cgs"]])
    un, _, mu = hcoolingx(pos, rho, u, dt_sec / kw["unit_time_s"], table, device=device)
    return float(un[0] * kw["unit_u_cgs"]), float(mu[0])


class CoolingNetwork:
    """The analytic H/He cooling network of photolibs3.py:140-543 (sphx_network).

    J21 is the reference's Jcoeff (the UV background at the H0 edge in units of
    1e-21 erg/s/cm^2/Hz/sr): 1e-5 in photolibs3.py; `gpu_drivers()` gives the
    background of the GPU drivers' myPhotoLibsGPU.h (J21 = 50 on 200 frequency
    points per edge).  Unit factors convert code units to cgs (1 for cgs)."""

    def __init__(self, XH=0.76, J21=1e-5, unit_u_cgs=1.0, unit_rho_cgs=1.0, unit_time_s=1.0,
                 tol=1e-4, n_grid=0, max_iter=100, max_bracket=400):
        self.kw = dict(XH=XH, J21=J21, unit_u_cgs=unit_u_cgs, unit_rho_cgs=unit_rho_cgs,
                       unit_time_s=unit_time_s, tol=tol, n_grid=n_grid, max_iter=max_iter,
                       max_bracket=max_bracket)

    @classmethod
    def gpu_drivers(cls, **kw):
        return cls(**{"J21": 50.0, "n_grid": 200, **kw})

    def c(self) -> _lib.Network:
        return _lib.Network(**self.kw, pad=0)

    def rates(self):
        """RadiationField (photolibs3.py:230-329): gJH0, gJHe0, gJHep (s^-1) and the
        heating rates HRate_H0, HRate_He0, HRate_Hep (erg/s)."""
        out = (C.c_double * 6)()
        c = self.c()
        check(_lib.load().sphx_network_rates(C.byref(c), out), "network rates")
        return np.array(out[:])


def cool_network(rho, u, dt, network: CoolingNetwork, *, device=0, return_iters=False):
    """DoCooling_h of the network on arrays in the network's code units."""
    t = _t()
    dr, du = _rho_u(rho, u)
    n = len(dr)
    out = t.empty(n, dtype=t.float64, device="cuda")
    it = t.empty(n, dtype=t.int32, device="cuda")
    c = network.c()
    rc = check(_lib.load().sphx_cool_network(_context(device), n, ptr(dr), ptr(du), float(dt),
                                            C.byref(c), ptr(out), ptr(it), _lib.stream_handle()),
               "cool_network")
    if rc == _lib.CAPPED:
        warnings.warn("network cooling solve capped for some particles", RuntimeWarning)
    res = out.cpu().numpy()
    return (res, it.cpu().numpy()) if return_iters else res


def DoCooling_h(rho, u_old, dt, XH, network: CoolingNetwork | None = None):
    """DoCooling_h(rho, u_old, dt, XH) (photolibs3.py:489-543) in cgs: the u after an
    implicit step of dt seconds of the analytic H/He network (photolibs3.py's
    background unless `network` is given).  Scalars in, a float out, as the
    reference; arrays give arrays."""
    net = CoolingNetwork(**{**(network.kw if network is not None else {}), "XH": XH,
                            "unit_u_cgs": 1.0, "unit_rho_cgs": 1.0, "unit_time_s": 1.0})
    scalar = np.ndim(rho) == 0 and np.ndim(u_old) == 0
    out = cool_network(np.atleast_1d(rho), np.atleast_1d(u_old), dt, net)
    return float(out[0]) if scalar else out
