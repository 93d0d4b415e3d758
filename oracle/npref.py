"""Independent pure-numpy restatement of the oracle (small N only, O(N^2) memory).

Written separately from sph_oracle.c, from the same reference functions
(11_Dec_2022/GPU_sph/IC/libsx2_2t.py, work_on_smoothing/shear.py,
Evrard_collapse_new_h_algorithm/augsphx6.5x.py), so a transcription slip in
either one shows up as a disagreement.  Vectorised over the N x N pair
matrix instead of looping.
"""
from __future__ import annotations

import numpy as np

SIGMA = 1.0 / np.pi


def _pairs(pos):
    d = pos[:, None, :] - pos[None, :, :]          # d[i, j] = x_i - x_j
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return d, r2


def _w(q, hij):
    return np.where(q <= 1.0, 1.0 - 1.5 * q ** 2 + 0.75 * q ** 3,
                    0.25 * np.clip(2.0 - q, 0.0, None) ** 3) * SIGMA / hij ** 3


def _f(q, hij):
    qs = np.where(q > 0, q, 1.0)
    return np.where(q <= 1.0, -3.0 + 2.25 * q,
                    -0.75 * np.clip(2.0 - q, 0.0, None) ** 2 / qs) * SIGMA / hij ** 5


def density(pos, m, h):
    d, r2 = _pairs(pos)
    t = h[:, None] + h[None, :]
    nb = r2 <= t * t
    hij = 0.5 * t
    q = np.sqrt(r2) / hij
    return np.where(nb, m[None, :] * _w(q, hij), 0.0).sum(1), nb.sum(1)


def divcurl(pos, v, rho, m, h):
    d, r2 = _pairs(pos)
    t = h[:, None] + h[None, :]
    nb = r2 <= t * t
    hij = 0.5 * t
    F = np.where(nb, _f(np.sqrt(r2) / hij, hij), 0.0)
    vij = v[:, None, :] - v[None, :, :]
    g = F[..., None] * d
    wgt = m[None, :] / rho[:, None]
    div = (wgt * -(vij * g).sum(-1)).sum(1)
    curl = (wgt[..., None] * np.cross(vij, g)).sum(1)
    return div, np.linalg.norm(curl, axis=1)


def hydro(pos, v, rho, P, c, h, m, divV=None, curlV=None, gamma=5 / 3, alpha=1.0, beta=2.0,
          eta=0.1, av_form=0, balsara=0, av_w_eps=1e-5):
    d, r2 = _pairs(pos)
    t = h[:, None] + h[None, :]
    nb = r2 <= t * t
    hij = 0.5 * t
    r = np.sqrt(r2)
    F = np.where(nb, _f(r / hij, hij), 0.0)
    vij = v[:, None, :] - v[None, :, :]
    vr = (vij * d).sum(-1)
    rhob = 0.5 * (rho[:, None] + rho[None, :])
    if av_form == 0:
        mu = hij * vr / (r2 + hij ** 2 * eta ** 2)
        pi = (-alpha * 0.5 * (c[:, None] + c[None, :]) * mu + beta * mu ** 2) / rhob
    else:
        w = vr / (r + av_w_eps)
        pi = -0.5 * alpha * (c[:, None] + c[None, :] - 3.0 * w) * w / rhob
    pi = np.where(vr <= 0, pi, 0.0)
    if balsara:
        f = np.abs(divV) / (np.abs(divV) + curlV + 1e-4 * c / h)
        pi = pi * 0.5 * (f[:, None] + f[None, :])
    Pr = P / rho ** 2
    s = m[None, :] * (Pr[:, None] + Pr[None, :] + pi) * F
    acc = -(s[..., None] * d).sum(1)
    du = (m[None, :] * (Pr[:, None] + 0.5 * pi) * F * vr).sum(1)
    return acc, du


def gravity(pos, m, eps, G=1.0, potential=False):
    n = len(pos)
    d = pos[None, :, :] - pos[:, None, :]           # x_j - x_i
    r = np.sqrt((d ** 2).sum(-1))
    e = 0.5 * (eps[:, None] + eps[None, :])
    q = r / e
    off = ~np.eye(n, dtype=bool)
    rs = np.where(r > 0, r, 1.0)
    f_in = (4.0 / 3.0 - 1.2 * q ** 2 + 0.5 * q ** 3) / e ** 3
    f_mid = (-1 / 15 + 8 / 3 * q ** 3 - 3 * q ** 4 + 1.2 * q ** 5 - q ** 6 / 6) / rs ** 3
    f_out = 1.0 / rs ** 3
    f = np.where(q <= 1, f_in, np.where(q <= 2, f_mid, f_out)) * off
    acc = G * (f[..., None] * d * m[None, :, None]).sum(1)
    if not potential:
        return acc
    p_in = (-2 / e) * (q ** 2 / 3 - 0.15 * q ** 4 + q ** 5 / 20) + 1.4 / e
    p_mid = -1 / (15 * rs) - (1 / e) * (4 / 3 * q ** 2 - q ** 3 + 0.3 * q ** 4 - q ** 5 / 30) + 1.6 / e
    p_out = 1.0 / rs
    ph = np.where(q <= 1, p_in, np.where(q <= 2, p_mid, p_out)) * off
    return acc, -G * (ph * m[None, :]).sum(1)


def h_factor(nngb, k):
    """Same fixed-count Newton cube root as the C oracle (IEEE doubles, no FMA)."""
    if 8 * k <= nngb:
        return 2.0
    if k >= 8 * nngb:
        return 0.5
    x = float(nngb) / float(k)
    y = 1.0
    for _ in range(12):
        y2 = y * y
        y3 = y2 * y
        y = y - (y3 - x) / (3.0 * y2)
    return min(max(y, 0.5), 2.0)


def smoothing(pos, h_prev, k_prev=None, divV=None, dt=0.0, nngb=64, ndown=59, nup=69,
              max_iter=100):
    n = len(pos)
    hout = np.zeros(n)
    kout = np.zeros(n, np.int64)
    for i in range(n):
        d = pos[i] - pos
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        hp = float(h_prev[i])
        if k_prev is not None:
            s = 2.0 * (0.5 * hp * (1.0 + h_factor(nngb, int(k_prev[i]))) + ((hp * float(divV[i])) * dt) / 3.0)
            if not (0.0 < s < np.inf):
                s = 2.0 * hp
        else:
            s = 2.0 * hp
        lo = hi = None
        klo = khi = 0
        done = False
        it = 0
        k = 0
        while it < max_iter:
            k = int(np.count_nonzero(r2 <= s * s))
            it += 1
            if ndown <= k <= nup:
                done = True
                break
            if k < ndown:
                lo, klo = s, k
                if k >= n:
                    break
            else:
                hi, khi = s, k
            sn = s * h_factor(nngb, k)
            if lo is not None and hi is not None:
                if not (lo < sn < hi):
                    sn = 0.5 * (lo + hi)
                if not (lo < sn < hi):
                    break
            s = sn
        if not done:
            s, k = (hi, khi) if hi is not None else (lo, klo)
        hout[i] = 0.5 * s
        kout[i] = k
    return hout, kout
