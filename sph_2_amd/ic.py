"""Deterministic initial conditions and cooling tables.

Every random number comes from a counter-based SplitMix64 stream
(value = mix(seed, stream, counter)), so a given (seed, N) produces the same
particles on any machine and numpy version.

Workloads (11_Dec_2022/ paths of hassanfv/SPH_2 for the reference setups):
  evrard          rho ~ 1/r sphere, R = M = G = 1, u = 0.05, v = 0
                  (Evrard_collapse/augsphx6.5.py:406-457; the reference builds
                  it as a stretched lattice, GPU_sph/Important_Test_Simulations/
                  Evrard_collapse/IC/1.py:31-49; here r = R*sqrt(U)).
  plummer         Plummer sphere r = a / sqrt(U^(-2/3) - 1), mass fraction <= 0.999.
  turbulent_cloud uniform sphere with a Gaussian random velocity field,
                  P(k) ~ k^-4 on a 64^3 grid, trilinear interpolation, rms
                  Mach 2 at T = 1e4 K, 1e4 Msun in 20 pc (unit template
                  GPU_sph/IC/step_1.py:11-66).
  cooling_table   200 log-spaced T nodes on 10..1e8 K of a seeded piecewise
                  power-law Lambda(T).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# physical constants (cgs), as in the reference drivers
G_CGS = 6.6738e-8
MSUN = 1.98892e33
PC = 3.086e18
MH = 1.6726e-24
KB = 1.3807e-16


def splitmix64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n uint64 values of stream `stream` (counters offset .. offset+n-1)."""
    with np.errstate(over="ignore"):
        ctr = np.arange(offset, offset + n, dtype=np.uint64)
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
        z = base + (ctr + np.uint64(1)) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller from two uniform streams."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1_000_003, n)
    u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def isotropic(seed: int, stream: int, n: int) -> np.ndarray:
    ct = 2.0 * uniform(seed, stream, n) - 1.0
    ph = 2.0 * np.pi * uniform(seed, stream + 1, n)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    return np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)


@dataclass
class IC:
    pos: np.ndarray
    vel: np.ndarray
    u: np.ndarray
    m: np.ndarray
    h: np.ndarray
    eps: np.ndarray | None = None
    units: dict | None = None

    @property
    def n(self) -> int:
        return len(self.m)


def _h_guess(rho, m, nngb=64):
    """Smoothing length that holds nngb neighbours at number density rho/m."""
    nd = rho / m
    return 0.5 * (3.0 * nngb / (4.0 * np.pi * nd)) ** (1.0 / 3.0)


def evrard(n: int, seed: int = 0, u0: float = 0.05, nngb: int = 64) -> IC:
    """Evrard collapse: M = R = 1, rho = M / (2 pi R^2 r), u = 0.05, v = 0."""
    r = np.sqrt(uniform(seed, 0, n))
    pos = r[:, None] * isotropic(seed, 10, n)
    m = np.full(n, 1.0 / n)
    rho = 1.0 / (2.0 * np.pi * np.maximum(r, 1e-3))
    h = _h_guess(rho, m, nngb)
    return IC(pos=pos, vel=np.zeros((n, 3)), u=np.full(n, u0), m=m, h=h)


def plummer(n: int, seed: int = 0, a: float = 1.0, eps: float = 0.01) -> IC:
    """Plummer sphere of unit mass, scale a, positions only (v = 0)."""
    u = 0.999 * uniform(seed, 0, n)
    u = np.maximum(u, 1e-300)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * isotropic(seed, 10, n)
    m = np.full(n, 1.0 / n)
    rho = 3.0 / (4.0 * np.pi * a ** 3) * (1.0 + (r / a) ** 2) ** -2.5
    return IC(pos=pos, vel=np.zeros((n, 3)), u=np.zeros(n), m=m, h=_h_guess(rho, m),
              eps=np.full(n, eps))


def cloud_units(mass_msun: float = 1.0e4, radius_pc: float = 20.0) -> dict:
    """Code units of the cooling cloud: M = 1e4 Msun in R = 20 pc (mean n_H ~ 9
    cm^-3).  At this density the 200-node Lambda(T) table cools 1e4 K gas on a
    time scale of a few to tens of 1e-4 code-time steps (1 code time = 13.4
    Myr) toward the ~100 K balance with the 2e-26 erg/s heating, so every step
    solves the implicit energy equation by bracketing and bisection instead of
    stopping at the table's 10 K floor."""
    ul = radius_pc * PC
    um = mass_msun * MSUN
    ut = np.sqrt(ul ** 3 / (G_CGS * um))
    return {"unit_length_cm": ul, "unit_mass_g": um, "unit_time_s": ut,
            "unit_u_cgs": G_CGS * um / ul, "unit_rho_cgs": um / ul ** 3,
            "unit_velocity_cms": ul / ut}


def _grf_velocity(seed: int, ng: int = 64) -> np.ndarray:
    """Gaussian random vector field on ng^3 with power spectrum P(k) ~ k^-4."""
    k1 = np.fft.fftfreq(ng) * ng
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    amp = np.where(k > 0, k, 1.0) ** -2.0 * (k > 0)  # |v_k| ~ k^-2  <=>  P(k) ~ k^-4
    out = np.empty((3, ng, ng, ng))
    for c in range(3):
        re = normal(seed, 100 + 7 * c, ng ** 3).reshape(ng, ng, ng)
        im = normal(seed, 200 + 7 * c, ng ** 3).reshape(ng, ng, ng)
        out[c] = np.real(np.fft.ifftn((re + 1j * im) * amp))
    return out


def _trilinear(field: np.ndarray, pos: np.ndarray, lo: float = -1.0, hi: float = 1.0):
    ng = field.shape[-1]
    g = (pos - lo) / (hi - lo) * (ng - 1)
    g = np.clip(g, 0.0, ng - 1 - 1e-9)
    i0 = np.floor(g).astype(np.int64)
    f = g - i0
    out = np.zeros((len(pos),) + field.shape[:-3])
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                w = wx * wy * wz
                v = field[..., i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz]
                out += (w * v).T if v.ndim > 1 else w * v
    return out


def turbulent_cloud(n: int, seed: int = 0, mach: float = 2.0, T0: float = 1.0e4,
                    mu: float = 1.22, gamma: float = 5.0 / 3.0, units: dict | None = None,
                    ng: int = 64) -> IC:
    """Uniform sphere R = M = 1 with an rms Mach `mach` turbulent velocity field."""
    units = units or cloud_units()
    r = uniform(seed, 0, n) ** (1.0 / 3.0)
    pos = r[:, None] * isotropic(seed, 10, n)
    vf = _grf_velocity(seed, ng)
    vel = _trilinear(vf, pos)
    vel -= vel.mean(axis=0)
    cs = np.sqrt(KB * T0 / (mu * MH)) / units["unit_velocity_cms"]
    vel *= mach * cs / np.sqrt((vel ** 2).sum(axis=1).mean())
    u_cgs = KB * T0 / ((gamma - 1.0) * mu * MH)
    m = np.full(n, 1.0 / n)
    rho = np.full(n, 3.0 / (4.0 * np.pi))
    return IC(pos=pos, vel=vel, u=np.full(n, u_cgs / units["unit_u_cgs"]), m=m,
              h=_h_guess(rho, m), units=units)


def cooling_table(seed: int = 0, ntab: int = 200, tmin: float = 10.0, tmax: float = 1.0e8):
    """(T, Lambda) nodes of a seeded piecewise power-law cooling function.

    Breakpoints at log T = 1..8 with slopes drawn around a generic shape
    (rising to a peak near 1e5 K, dip, bremsstrahlung tail ~ T^0.5)."""
    logT_b = np.array([1.0, 2.0, 3.0, 4.0, 4.3, 5.0, 5.5, 6.5, 7.3, 8.0])
    base = np.array([2.0, 1.5, 0.8, 6.0, 1.2, -0.2, -0.7, -0.5, 0.5])
    jit = 0.2 * (2.0 * uniform(seed, 900, len(base)) - 1.0)
    slopes = base + jit
    logL = np.empty(len(logT_b))
    logL[0] = -29.0
    for k in range(len(slopes)):
        logL[k + 1] = logL[k] + slopes[k] * (logT_b[k + 1] - logT_b[k])
    logL += -21.7 - logL.max()  # peak value ~ 2e-22 erg cm^3 / s
    T = 10.0 ** np.linspace(np.log10(tmin), np.log10(tmax), ntab)
    L = 10.0 ** np.interp(np.log10(T), logT_b, logL)
    return T, L
