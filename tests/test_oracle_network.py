"""Pins of the oracle's analytic H/He cooling network (photolibs3.py:140-543).

The reference keeps two recorded cooling steps of its network:
  GPU_sph/Cooling_v1/single_check.py:36 (commented tuple, photolibs3 era):
      u = 1.1884e13 erg/g, rho = 8.9824e-23 g/cm^3, dt = 6.3114e9 s (200 yr),
      delta_u = 7.9720e12 erg/g (5 significant digits);
  GPU_sph/Cooling_v1/comparex2.py:31-33 (photolibs2 era):
      u = 2.4789e13, rho = 2.4968e-22, dt = 500 yr, delta_u = 2.2955e13.
The first is reproduced to its printed digits with the GPU drivers' background
(J21 = 50 on 200 points per edge, myPhotoLibsGPU.h:166-296); the second, from an
earlier version of the functions ("DO NOT USE version before v3",
photolibs3.py:1), to 0.3 % with photolibs3.py's background.  The rest is fixed
by the mathematics: an independent numpy integration of the background rates,
textbook threshold cross sections, the abundance sum rules and ionisation
limits, the fully ionised T(u), and the bracket the solve returns.
"""
import numpy as np
import pytest

import oracle

XH = 0.76
YEAR = 3600.0 * 24.0 * 365.24
GPU = oracle.NetParams(J21=50.0, n_grid=200)
PY = oracle.NetParams()


def test_reference_recorded_cooling_steps():
    u, it, st = oracle.net_cool(8.9824e-23, 1.1884e13, 6.3114e9, GPU)
    assert st[0] == 0
    du = 1.1884e13 - u[0]
    assert abs(du / 7.9720e12 - 1.0) < 1e-5, du
    u2, _, st2 = oracle.net_cool(2.4968e-22, 2.4789e13, 500 * YEAR, PY)
    assert st2[0] == 0
    assert abs((2.4789e13 - u2[0]) / 2.2955e13 - 1.0) < 3e-3


def _np_rates(J21, n_grid):
    """The background integrals by vectorised numpy (np.trapezoid) for the check."""
    h_ev, c = 4.135667696e-15, 29979245800.0
    edges = ((13.60, (4.298e-1, 5.475e4, 3.288e1, 2.963, 0.0, 0.0, 0.0)),
             (24.59, (1.361e1, 9.492e2, 1.469e0, 3.188, 2.039, 4.434e-1, 2.136)),
             (54.42, (1.720e0, 1.369e4, 3.288e1, 2.963, 0.0, 0.0, 0.0)))
    g, H = [], []
    for E, (E0, s0, ya, P, yw, y0, y1) in edges:
        nu0 = E / h_ev
        a = nu0 / 1e16
        if n_grid:
            nu = (a + np.arange(n_grid) * ((6.0 - a) / n_grid)) * 1e16
        else:
            nu = np.arange(a, 6.0, 0.02) * 1e16
        x = h_ev * nu / E0 - y0
        y = np.sqrt(x ** 2 + y1 ** 2)
        sig = s0 * (((x - 1) ** 2 + yw ** 2) * y ** (0.5 * P - 5.5)) * (1 + np.sqrt(y / ya)) ** (-P)
        sig *= 1e-18
        nu_edge = c / (c / nu0 * 1e8 * 1e-8)
        J = J21 * 1e-21 / (nu / nu_edge)
        g.append(np.trapezoid(4 * np.pi * J * sig / 6.626196e-27 / nu, nu))
        H.append(np.trapezoid(4 * np.pi * J * sig * (nu - nu0) / nu, nu))
    return np.array(g + H)


@pytest.mark.parametrize("p", [PY, GPU])
def test_background_rates_against_numpy_integration(p):
    ref = _np_rates(p.J21, p.n_grid)
    got = oracle.net_rates(p)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    # linear in the background strength
    q = oracle.NetParams(J21=2 * p.J21, n_grid=p.n_grid)
    assert np.allclose(oracle.net_rates(q), 2 * got, rtol=1e-14)


def test_threshold_cross_sections():
    """Verner et al. (1996) at the edges: H0 6.30e-18 cm^2; He+ is hydrogenic
    (sigma_H / Z^2 = 1.58e-18); He0 7.4e-18."""
    h_ev = 4.135667696e-15
    s = lambda E, c: oracle.lib().orc_net_sigma(E / h_ev, *c)
    assert s(13.60, (4.298e-1, 5.475e4, 3.288e1, 2.963, 0.0, 0.0, 0.0)) == pytest.approx(6.30e-18, rel=0.01)
    assert s(54.42, (1.720e0, 1.369e4, 3.288e1, 2.963, 0.0, 0.0, 0.0)) == pytest.approx(1.58e-18, rel=0.01)
    assert s(24.59, (1.361e1, 9.492e2, 1.469e0, 3.188, 2.039, 4.434e-1, 2.136)) == pytest.approx(7.4e-18, rel=0.03)


def test_abundance_sum_rules_and_limits():
    y = 0.24 / (4.0 - 4.0 * 0.24)
    rf = oracle.net_rates(PY)
    for T in (1e4, 3e4, 1e5, 1e6, 1e7):
        for nH in (1e-4, 1e-1, 1e2):
            ab, it = oracle.net_abundance(T, nH, rf)
            nH0, nHe0, nHp, ne, nHep, nHepp = ab
            assert nH0 + nHp == pytest.approx(1.0, abs=1e-15)
            assert nHe0 + nHep + nHepp == pytest.approx(y, rel=1e-13)
            assert min(ab) >= 0.0 and it < 100
            # the fixed point: ne equals the electrons of the ions to the loop's tolerance
            assert ne == pytest.approx(nHp + nHep + 2 * nHepp, abs=2e-4)
    hot, _ = oracle.net_abundance(1e7, 1.0, np.zeros(6))
    assert hot[2] > 0.999 and hot[5] > 0.999 * y          # collisionally fully ionised
    cold, _ = oracle.net_abundance(5e3, 1.0, np.zeros(6))
    assert cold[0] > 0.999                                  # neutral without a background


def test_temperature_of_u():
    """Fully ionised gas: mu = (1 + 4 y) / (1 + y + 1 + 2 y); neutral gas without a
    background: mu = (1 + 4 y) / (1 + y); the damped fixed point reaches both to its
    1e-4 tolerance; T -> u -> T is the identity to that tolerance."""
    kB, mH = 1.3807e-16, 1.6726e-24
    yHe = (1 - XH) / (4 * XH)
    mu_ion = (1 + 4 * yHe) / (1 + yHe + 1 + 2 * yHe)
    rf = oracle.net_rates(PY)
    u = 1e7 / ((2 / 3) * mH / kB * mu_ion)
    assert oracle.net_u_to_temp(u, 1.0, XH, rf) == pytest.approx(1e7, rel=3e-4)
    mu_neu = (1 + 4 * yHe) / (1 + yHe)
    u = 3e3 / ((2 / 3) * mH / kB * mu_neu)
    assert oracle.net_u_to_temp(u, 1.0, XH, np.zeros(6)) == pytest.approx(3e3, rel=3e-4)
    for T in (2e4, 1e5, 1e6):
        uu = oracle.net_temp_to_u(T, 0.1, XH, rf)
        assert oracle.net_u_to_temp(uu, 0.1, XH, rf) == pytest.approx(T, rel=1e-3)


@pytest.mark.parametrize("p", [PY, GPU])
def test_solve_returns_a_bracket_of_the_implicit_equation(p):
    """At the returned u the backward-Euler residual g(v) = v - u_old - nH^2/rho
    rate(v) dt changes sign within the stopping width; dt = 0 leaves u; more dt
    cools more (net cooling) or heats more (net heating)."""
    rng = np.random.default_rng(3)
    rf = oracle.net_rates(p)
    rho = 10 ** rng.uniform(-26, -21, 200)
    u0 = 10 ** rng.uniform(11.5, 14.5, 200)
    dt = 300 * YEAR
    u, it, st = oracle.net_cool(rho, u0, dt, p, rf)
    assert np.all(st == 0) and np.all(it < 100)
    for r, a, b in zip(rho, u0, u):
        nH = XH * r / 1.6726e-24
        g = lambda v: v - a - nH * nH / r * oracle.net_rate(v, nH, XH, rf) * dt
        w = 1.5e-4 * b
        assert g(b - w) <= 0.0 <= g(b + w)
    u_zero, _, _ = oracle.net_cool(rho, u0, 0.0, p, rf)
    assert np.array_equal(u_zero, u0)
    u_long, _, _ = oracle.net_cool(rho, u0, 3 * dt, p, rf)
    moved, moved_more = u - u0, u_long - u0
    big = np.abs(moved) > 1e-3 * u0
    assert np.all(np.sign(moved_more[big]) == np.sign(moved[big]))
    assert np.all(np.abs(moved_more[big]) >= np.abs(moved[big]) * (1 - 1e-3))


def test_library_background_rates_match_the_oracle():
    """sphx_network_rates (host code of the library, no GPU needed) against the
    oracle's restatement, both backgrounds."""
    from sph_2_amd import libsx

    for p, net in ((PY, libsx.CoolingNetwork()), (GPU, libsx.CoolingNetwork.gpu_drivers())):
        assert np.allclose(net.rates(), oracle.net_rates(p), rtol=1e-14, atol=0)
