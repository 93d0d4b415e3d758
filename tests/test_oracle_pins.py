"""Pins of the FP64 oracle (CPU only).

Each test checks one oracle function against something the oracle's own
formula cannot reproduce by construction: a closed form, a conservation law,
a hand-computed value, brute force, an independent restatement (npref), or a
value the reference itself recorded (tests/golden).
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
from oracle import npref
from sph_2_amd import ic

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def kernel_w(r, h=1.0):
    """W(r, h) through the oracle's density: a massless self plus one unit mass."""
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    rho, _ = oracle.density(pos, np.array([0.0, 1.0]), np.array([h, h]), rows=[0])
    return rho[0]


def lattice(nside, spacing=1.0):
    g = np.arange(nside) * spacing
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


# ---------------------------------------------------------------- kernel
def test_kernel_normalised():
    """4 pi int_0^{2h} W r^2 dr = 1 for the M4 kernel with sigma = 1/pi."""
    for h in (0.5, 1.0, 1.7):
        val, err = integrate.quad(lambda r: 4 * math.pi * r * r * kernel_w(r, h), 0, 2 * h,
                                  points=[h], limit=200)
        assert abs(val - 1.0) < 1e-9


def test_kernel_continuity_and_support():
    h = 1.0
    for q0 in (1.0, 2.0):
        a, b = kernel_w(q0 - 1e-7, h), kernel_w(q0 + 1e-7, h)
        assert abs(a - b) < 1e-6
    assert kernel_w(2.0 + 1e-9, h) == 0.0
    assert abs(kernel_w(0.0, h) - 1.0 / math.pi) < 1e-15


def test_gradient_matches_dW_dr():
    """grad_i W = F (x_i - x_j): F r must equal dW/dr (pure pressure pair)."""
    for r in (0.3, 0.9, 1.2, 1.8):
        pos = np.array([[r, 0.0, 0.0], [0.0, 0.0, 0.0]])
        # a_i = -m_j (P_i/rho_i^2 + P_j/rho_j^2) F r_ij with P/rho^2 = 0.5 each, m_j = 1
        acc, _ = oracle.hydro(pos, np.zeros((2, 3)), np.ones(2), np.array([0.5, 0.5]),
                              np.zeros(2), np.ones(2), np.array([0.0, 1.0]), rows=[0])
        dwdr = (kernel_w(r + 1e-6) - kernel_w(r - 1e-6)) / 2e-6
        assert abs(-acc[0, 0] - dwdr) < 1e-7


# ---------------------------------------------------------------- density
def test_lattice_density():
    """Interior of a unit cubic lattice with h = 1.3: rho ~ m / spacing^3."""
    pos = lattice(11)
    n = len(pos)
    centre = np.argmin(((pos - 5.0) ** 2).sum(1))
    rho, k = oracle.density(pos, np.ones(n), np.full(n, 1.3), rows=[centre])
    assert abs(rho[0] - 1.0) < 0.02
    # neighbours within h_i + h_j = 2.6 spacings on the lattice: brute force count
    d2 = ((pos - pos[centre]) ** 2).sum(1)
    assert k[0] == int((d2 <= 2.6 ** 2).sum())


def test_density_permutation_symmetric():
    rng = np.random.default_rng(7)
    n = 500
    pos, m, h = rng.random((n, 3)), rng.random(n), 0.08 + 0.05 * rng.random(n)
    rho, _ = oracle.density(pos, m, h)
    perm = rng.permutation(n)
    rho2, _ = oracle.density(pos[perm], m[perm], h[perm])
    assert np.allclose(rho2, rho[perm], rtol=1e-14, atol=0)


def test_density_matches_numpy_restatement():
    rng = np.random.default_rng(11)
    n = 700
    pos, m, h = rng.random((n, 3)), rng.random(n), 0.07 + 0.05 * rng.random(n)
    rho, k = oracle.density(pos, m, h)
    rho2, k2 = npref.density(pos, m, h)
    assert np.array_equal(k, k2)
    assert np.max(np.abs(rho / rho2 - 1)) < 1e-13


# ---------------------------------------------------------------- div / curl
def test_divcurl_linear_field():
    """v = A x on a lattice: div ~ tr A, |curl| ~ |axial vector of A - A^T|."""
    pos = lattice(13) * 0.1
    n = len(pos)
    A = np.array([[0.3, 0.2, -0.1], [-0.4, 0.5, 0.25], [0.1, -0.3, -0.2]])
    v = pos @ A.T
    h = np.full(n, 0.13)
    m = np.full(n, 1e-3)
    centre = np.argmin(((pos - 0.6) ** 2).sum(1))
    rho, _ = oracle.density(pos, m, h)
    dv, cv = oracle.divcurl(pos, v, rho, m, h, rows=[centre])
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert abs(dv[0] - np.trace(A)) < 0.03 * abs(np.trace(A)) + 1e-3
    assert abs(cv[0] - np.linalg.norm(curl)) < 0.03 * np.linalg.norm(curl)


def test_divcurl_rigid_rotation():
    pos = lattice(13) * 0.1
    n = len(pos)
    w = np.array([0.0, 0.0, 2.0])
    v = np.cross(w, pos - 0.6)
    h, m = np.full(n, 0.13), np.full(n, 1e-3)
    c = np.argmin(((pos - 0.6) ** 2).sum(1))
    rho, _ = oracle.density(pos, m, h)
    dv, cv = oracle.divcurl(pos, v, rho, m, h, rows=[c])
    assert abs(dv[0]) < 1e-10
    assert abs(cv[0] - 2 * np.linalg.norm(w)) < 0.06 * 2 * np.linalg.norm(w)


# ---------------------------------------------------------------- hydro
HEAD_ON = {0: (-2.344964998207222, 2.344964998207222),   # Monaghan-Gingold 1983
           1: (-1.909825895042153, 1.909825895042153)}   # Gadget signal velocity


@pytest.mark.parametrize("av_form", [0, 1])
def test_head_on_pair(av_form):
    """Two unit-mass particles, h = 1, 1 apart, closing at 2; P = 0, c = rho = 1.
    Expected values worked out by hand from augsphx6.5x.py:103-139 (M83) and
    libsx2_2t.py:158-169 (Gadget form)."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    v = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    acc, du = oracle.hydro(pos, v, np.ones(2), np.zeros(2), np.ones(2), np.ones(2), np.ones(2),
                           p=oracle.Params(av_form=av_form))
    ax, dui = HEAD_ON[av_form]
    assert abs(acc[0, 0] - ax) < 1e-12 and abs(acc[1, 0] + ax) < 1e-12
    assert np.all(acc[:, 1:] == 0)
    assert abs(du[0] - dui) < 1e-12 and abs(du[1] - dui) < 1e-12


def test_receding_pair_has_no_viscosity():
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    v = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    acc, du = oracle.hydro(pos, v, np.ones(2), np.zeros(2), np.ones(2), np.ones(2), np.ones(2))
    assert np.all(acc == 0) and np.all(du == 0)


def test_static_zero_pressure_is_force_free():
    rng = np.random.default_rng(2)
    n = 300
    pos = rng.random((n, 3))
    acc, du = oracle.hydro(pos, np.zeros((n, 3)), np.ones(n), np.zeros(n), np.ones(n),
                           np.full(n, 0.1), np.ones(n))
    assert np.all(acc == 0) and np.all(du == 0)


@pytest.mark.parametrize("av_form,balsara", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_hydro_conservation(av_form, balsara):
    """Symmetric SPH: sum m a = 0 and sum m v.a + sum m du/dt = 0 to round-off."""
    rng = np.random.default_rng(3)
    n = 800
    pos = rng.random((n, 3))
    v = rng.standard_normal((n, 3))
    m = 0.5 + rng.random(n)
    h = 0.07 + 0.03 * rng.random(n)
    rho, _ = oracle.density(pos, m, h)
    P, c = oracle.eos(rho, 0.5 + rng.random(n), 5 / 3)
    dv, cv = oracle.divcurl(pos, v, rho, m, h)
    acc, du = oracle.hydro(pos, v, rho, P, c, h, m, dv, cv,
                           oracle.Params(av_form=av_form, balsara=balsara))
    scale = (m[:, None] * np.abs(acc)).sum()
    assert np.abs((m[:, None] * acc).sum(0)).max() < 1e-12 * scale
    e_scale = np.abs(m * du).sum()
    assert abs((m[:, None] * v * acc).sum() + (m * du).sum()) < 1e-12 * e_scale


@pytest.mark.parametrize("av_form,balsara", [(0, 0), (1, 1)])
def test_hydro_matches_numpy_restatement(av_form, balsara):
    rng = np.random.default_rng(4)
    n = 600
    pos, v = rng.random((n, 3)), rng.standard_normal((n, 3))
    m, h = rng.random(n), 0.08 + 0.03 * rng.random(n)
    rho, _ = oracle.density(pos, m, h)
    P, c = oracle.eos(rho, rng.random(n) + 0.1, 5 / 3)
    dv, cv = oracle.divcurl(pos, v, rho, m, h)
    a, du = oracle.hydro(pos, v, rho, P, c, h, m, dv, cv,
                         oracle.Params(av_form=av_form, balsara=balsara))
    a2, du2 = npref.hydro(pos, v, rho, P, c, h, m, dv, cv, av_form=av_form, balsara=balsara)
    assert np.abs(a - a2).max() < 1e-12 * np.abs(a2).max()
    assert np.abs(du - du2).max() < 1e-12 * np.abs(du2).max()


# ---------------------------------------------------------------- gravity
def test_gravity_point_mass_far_field():
    """q > 2: exactly Newtonian, a = G m / r^2 toward the other particle."""
    pos = np.array([[0.0, 0.0, 0.0], [0.7, 0.0, 0.0]])
    acc, pot = oracle.gravity(pos, np.array([1.0, 2.0]), np.full(2, 0.1), 1.5, potential=True)
    assert abs(acc[0, 0] - 1.5 * 2.0 / 0.49) < 1e-13
    assert abs(acc[1, 0] + 1.5 * 1.0 / 0.49) < 1e-13
    assert abs(pot[0] + 1.5 * 2.0 / 0.7) < 1e-13


def test_gravity_force_is_minus_grad_potential():
    """Spline-softened force and potential are consistent across all branches."""
    m = np.array([0.0, 1.0])
    for r in (0.02, 0.07, 0.11, 0.15, 0.19, 0.25):
        def phi(x):
            pos = np.array([[x, 0.0, 0.0], [0.0, 0.0, 0.0]])
            return oracle.gravity(pos, m, np.full(2, 0.1), 1.0, potential=True)[1][0]
        pos = np.array([[r, 0.0, 0.0], [0.0, 0.0, 0.0]])
        a = oracle.gravity(pos, m, np.full(2, 0.1), 1.0)[0, 0]
        dphi = (phi(r + 1e-6) - phi(r - 1e-6)) / 2e-6
        assert abs(a + dphi) < 1e-6 * max(1.0, abs(a))


def test_gravity_newton_third_law():
    rng = np.random.default_rng(5)
    n = 1000
    pos = rng.random((n, 3))
    m = rng.random(n)
    acc = oracle.gravity(pos, m, 0.02 + 0.05 * rng.random(n), 1.0)
    assert np.abs((m[:, None] * acc).sum(0)).max() < 1e-12 * (m[:, None] * np.abs(acc)).sum()


def test_gravity_plummer_mean_field():
    """Plummer sphere: radial acceleration vs -G M r / (r^2 + a^2)^{3/2}."""
    p = ic.plummer(20000, seed=1)
    rows = np.arange(0, 20000, 10)
    acc = oracle.gravity(p.pos, p.m, np.full(len(p.m), 0.01), 1.0, rows=rows)
    r = np.linalg.norm(p.pos[rows], axis=1)
    ar = (acc * p.pos[rows]).sum(1) / r
    sel = (r > 0.3) & (r < 3.0)
    ana = -0.999 * r / (r * r + 1.0) ** 1.5
    ratio = ar[sel] / ana[sel]
    assert abs(np.median(ratio) - 1.0) < 0.02


def test_evrard_potential_energy():
    """rho ~ 1/r sphere of unit mass and radius: W = -(2/3) G M^2 / R."""
    e = ic.evrard(8000, seed=2)
    pe = oracle.pe(e.pos, e.m, np.full(len(e.m), 0.005), 1.0)
    assert abs(pe / (-2.0 / 3.0) - 1.0) < 0.01


def test_gravity_matches_numpy_restatement():
    rng = np.random.default_rng(6)
    n = 700
    pos, m, eps = rng.random((n, 3)), rng.random(n), 0.02 + 0.05 * rng.random(n)
    a, p = oracle.gravity(pos, m, eps, 1.0, potential=True)
    a2, p2 = npref.gravity(pos, m, eps, 1.0, potential=True)
    assert np.abs(a - a2).max() < 1e-12 * np.abs(a2).max()
    assert np.abs(p - p2).max() < 1e-12 * np.abs(p2).max()


def test_reference_recorded_potential_energy():
    """getPE on the reference's own Evrard_1472 IC against the value its driver
    recorded at t = 0 (Evrard_collapse/Energy.pkl).  The run used softening
    0.05 (the script in the repository now reads 0.005): 0.05 reproduces the
    record to float32 input precision, 0.005 misses it by 1.4e-5."""
    icf = np.load(f"{GOLDEN}/evrard_1472_ic.npz")
    en = np.load(f"{GOLDEN}/evrard_1472_energy.npz")
    pos = np.stack([icf["x"], icf["y"], icf["z"]], 1).astype(np.float64)
    n = len(pos)
    pe = oracle.pe(pos, np.full(n, 1.0 / n), np.full(n, 0.05), 1.0)
    assert abs(pe / en["PE"][0] - 1.0) < 1e-9
    pe_wrong = oracle.pe(pos, np.full(n, 1.0 / n), np.full(n, 0.005), 1.0)
    assert abs(pe_wrong / en["PE"][0] - 1.0) > 1e-5
    assert en["KE"][0] == 0.0 and abs(en["U"][0] - 0.05) < 1e-15


# ---------------------------------------------------------------- smoothing
def test_smoothing_brute_force_counts():
    """Converged h: #{j : r_ij <= 2h} (numpy count) is the returned k, in band."""
    e = ic.evrard(3000, seed=3)
    h, k, it, st = oracle.smoothing(e.pos, e.h)
    assert st.sum() == 0
    d2 = ((e.pos[:, None, :] - e.pos[None, :, :]) ** 2).sum(-1)
    kk = (d2 <= (2 * h[:, None]) ** 2).sum(1)
    assert np.array_equal(kk, k)
    assert k.min() >= 59 and k.max() <= 69


def test_smoothing_knn_relation():
    """Converged 2h lies between the 59th and the 70th nearest distances
    (self included): the band condition restated through sorting."""
    rng = np.random.default_rng(8)
    pos = rng.random((1500, 3))
    h, k, _, _ = oracle.smoothing(pos, np.full(1500, 0.05))
    d = np.sort(np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)), axis=1)
    assert np.all(d[:, 58] <= 2 * h * (1 + 1e-15))
    assert np.all(d[:, 69] > 2 * h * (1 - 1e-15))


def test_smoothing_matches_numpy_restatement_with_predictor():
    rng = np.random.default_rng(9)
    pos = rng.random((800, 3))
    h0 = np.full(800, 0.07)
    h, k, _, _ = oracle.smoothing(pos, h0)
    kp = np.clip(k + rng.integers(-20, 20, 800), 1, None)
    dv = rng.standard_normal(800)
    h1, k1, _, _ = oracle.smoothing(pos, h, kp, dv, 0.01)
    h2, k2 = npref.smoothing(pos, h, kp, dv, 0.01)
    assert np.array_equal(h1, h2) and np.array_equal(k1, k2)


def test_smoothing_lattice_degeneracy_caps():
    """On a perfect lattice the count jumps over the band: the iteration caps
    and returns the smallest over-full radius."""
    pos = lattice(9) * 0.1
    n = len(pos)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 0.12))
    c = np.argmin(((pos - 0.4) ** 2).sum(1))
    d2 = ((pos - pos[c]) ** 2).sum(1)
    assert k[c] == int((d2 <= (2 * h[c]) ** 2).sum())
    if st[c]:
        assert k[c] > 69


def _chain(n=201):
    """Unit-spaced particles on a line: an interior particle has exactly
    1 + 2 floor(s) neighbours within s (sqrt of a square of an integer)."""
    pos = np.zeros((n, 3))
    pos[:, 0] = np.arange(float(n))
    return pos


@pytest.mark.parametrize("h_mode,h0,h_want,k_want,passes", [
    (2, 14.0, 14.504, 59, 37),   # up: s = 28 + 36 x 0.028 = 29.008, 57 -> 59 neighbours
    (3, 14.0, 14.56, 59, 5),     # ten times the step: s = 28 + 4 x 0.28 = 29.12
    (2, 18.0, 17.496, 69, 29),   # down: s = 36 - 28 x 0.036 = 34.992, 73 -> 69
    (4, 14.0, 14.56, 59, 9),     # the LIBV7 driver's 0.01 h steps: s = 28 + 8 x 0.14
])
def test_smoothing_step_rule_on_a_chain(h_mode, h0, h_want, k_want, passes):
    """The reference's own rules (LIBSX3:947-1006; do_smoothingX4 with coeff
    0.01; the LIBV7 driver's coeff 0.005) on a chain, where the count at each
    radius is known in closed form: the first radius on the grid
    2 h0 +- m (2 coeff h0) whose count lies in 59..69, reached in m + 1 counts."""
    pos = _chain()
    h, k, it, st = oracle.smoothing(pos, np.full(len(pos), h0), p=oracle.Params(h_mode=h_mode),
                                    rows=[100])
    assert abs(h[0] - h_want) < 1e-9
    assert (k[0], it[0], st[0]) == (k_want, passes, 0)


@pytest.mark.parametrize("h_mode", [2, 4])
def test_smoothing_step_rule_eq31_start(h_mode):
    """The Gadget-2 eq. 31 start of smoothing_fast_new_mpi3 (LIBSX3:961) and
    of the LIBV7 driver (LIBV7:265-266), 2 (0.5 h (1 + (64/k_prev)^(1/3)) +
    h divV dt / 3), in exact binary values: h = 16, k_prev = 64, divV =
    0.1875, dt = 0.5 give s = 33, whose 67 chain neighbours already lie in
    the band: one count, h = 16.5."""
    pos = _chain()
    n = len(pos)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 16.0), np.full(n, 64, np.int32),
                                    np.full(n, 0.1875), 0.5, p=oracle.Params(h_mode=h_mode),
                                    rows=[100])
    assert (h[0], k[0], it[0], st[0]) == (16.5, 67, 1, 0)


def test_smoothing_step_rule_pass_limit():
    """From h0 = 10 the chain needs 450 steps of 0.02 to reach 59 neighbours:
    after 101 counts the rule takes the largest radius tried, 20 + 101 x 0.02,
    and reports the last count (taken at 20 + 100 steps, accumulated by
    repeated addition as the reference does)."""
    pos = _chain()
    h, k, it, st = oracle.smoothing(pos, np.full(len(pos), 10.0), p=oracle.Params(h_mode=2),
                                    rows=[100])
    s = 20.0
    for _ in range(100):
        s = s + 0.002 * 10.0
    assert abs(h[0] - 0.5 * (s + 0.02)) < 1e-12 and abs(h[0] - 11.01) < 1e-9
    assert (k[0], it[0], st[0]) == (1 + 2 * math.floor(s), 101, 1)


def test_smoothing_step_rule_converges_on_the_last_pass():
    """A row whose 101st count is the first in the band keeps that radius, as
    the reference tests the band before its counter: from h0 = 19.43 the chain
    count falls from 71 to 69 exactly at the 100th step of 0.002 h0 (s = 34.974
    < 35 <= 35.0129), so h = 17.487, not the largest radius tried (the start)."""
    pos = _chain()
    h, k, it, st = oracle.smoothing(pos, np.full(len(pos), 19.43), p=oracle.Params(h_mode=2),
                                    rows=[100])
    assert abs(h[0] - 17.487) < 1e-9
    assert (k[0], it[0], st[0]) == (69, 101, 0)


@pytest.mark.parametrize("h0", [1.2, 1.3])
def test_smoothing_halving_rule_on_a_lattice_shell(h0):
    """The LIBV7 driver's rule (LIBV7:255-328) where the count jumps across
    the band: on a cubic lattice a particle has 57 neighbours within
    [sqrt 5, sqrt 6) and 81 at sqrt 6, so the radius crosses sqrt 6 back and
    forth; halving the step on every crossing closes in on sqrt 6, and after
    501 counts h is sqrt(6)/2 to round-off (a fixed step would keep it a step
    away).  The last count is 57, or 70 where the reference stops counting."""
    g = np.arange(11.0)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    c = int(np.argmin(((pos - 5.0) ** 2).sum(1)))
    h, k, it, st = oracle.smoothing(pos, np.full(len(pos), h0), p=oracle.Params(h_mode=4),
                                    rows=[c])
    assert abs(h[0] - math.sqrt(6.0) / 2) < 1e-12
    assert k[0] in (57, 70) and (it[0], st[0]) == (501, 1)


@pytest.mark.parametrize("h0,steps,k_want", [(14.0, 12, 59), (18.0, -10, 69)])
def test_smoothing_relative_rule_on_a_chain(h0, steps, k_want):
    """h_smooth_fast (LIBSX:1133-1184) on a chain, where the strict count is
    #{j: |i - j| < s} = 2 ceil(s) - 1: from h = 14 (s = 28, 55 neighbours) h
    grows by 0.3 % per count until 2h > 29 (59 neighbours), twelve steps since
    ln(14.5 / 14) / ln(1.003) = 11.7; from h = 18 (71 neighbours) it shrinks
    until 2h <= 35 (69 neighbours), ten steps.  h is h0 (1 +- 0.003)^steps to
    round-off."""
    pos = _chain()
    h, k, it, st = oracle.smoothing(pos, np.full(len(pos), h0), p=oracle.Params(h_mode=5),
                                    rows=[100])
    assert abs(h[0] - h0 * (1.0 + math.copysign(0.003, steps)) ** abs(steps)) < 1e-12
    assert (k[0], it[0], st[0]) == (k_want, abs(steps) + 1, 0)


def test_smoothing_relative_rule_counts_strictly():
    """h_smooth_fast counts dist < 2h: at h = 14.5 the particles at distance
    29 are out (57 neighbours), so h takes one step to 14.5435, where the
    inclusive count of the other rules already holds 59 and stops at once.
    With augsphx6.5x.py's band 40..60 (augsphx6.5x.py:426-465), h = 10 counts
    39 and takes one step."""
    pos = _chain()
    n = len(pos)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 14.5), p=oracle.Params(h_mode=5), rows=[100])
    assert abs(h[0] - 14.5435) < 1e-12 and (k[0], it[0]) == (59, 2)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 14.5), p=oracle.Params(h_mode=2), rows=[100])
    assert h[0] == 14.5 and (k[0], it[0]) == (59, 1)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 10.0), p=oracle.Params(h_mode=5, nngb=50, ndown=40, nup=60),
                                    rows=[100])
    assert abs(h[0] - 10.03) < 1e-12 and (k[0], it[0]) == (41, 2)


def test_smoothing_relative_rule_floor():
    """h_smooth_fast_h_minimum_set (LIBSX:1192-1245) on the chain: from h = 19
    (75 neighbours) h shrinks by 0.3 % a step; without a floor it needs 28
    steps to reach 2h <= 35 (69 neighbours); with minimum_h = 18.5 the count
    after the eleventh step is made at 19 x 0.997^11 = 18.38 < 18.5 (73
    neighbours), so h = 18.5 there, after 12 counts."""
    pos = _chain()
    n = len(pos)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 19.0), p=oracle.Params(h_mode=5), rows=[100])
    assert abs(h[0] - 19.0 * 0.997 ** 28) < 1e-12 and (k[0], it[0]) == (69, 29)
    h, k, it, st = oracle.smoothing(pos, np.full(n, 19.0), p=oracle.Params(h_mode=5, h_min=18.5),
                                    rows=[100])
    assert h[0] == 18.5 and (k[0], it[0], st[0]) == (73, 12, 0)
    # a floor the walk never goes under changes nothing
    h, k, it, st = oracle.smoothing(pos, np.full(n, 19.0), p=oracle.Params(h_mode=5, h_min=17.0),
                                    rows=[100])
    assert abs(h[0] - 19.0 * 0.997 ** 28) < 1e-12 and it[0] == 29


def test_smoothing_step_rule_walks_the_grid():
    """On an Evrard sphere from the IC guess: every row that stops inside the
    pass limit ends on the grid 2 h0 + m 0.002 h0 (|m| <= 100) with its count
    (sqrt(r^2) <= 2h, self included) in 59..69, and every grid radius it
    passed on the way has a count on the side it stepped away from; rows at
    the limit end on the largest radius they tried."""
    e = ic.evrard(1500, seed=4)
    rows = np.arange(0, 1500, 7)
    h, k, it, st = oracle.smoothing(e.pos, e.h, p=oracle.Params(h_mode=2), rows=rows)
    for r, i in enumerate(rows):
        d = np.sqrt(((e.pos - e.pos[i]) ** 2).sum(1))
        s0, step = 2 * e.h[i], 0.002 * e.h[i]
        mf = (2 * h[r] - s0) / step
        m = int(round(mf))
        assert abs(mf - m) < 1e-6 and abs(m) <= 101
        if st[r]:
            assert it[r] == 101 and m >= 0
            continue
        assert 59 <= (d <= 2 * h[r]).sum() <= 69 and k[r] == (d <= 2 * h[r]).sum()
        assert it[r] == abs(m) + 1
        for j in range(abs(m)):
            c = (d <= s0 + np.sign(m) * j * step).sum()
            assert (c < 59) if m > 0 else (c > 69)


def test_h_factor_cube_root():
    for k in (9, 27, 50, 64, 100, 300, 511):
        assert abs(oracle.h_factor(64, k) - (64 / k) ** (1 / 3)) < 1e-15
    assert oracle.h_factor(64, 0) == 2.0 and oracle.h_factor(64, 4) == 2.0
    assert oracle.h_factor(64, 10 ** 6) == 0.5


# ---------------------------------------------------------------- cooling
def _cool(L, heat=0.0, tol=1e-12):
    T = 10.0 ** np.linspace(1, 8, 200)
    return oracle.CoolParams(1.0, 1.0, 1.0, heat=heat, tol=tol, T=T, L=np.full(200, L))


def test_cooling_off_leaves_u():
    cp = _cool(0.0)
    u, _ = oracle.cool(np.array([1e-22, 1e-20]), np.array([1e12, 1e13]), 1e10, cp)
    assert np.array_equal(u, [1e12, 1e13])


def test_cooling_constant_lambda_closed_form():
    """Constant Lambda: backward Euler root u = u_old - dt nH^2 Lambda / rho."""
    cp = _cool(1e-23)
    rho = np.array([1e-22, 3e-22])
    u0 = np.array([2e12, 5e12])
    dt = 1e9          # Delta u ~ 10% of u0: the root stays positive
    u, _ = oracle.cool(rho, u0, dt, cp)
    nH = 0.76 * rho / 1.6726e-24
    expect = u0 - dt * nH * nH * 1e-23 / rho
    assert np.all(np.abs(u / expect - 1.0) < 1e-10)


def test_cooling_monotone_in_dt():
    T, L = ic.cooling_table(seed=0)
    cp = oracle.CoolParams(1.0, 1.0, 1.0, T=T, L=L)
    prev = 1e13
    for dt in (1e9, 1e10, 1e11, 1e12):
        u, _ = oracle.cool(np.array([1e-22]), np.array([1e13]), dt, cp)
        assert u[0] <= prev
        prev = u[0]


def test_cooling_table_exact_at_nodes():
    T, L = ic.cooling_table(seed=0)
    cp = oracle.CoolParams(1.0, 1.0, 1.0, T=T, L=L)
    for k in (0, 17, 100, 199):
        assert oracle.lam(T[k], cp) == L[k]
    assert oracle.lam(1.0, cp) == L[0] and oracle.lam(1e9, cp) == L[-1]


# ---------------------------------------------------------------- energies
def test_kinetic_energy_vs_fsum():
    rng = np.random.default_rng(10)
    v, m = rng.standard_normal((5000, 3)), rng.random(5000)
    ke = oracle.ke(v, m)
    ref = math.fsum(0.5 * m * (v * v).sum(1))
    assert abs(ke - ref) < 1e-13 * ref


# ---------------------------------------------------------------- step
def test_step_energy_budget():
    """Evrard N=1500: total energy drift over 5 KDK steps stays small."""
    e = ic.evrard(1500, seed=4)
    p = oracle.Params()
    s = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)

    def energy(s):
        pe = oracle.pe(s.pos, s.m, s.h, 1.0)
        return oracle.ke(s.vel, s.m) + float((s.m * s.u).sum()) + pe

    e0 = energy(s)
    for _ in range(5):
        s = oracle.step(s, 1e-3, p)
    assert abs(energy(s) - e0) < 2e-4 * abs(e0)


def test_cooling_floor_at_table_minimum():
    """Net cooling that would overshoot the lowest table temperature stops at
    u(T_min); inputs below the floor are raised to it."""
    cp = _cool(1e-20)
    tfac = (5.0 / 3.0 - 1.0) * cp.mu * 1.6726e-24 / 1.3807e-16
    u_floor = 10.0 / tfac
    u, _ = oracle.cool(np.array([1e-20]), np.array([1e12]), 1e12, cp)
    assert u[0] == u_floor
    u, _ = oracle.cool(np.array([1e-30]), np.array([-5.0]), 1.0, _cool(0.0))
    assert u[0] == u_floor


# ---------------------------------------------------------------------------
# Global adaptive time step (myCppSPHLibs_v2.h:491-522 + hfv_sph_gpu_v2_evrard.cu:517-561)
# ---------------------------------------------------------------------------
def _dt_case():
    h = np.array([0.1, 0.2, 0.05, 0.3, 0.2])
    c = np.array([1.0, 2.0, 0.5, 0.1, 0.3])
    acc = np.zeros((5, 3))
    acc[2] = [3.0, 4.0, 0.0]           # |a| = 5
    acc_g = np.zeros((5, 3))
    acc_g[1] = [0.0, 0.0, 2.0]         # |a_g| = 2
    divV = np.array([0.0, 1.0, -3.0, 0.5, 0.0])
    return h, c, acc, acc_g, divV


def test_global_dt_hand_values():
    h, c, acc, acc_g, divV = _dt_case()
    dt, parts = oracle.global_dt(h, c, acc, acc_g, divV, 0.25, 1e-9, 1.0)
    # hmin = 0.05; dt_f = sqrt(0.05/2); dt_kin = sqrt(0.05/5) = 0.1;
    # dt_cfl = 0.25 * 0.05 / 2; h/|h divV/3| = 3/|divV| -> min 1 (divV = -3)
    assert parts == (math.sqrt(0.025), 0.1, 0.00625, 0.25)
    assert dt == 0.25 * 0.00625
    # clamps of the driver
    assert oracle.global_dt(h, c, acc, acc_g, divV)[0] == 1e-3
    assert oracle.global_dt(h * 1e-6, c, acc, acc_g, divV)[0] == 1e-5


def test_global_dt_each_limit_can_bind():
    h, c, acc, acc_g, divV = _dt_case()
    big = 1e9
    c_slow = c * 1e-6
    # gravity binds: huge |a_g| (|a_total| kept small)
    g = acc_g.copy(); g[1] = [0, 0, 4e6]
    dt, p = oracle.global_dt(h, c_slow, acc, g, divV * 0, 0.25, 1e-30, big)
    assert dt == 0.25 * math.sqrt(0.05 / 4e6) and p[0] < min(p[1:])
    # total acceleration binds
    a = acc.copy(); a[2] = [0, 0, 9e6]
    dt, p = oracle.global_dt(h, c_slow, a, acc_g, divV * 0, 0.25, 1e-30, big)
    assert dt == 0.25 * math.sqrt(0.05 / 9e6) and p[1] < min(p[0], p[2], p[3])
    # density change binds; divV = 0 everywhere gives an infinite dt_dens
    dt, p = oracle.global_dt(h, c_slow, acc * 0, acc_g * 0, divV * 1e7, 0.25, 1e-30, big)
    assert p[0] == math.inf and p[1] == math.inf
    assert dt == pytest.approx(0.25 * 0.25 * (3.0 / 3e7), rel=1e-15)
    assert oracle.global_dt(h, c, acc, acc_g, divV * 0, 0.25, 1e-30, big)[1][3] == math.inf


def test_global_dt_vsig_is_the_reference_double_loop():
    """dt_array's v_sig[i] = max_j max(c_i, c_j) (myCppSPHLibs_v2.h:506-517),
    reduced with max_finder: the global maximum of c."""
    rng = np.random.default_rng(5)
    c = rng.random(40)
    vsig = max(max(max(ci, cj) for cj in c) for ci in c)
    h = rng.random(40) + 0.5
    acc = np.zeros((40, 3))
    dt, p = oracle.global_dt(h, c, acc, acc, np.zeros(40), 0.25, 1e-30, 1e30)
    assert p[2] == 0.25 * h.min() / vsig


# ---------------------------------------------------------------------------
# 2-D (u, rho) cooling table (myCppSPHLibs_v3.h:794-859)
# ---------------------------------------------------------------------------
def _grid_units():
    return dict(unit_u_cgs=2.0e10, unit_rho_cgs=3.0e-23, unit_time_s=5.0e13)


def _grid_nodes():
    un = 10.0 ** (11 + np.arange(40) * (2.0 / 40))
    rn = 10.0 ** (-28 + np.arange(30) * (6.0 / 30))
    return un, rn


def test_cool_grid_bin_is_the_reference_scan():
    """With du = 0 the result is the node of the chosen bin: the first j with
    g[j] <= x <= g[j+1], restated here with numpy's searchsorted."""
    un, rn = _grid_nodes()
    U = _grid_units()
    du = np.zeros((len(un), len(rn)))
    rng = np.random.default_rng(9)
    xu = np.concatenate([10.0 ** rng.uniform(11, 13, 500), un[1:-1], un[1:-1] * (1 + 1e-14)])
    rho = np.full(len(xu), rn[3] * 1.5)
    got = oracle.cool_grid_apply(rho / U["unit_rho_cgs"], xu / U["unit_u_cgs"], 1.0, un, rn, du,
                                 1.0, **U)
    uc = (xu / U["unit_u_cgs"]) * U["unit_u_cgs"]
    j = np.minimum(np.searchsorted(un[1:], uc, side="left"), len(un) - 2)
    assert np.array_equal(got, un[j] / U["unit_u_cgs"])


def test_cool_grid_clamps_scales_and_floors():
    un, rn = _grid_nodes()
    U = _grid_units()
    rng = np.random.default_rng(2)
    du = rng.random((len(un), len(rn))) * un[:, None] * 0.1
    cu, cr = U["unit_u_cgs"], U["unit_rho_cgs"]
    # below / above the table: bins 0 and n-2
    u = np.array([un[0] * 0.5, un[-1] * 3.0]) / cu
    r = np.array([rn[0] * 0.1, rn[-1] * 10.0]) / cr
    dt_ref = 7.0e10
    out = oracle.cool_grid_apply(r, u, dt_ref / U["unit_time_s"], un, rn, du, dt_ref, **U)
    assert out[0] * cu == pytest.approx(max(un[0] - du[0, 0], un[0]), rel=1e-15)  # floored
    assert out[1] * cu == pytest.approx(un[-2] - du[-2, -2], rel=1e-15)
    # linear in dt / dt_ref
    x = np.array([un[7] * 1.001]) / cu
    y = np.array([rn[4] * 1.001]) / cr
    d1 = un[7] - oracle.cool_grid_apply(y, x, dt_ref / U["unit_time_s"], un, rn, du, dt_ref, **U)[0] * cu
    d2 = un[7] - oracle.cool_grid_apply(y, x, 2 * dt_ref / U["unit_time_s"], un, rn, du, dt_ref, **U)[0] * cu
    assert d1 == pytest.approx(du[7, 4], rel=1e-9) and d2 == pytest.approx(2 * du[7, 4], rel=1e-9)
    # floor at the first u node
    big = np.full_like(du, 1e30)
    assert oracle.cool_grid_apply(y, x, 1.0, un, rn, big, 1.0, **U)[0] * cu == pytest.approx(un[0])


def test_cool_grid_reproduces_the_solve_at_its_nodes():
    """Table built from the implicit solve over dt_ref: at a node and dt = dt_ref
    the lookup equals the solve."""
    units = ic.cloud_units()
    T, L = ic.cooling_table(seed=0)
    cp = oracle.CoolParams(units["unit_u_cgs"], units["unit_rho_cgs"], units["unit_time_s"],
                           T=T, L=L, heat=2e-26)
    tfac = (cp.gamma - 1.0) * cp.mu * 1.6726e-24 / 1.380649e-16
    un = 10.0 ** (4 + np.arange(30) * (2.0 / 30)) / tfac
    rn = 10.0 ** (-4 + np.arange(20) * (6.0 / 20)) * 1.6726e-24
    dt_ref = 200.0 * 365.24 * 24 * 3600
    du = oracle.cool_grid_build(un, rn, dt_ref, cp)
    assert np.all(np.isfinite(du))
    j, i = np.meshgrid(np.arange(1, 29), np.arange(1, 19), indexing="ij")
    j, i = j.ravel(), i.ravel()
    u = un[j] * (1 + 1e-12) / units["unit_u_cgs"]
    rho = rn[i] * (1 + 1e-12) / units["unit_rho_cgs"]
    dt = dt_ref / units["unit_time_s"]
    got = oracle.cool_grid_apply(rho, u, dt, un, rn, du, dt_ref, units["unit_u_cgs"],
                                 units["unit_rho_cgs"], units["unit_time_s"])
    want, _ = oracle.cool(rn[i] / units["unit_rho_cgs"], un[j] / units["unit_u_cgs"], dt, cp)
    want = np.maximum(want, un[0] / units["unit_u_cgs"])
    assert np.max(np.abs(got / want - 1)) < 1e-9


# ---------------------------------------------------------------------------
# kNN smoothing length (do_smoothingX, LIBSX:762-783)
# ---------------------------------------------------------------------------
def test_knn_h_equals_partition_of_distances():
    rng = np.random.default_rng(12)
    pos = rng.random((1500, 3)) ** 2           # uneven density
    h = oracle.knn_h(pos)
    d2 = ((pos[None, :, :] - pos[:, None, :]) ** 2).sum(-1)
    ref = 0.5 * np.sqrt(np.partition(d2, 64, axis=1)[:, 64])
    assert np.array_equal(h, ref)
    rows = np.arange(100, 400)
    assert np.array_equal(oracle.knn_h(pos, rows=rows), ref[rows])
    assert np.array_equal(oracle.knn_h(pos, kth=10), 0.5 * np.sqrt(np.partition(d2, 10, axis=1)[:, 10]))


def test_knn_h_cubic_lattice_shells():
    """Interior lattice point: shells 0, 1, sqrt2, sqrt3, 2, sqrt5 hold 1+6+12+8+6+24 = 57
    points, so rank 64 lies on the sqrt6 shell: h = sqrt(6) a / 2."""
    a = 0.1
    g = np.arange(12) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    centre = np.where(np.all(np.abs(pos - 0.55) < 0.06, axis=1))[0]
    h = oracle.knn_h(pos, rows=centre)
    assert np.allclose(h, 0.5 * np.sqrt(6.0) * a, rtol=1e-14)


def test_reference_recorded_energy_history():
    """The whole composed step against the reference's own execution: its
    Python Evrard driver (Evrard_collapse/augsphx6.5_Monagh_visc.py, the
    Monaghan-Gingold variant of augsphx6.5.py:456-567) on its N = 1472 IC with
    kNN h each step, alpha = 1, beta = 2, eta = 0.1, softening 0.05, dt = 1e-3,
    KDK with the u trapezoid, recorded U, KE (at the half-kicked velocity) and
    PE in Energy.pkl.  The oracle follows that record to ~1e-7 through the
    collapse (the reference keeps positions in float32, hence 1e-7 bars);
    checked here over the first 120 steps."""
    icf = np.load(f"{GOLDEN}/evrard_1472_ic.npz")
    en = np.load(f"{GOLDEN}/evrard_1472_energy.npz")
    pos = np.stack([icf["x"], icf["y"], icf["z"]], 1).astype(np.float64)
    n = len(pos)
    m = np.full(n, 1.0 / n)
    eps = np.full(n, 0.05)
    dt = 1e-3
    p = oracle.Params(h_mode=1, knn_kth=64, av_form=0, alpha=1.0, beta=2.0, eta=0.1, balsara=0,
                      h_predictor=False, eps_from_h=False)
    s = oracle.bootstrap(oracle.State(pos.copy(), np.zeros((n, 3)), np.full(n, 0.05), m,
                                      oracle.knn_h(pos), eps=eps), p)
    for k in range(120):
        s = oracle.step(s, dt, p)
        if (k + 1) % 30 == 0:
            vmid = s.vel - s.acc * (0.5 * dt)
            ke = 0.5 * (m * (vmid ** 2).sum(1)).sum()
            u = (m * s.u).sum()
            pe = oracle.pe(s.pos, m, eps, 1.0)
            assert abs(u / en["U"][k + 1] - 1) < 1e-7, k
            assert abs(ke / en["KE"][k + 1] - 1) < 1e-6, k
            assert abs(pe / en["PE"][k + 1] - 1) < 1e-7, k


# ---------------------------------------------------------------- block time steps
def _boot(n=512, seed=5):
    e = ic.evrard(n, seed=seed)
    p = oracle.Params()
    return oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p), p


def test_block_levels_closed_form():
    """The level is the smallest b with dt_max / 2^b <= dt_i: b = ceil(log2(dt_max /
    dt_i)) clamped to [0, levels], with a floor flag below the shortest step."""
    rng = np.random.default_rng(3)
    dt_i = 10.0 ** rng.uniform(-4, 0.5, 5000)
    b, fl = oracle.block_level(dt_i, 1.0, 6)
    want = np.clip(np.ceil(np.log2(1.0 / dt_i)), 0, 6).astype(int)
    assert np.array_equal(b, want)
    assert np.array_equal(fl, dt_i < 2.0 ** -6)
    b, fl = oracle.block_level(np.array([0.5, 0.25, np.nan, np.inf]), 1.0, 3)
    assert list(b) == [1, 2, 3, 0] and list(fl) == [False, False, True, False]


def test_block_steps_with_one_level_are_the_global_step():
    """levels = 0: every particle steps with dt_max, which is the fixed-dt KDK
    step, bit for bit."""
    s0, p = _boot()
    want = s0
    for _ in range(3):
        want = oracle.step(want, 2e-3, p)
    got, bins, _ = oracle.block_run(s0, 2e-3, 0, 3, p)
    assert np.all(bins == 0)
    for f in ("pos", "u", "h", "rho"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    # between substeps a block-stepped particle carries the opening half kick of
    # its next step (its drift velocity)
    assert np.array_equal(got.vel, want.vel + want.acc * (0.5 * 2e-3))


def test_block_steps_all_on_the_shortest_level_are_the_global_step():
    """With dt_max far above every particle's criterion all particles sit on
    the shortest level: again the fixed-dt step with dt_max / 2^levels."""
    s0, p = _boot()
    dti = oracle.particle_dt(s0.h, s0.c, s0.acc, s0.acc_g, s0.divV)
    big = 8.0 * 1.25 * dti.max()          # even the shortest step, big / 8, is too long
    got, bins, floors = oracle.block_run(s0, big, 3, 4, p)
    assert floors[0] == len(bins) and np.all(bins == 3)
    want = s0
    for _ in range(4):
        want = oracle.step(want, big / 8, p)
    assert np.all(np.isfinite(want.pos))
    for f in ("pos", "u"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    assert np.array_equal(got.vel, want.vel + want.acc * (0.5 * (big / 8)))
