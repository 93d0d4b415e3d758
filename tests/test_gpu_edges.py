"""Edge inputs of the drop-in functions on the GPU: no particles, one
particle, and row ranges at the ends of the array (the reference's numpy
functions accept all of these and return arrays of the matching shape)."""
import numpy as np
import pytest

from sph_2_amd import ic

pytestmark = pytest.mark.gpu

G53 = 5.0 / 3.0


def test_no_particles(gpu):
    from sph_2_amd import libsx

    z3, z = np.zeros((0, 3)), np.zeros(0)
    assert libsx.getDensity(z3, z, z).shape == (0,)
    assert libsx.getAcc_sph(z3, z3, z, z, z, z, z, G53, 0.1, 1.0, 2.0).shape == (0, 3)
    assert libsx.get_dU(z3, z3, z, z, z, z, z, G53, 0.1, 1.0, 2.0).shape == (0,)
    dv, cv = libsx.div_curlVel(z3, z3, z, z, z)
    assert dv.shape == (0,) and cv.shape == (0,)
    assert libsx.getAcc_g_smth(z3, z, 1.0, z).shape == (0, 3)
    assert libsx.getAcc_g_smth(z3, z, 1.0, 0.01, mode="fp64").shape == (0, 3)
    assert libsx.getAcc_g_bh(z3, z, 1.0, z).shape == (0, 3)
    h, k, it = libsx.smoothing_h(z3, z)
    assert h.shape == (0,) and k.shape == (0,)
    assert libsx.do_smoothingX((z3, z3)).shape == (0,)
    off, idx = libsx.neighbors(z3, z)
    assert list(off) == [0] and idx.shape == (0,)
    keys, perm = libsx.morton_sort(z3)
    assert keys.shape == (0,) and perm.shape == (0,)
    P, c = libsx.getPressure_csound(z, z, G53)
    assert P.shape == (0,) and c.shape == (0,)
    assert libsx.getKE(z3, z) == 0.0
    assert libsx.getPE(z3, z, 1.0, 0.01) == 0.0


def test_one_particle(gpu):
    """A lone particle: its density is its own kernel term m W(0, h) =
    m / (pi h^3); it feels no gravity and no pressure force."""
    from sph_2_amd import libsx

    pos = np.array([[0.3, -0.2, 0.1]])
    m, h = np.array([2.0]), np.array([0.1])
    for mode in ("mixed", "fp64"):
        rho = libsx.getDensity(pos, m, h, mode=mode)
        assert abs(rho[0] / (2.0 / (np.pi * 0.1 ** 3)) - 1) < 1e-6, (mode, rho)
        assert np.array_equal(libsx.getAcc_g_smth(pos, m, 1.0, h, mode=mode), np.zeros((1, 3)))
        P = np.array([1.0])
        acc = libsx.getAcc_sph(pos, np.zeros((1, 3)), rho, P, P, h, m, G53, 0.1, 1.0, 2.0,
                               mode=mode)
        assert np.array_equal(acc, np.zeros((1, 3))), (mode, acc)


def test_row_ranges_at_the_ends(gpu):
    """Rows [n-1, n), [0, 1) and the empty ranges [0, 0) and [n, n) of the
    _mpi variants are the matching slices of the full result."""
    from sph_2_amd import libsx

    e = ic.evrard(3000, seed=12)
    n = e.n
    full = libsx.getDensity(e.pos, e.m, e.h)
    g = libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.h)
    for a, b in ((n - 1, n), (0, 1), (0, 0), (n, n), (1234, 2345)):
        assert np.array_equal(libsx.getDensity_mpi(a, b, e.pos, e.m, e.h), full[a:b]), (a, b)
        assert np.array_equal(libsx.getAcc_g_smth_mpi(a, b, e.pos, e.m, 1.0, e.h), g[a:b]), (a, b)
    with pytest.raises(ValueError):
        libsx.getDensity_mpi(5, 4, e.pos, e.m, e.h)


def test_plan_refuses_a_short_array(gpu):
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(2048, seed=2)
    pl = SphPlan(2048)
    with pytest.raises(ValueError, match="h must have shape"):
        pl.load(e.pos, e.vel, e.u, e.m, e.h[:-1])
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.close()
