"""Failure paths forced from the host (no kernel is made to fault): coincident
particles drive the smoothing-length iteration into its cap, which is counted
and reported instead of printed; a non-finite state is counted, and in a
served batch it stays inside its own run."""
import warnings

import numpy as np
import pytest

from sph_2_amd import ic

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "u", "h", "rho", "acc", "k")


def test_coincident_particles_cap_the_h_iteration(gpu):
    from sph_2_amd import libsx
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(2048, seed=3)
    pos = e.pos.copy()
    pos[:200] = [0.3, 0.1, -0.05]               # 200 particles on one point
    pl = SphPlan(2048)
    pl.load(pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(2, 1e-4)
    c = pl.counters()
    st = pl.state()
    pl.close()
    assert c["h_capped"] > 0
    # the cap leaves the coincident particles a vanishing h, as the reference's
    # does; whatever that does to the state (their density overflows, and the
    # direct-sum gravity then carries a non-finite position to every particle)
    # is what the counter reports, particle for particle
    bad = ~(np.isfinite(st["pos"]).all(1) & np.isfinite(st["vel"]).all(1) &
            np.isfinite(st["u"]))
    assert c["nonfinite"] == int(bad.sum())
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        h, k, it = libsx.smoothing_h(pos, e.h)
    assert np.all(np.isfinite(h))
    assert any(issubclass(x.category, RuntimeWarning) for x in w)


def test_nonfinite_state_is_counted_and_stays_in_its_run(gpu):
    from sph_2_amd.serve import Batch

    a = ic.evrard(3000, seed=41)
    b = ic.evrard(2500, seed=42)
    b.u = b.u.copy()
    b.u[7] = np.nan                              # a corrupt request
    outs = []
    for mixed in (False, True):
        bt = Batch(20000, 8)
        if mixed:
            rb = bt.admit(b, 1e-4)
        ra = bt.admit(a, 1e-4)
        bt.step(3)
        c = bt.counters()
        assert (c["nonfinite"] > 0) == mixed
        outs.append(bt.retire(ra))
        if mixed:
            bad = bt.retire(rb)
            assert not np.all(np.isfinite(bad["u"]))
        bt.close()
    for f in FIELDS:
        assert np.array_equal(outs[0][f], outs[1][f]), f
