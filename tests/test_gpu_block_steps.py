"""Block time steps (sphx_plan_set_block_steps) on the GPU against the oracle's
restatement (oracle.block_run), and against the fixed-dt step they reduce to."""
import numpy as np
import pytest

import oracle
from sph_2_amd import ic

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_block_steps_follow_the_oracle(gpu):
    """Evrard N = 2048, 4 levels below dt_max = 0.05, 16 substeps (two cycles of
    the longest step): the particles' levels agree, and so does the state."""
    from sph_2_amd.plan import SphPlan

    n, levels, dt_max, nsub = 2048, 3, 0.05, 16
    e = ic.evrard(n, seed=5)
    p = oracle.Params()
    s0 = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    want, bins_o, floors = oracle.block_run(s0, dt_max, levels, nsub, p)
    assert len(set(bins_o.tolist())) >= 3           # the levels are really mixed
    pl = SphPlan(n)
    pl.set_block_steps(levels)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(nsub, dt_max)
    st = pl.state()
    bins = pl.block_bins()
    t = pl.time()[0]
    c = pl.counters()
    pl.close()
    assert np.mean(bins == bins_o) > 0.999
    same = bins == bins_o
    # mixed precision against the FP64 oracle: accelerations agree to ~1e-7
    # relative, which the kicks carry into the velocities and positions
    e_pos = rel_rms(st["pos"][same], want.pos[same])
    e_vel = rel_rms(st["vel"][same], want.vel[same])
    e_u = float(np.max(np.abs(st["u"][same] / want.u[same] - 1)))
    assert e_pos < 1e-8 and e_vel < 1e-5 and e_u < 1e-6, (e_pos, e_vel, e_u)
    assert abs(t - nsub * dt_max / 2 ** levels) < 1e-15
    assert c["bin_floor"] == floors[-1]


def test_one_level_is_the_fixed_step(gpu):
    """levels = 0 is the fixed-dt KDK step, bit for bit (the velocity carries the
    opening half kick of the next step)."""
    from sph_2_amd.plan import SphPlan

    n, dt = 2048, 2e-3
    e = ic.evrard(n, seed=6)
    outs = []
    for blk in (False, True):
        pl = SphPlan(n)
        if blk:
            pl.set_block_steps(0)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        pl.step(3, dt)
        outs.append(pl.state())
        pl.close()
    glob, blk = outs
    # levels = 0 switches block stepping off: the same path
    for f in glob:
        assert np.array_equal(glob[f], blk[f]), f


def test_all_levels_equal_one_level_at_the_shortest_step(gpu):
    """A longest step far above every particle's criterion puts every particle
    on the shortest level: the fixed-dt step with dt_max / 2^levels, bit for bit
    apart from the opening half kick the velocity carries."""
    from sph_2_amd.plan import SphPlan

    n, levels = 2048, 2
    e = ic.evrard(n, seed=7)
    p = oracle.Params()
    s0 = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    big = 2 ** levels * 1.25 * oracle.particle_dt(s0.h, s0.c, s0.acc, s0.acc_g, s0.divV).max()
    pb = SphPlan(n)
    pb.set_block_steps(levels)
    pb.load(e.pos, e.vel, e.u, e.m, e.h)
    pb.init()
    pb.step(4, big)
    blk = pb.state()
    assert np.all(pb.block_bins() == levels)
    pb.close()
    pg = SphPlan(n)
    pg.load(e.pos, e.vel, e.u, e.m, e.h)
    pg.init()
    pg.step(4, big / 2 ** levels)
    glob = pg.state()
    pg.close()
    for f in ("pos", "u", "h", "rho", "acc"):
        assert np.array_equal(glob[f], blk[f]), f
    assert np.array_equal(blk["vel"], glob["vel"] + glob["acc"] * (0.5 * (big / 2 ** levels)))


def test_block_step_errors(gpu):
    from sph_2_amd._lib import SphxError
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(1024, seed=1)
    pf = SphPlan(1024, mode="fp64")
    with pytest.raises(SphxError, match="mixed"):
        pf.set_block_steps(2)
    pf.close()
    pl = SphPlan(1024)
    pl.set_adaptive_dt(True)
    with pytest.raises(SphxError, match="adaptive"):
        pl.set_block_steps(2)
    pl.set_adaptive_dt(False)
    pl.set_block_steps(2)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(2, 0.01)
    with pytest.raises(SphxError, match="longest step"):
        pl.step(1, 0.02)
    with pytest.raises(SphxError, match="inside block steps"):
        pl.set_block_steps(3)
    with pytest.raises(SphxError, match="checkpoints"):
        pl.save("/tmp/never_written.ckpt")
    pl.init()                                       # a new start allows new settings
    pl.set_block_steps(3)
    pl.step(1, 0.02)
    pl.close()
