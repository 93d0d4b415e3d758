"""Global adaptive time step on the GPU (myCppSPHLibs_v2.h:491-522 dt_array and
the host reduction of hfv_sph_gpu_v2_evrard.cu:517-561) against the oracle."""
import numpy as np
import pytest

import oracle
from sph_2_amd import ic

pytestmark = pytest.mark.gpu

N = 2048


def test_drop_in_bit_exact_random(gpu):
    from sph_2_amd import libsx

    rng = np.random.default_rng(17)
    n = 10007                                   # ragged last bucket
    h = rng.random(n) * 0.1 + 1e-3
    c = rng.random(n)
    acc = rng.normal(size=(n, 3))
    acc_g = rng.normal(size=(n, 3)) * 3
    divV = rng.normal(size=n)
    divV[::7] = 0.0                             # infinite density limit on those rows
    for args in ((0.25, 1e-5, 1e-3), (0.3, 1e-12, 10.0)):
        got = libsx.global_dt(h, c, acc, acc_g, divV, *args)
        ref = oracle.global_dt(h, c, acc, acc_g, divV, *args)
        assert got == ref


def test_drop_in_bit_exact_on_evrard_forces(gpu):
    from sph_2_amd import libsx

    e = ic.evrard(4096, seed=0)
    s = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), oracle.Params())
    s = oracle.step(s, 1e-4, oracle.Params())
    got = libsx.global_dt(s.h, s.c, s.acc, s.acc_g, s.divV, 0.25, 1e-9, 1.0)
    ref = oracle.global_dt(s.h, s.c, s.acc, s.acc_g, s.divV, 0.25, 1e-9, 1.0)
    assert got == ref
    assert 1e-9 < got[0] < 1.0                  # the criterion, not a clamp, decided


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_plan_adaptive_follows_the_oracle(gpu, mode, tol):
    """Seed dt, then each step's dt from the previous step's forces, as the
    driver's loop does; simulated time is the sum of the dt used."""
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(N, seed=3)
    p = oracle.Params()
    lim = dict(c_cfl=0.25, dt_min=1e-9, dt_max=1.0)
    s = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    pl = SphPlan(N, mode=mode)
    pl.set_adaptive_dt(True, **lim)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    dt = 2e-4                                    # the driver's starting dt
    t_ref = 0.0
    for k in range(4):
        s = oracle.step(s, dt, p)
        t_ref += dt
        dt = oracle.next_dt(s, **lim)
        pl.step(1, 2e-4 if k == 0 else 123.0)    # later dt arguments are ignored
        t, dt_gpu, parts = pl.time()
        assert abs(t - t_ref) <= tol * t_ref
        assert abs(dt_gpu - dt) <= tol * dt, (k, dt_gpu, dt, parts)
    st = pl.state()
    assert np.sqrt(((st["pos"] - s.pos) ** 2).sum() / (s.pos ** 2).sum()) < 1e-9
    pl.close()


def test_plan_fixed_dt_time_and_switch(gpu):
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(N, seed=4)
    pl = SphPlan(N)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(3, 1e-4)
    t, dt, _ = pl.time()
    assert t == pytest.approx(3e-4, rel=1e-15) and dt == 1e-4
    pl.set_adaptive_dt(True, dt_min=1e-9, dt_max=1.0)
    pl.step(1, 5e-5)                             # seeds the first adaptive step
    t, dt, parts = pl.time()
    assert t == pytest.approx(3.5e-4, rel=1e-15)
    assert dt == pytest.approx(0.25 * min(parts), rel=1e-15)
    pl.set_adaptive_dt(False)
    pl.step(1, 1e-4)
    t2, dt2, _ = pl.time()
    assert dt2 == 1e-4 and t2 == pytest.approx(t + 1e-4, rel=1e-15)
    pl.close()


def test_plan_adaptive_graph_matches_eager(gpu):
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(6000, seed=5)
    outs = []
    for g in (False, True):
        pl = SphPlan(6000)
        pl.set_graph(g)
        pl.set_adaptive_dt(True, dt_min=1e-9, dt_max=1.0)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        times = []
        for _ in range(3):
            pl.step(2, 1e-4)
            times.append(pl.time())
        outs.append((pl.state(), times))
        pl.close()
    assert outs[0][1] == outs[1][1]
    for f in outs[0][0]:
        assert np.array_equal(outs[0][0][f], outs[1][0][f]), f


def test_batch_adaptive_run_alone_equals_in_the_mix(gpu):
    from sph_2_amd.serve import Batch, default_cooling

    a = ic.evrard(3000, seed=11)
    b = ic.turbulent_cloud(5000, seed=12, ng=32)
    cool = default_cooling()
    solo = Batch(20000, 8, cooling=cool)
    ra = solo.admit(a, 1e-4)
    solo.set_adaptive_dt(ra, dt_min=1e-9, dt_max=1.0)
    solo.step(4)
    t_alone = solo.time(ra)
    alone = solo.retire(ra)
    solo.close()

    mix = Batch(20000, 8, cooling=cool)
    rb = mix.admit(b, 2e-4, cooling=True)
    mix.step(1)
    ra = mix.admit(a, 1e-4)
    mix.set_adaptive_dt(ra, dt_min=1e-9, dt_max=1.0)
    mix.step(2)
    mix.retire(rb)                               # compaction moves a's segment and params
    mix.step(2)
    assert mix.time(ra) == t_alone
    in_mix = mix.retire(ra)
    mix.close()
    for f in alone:
        assert np.array_equal(alone[f], in_mix[f]), f
