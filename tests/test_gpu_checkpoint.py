"""Exact checkpoints: a saved and restored run continues bit-identically to the
run that was never stopped (plan and served batch), and the reference's
restart CSV feeds the plan."""
import numpy as np
import pytest

from sph_2_amd import ic

pytestmark = pytest.mark.gpu


def _same(a, b):
    for f in a:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_plan_restore_continues_bit_identically(gpu, tmp_path, mode):
    from sph_2_amd.plan import SphPlan, checkpoint_info

    e = ic.evrard(3000, seed=8)
    ck = tmp_path / "run.ckpt"
    a = SphPlan(3000, mode=mode)
    a.set_adaptive_dt(True, dt_min=1e-9, dt_max=1.0)
    a.load(e.pos, e.vel, e.u, e.m, e.h)
    a.init()
    a.step(3, 2e-4)
    a.save(ck)
    before = a.state()
    a.step(2, 2e-4)
    want, t_want = a.state(), a.time()
    a.close()
    info = checkpoint_info(ck)
    assert info["n"] == 3000 and info["mode"] == mode and info["steps"] == 3

    b = SphPlan(4096, mode=mode)                 # larger capacity is fine
    b.restore(ck)
    _same(before, b.state())                     # restored state is the saved state
    b.step(2, 123.0)                              # adaptive: the dt argument is ignored
    _same(want, b.state())
    assert b.time() == t_want
    b.close()


def test_batch_run_saved_and_restored_elsewhere(gpu, tmp_path):
    from sph_2_amd.serve import Batch, default_cooling

    a = ic.evrard(3000, seed=11)
    b = ic.turbulent_cloud(4000, seed=12, ng=32)
    c = ic.evrard(2000, seed=13)
    cool = default_cooling()
    solo = Batch(20000, 8, cooling=cool)
    rb = solo.admit(b, 2e-4, cooling=True)
    solo.step(5)
    uninterrupted = solo.retire(rb)
    solo.close()

    ck = tmp_path / "cloud.ckpt"
    first = Batch(20000, 8, cooling=cool)
    ra = first.admit(a, 1e-4)
    rb = first.admit(b, 2e-4, cooling=True)
    first.step(3)
    first.save(rb, ck)                           # preempt the cloud after 3 steps
    first.retire(rb)
    first.step(1)
    first.close()

    second = Batch(20000, 8, cooling=cool, graph=True)
    rc = second.admit(c, 1e-4)
    second.step(1)
    rb2 = second.restore(ck)
    assert rb2 is not None
    second.step(2)
    resumed = second.retire(rb2)
    second.retire(rc)
    second.close()
    _same(uninterrupted, resumed)


def test_batch_restore_refused_when_full(gpu, tmp_path):
    from sph_2_amd.serve import Batch, default_cooling

    a = ic.evrard(3000, seed=11)
    cool = default_cooling()
    bt = Batch(5000, 4, cooling=cool)
    ra = bt.admit(a, 1e-4)
    ck = tmp_path / "a.ckpt"
    bt.save(ra, ck)
    assert bt.restore(ck) is None                # 3000 + 3000 > 5000
    bt.retire(ra)
    assert bt.restore(ck) is not None
    bt.close()


def test_plan_from_reference_restart_csv(gpu, tmp_path):
    """A restart CSV (x..u, dudt) written by the library loads into a plan whose
    first step matches one loaded from the same arrays directly."""
    from sph_2_amd import formats
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(2048, seed=2)
    f = tmp_path / "RESTART_0.csv"
    formats.write_restart_csv(f, e.pos, e.vel, e.m, e.h, e.h, e.u, np.zeros(2048))
    r = formats.read_ic_csv(f, restart=True)
    outs = []
    for src in (dict(pos=e.pos, vel=e.vel, u=e.u, m=e.m, h=e.h),
                dict(pos=r["pos"], vel=r["vel"], u=r["u"], m=r["m"], h=r["h"])):
        pl = SphPlan(2048)
        pl.load(**src)
        pl.init()
        pl.step(1, 1e-4)
        outs.append(pl.state())
        pl.close()
    _same(outs[0], outs[1])
