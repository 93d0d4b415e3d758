"""Served runs on the GPU: a run's result is the same alone and in the mix,
graph replays match eager steps, admission is refused beyond the budget."""
import numpy as np
import pytest

from sph_2_amd import ic

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "u", "h", "rho", "acc", "k")


def _same(a, b):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f


@pytest.fixture(scope="module")
def runs():
    from sph_2_amd.serve import default_cooling

    a = ic.evrard(3000, seed=11)
    b = ic.turbulent_cloud(5000, seed=12, ng=32)
    c = ic.evrard(2100, seed=13)
    return a, b, c, default_cooling()


def test_alone_equals_in_the_mix(gpu, runs):
    from sph_2_amd.serve import Batch

    a, b, c, cool = runs
    solo = Batch(20000, 16, cooling=cool)
    ra = solo.admit(a, 1e-4)
    solo.step(5)
    alone = solo.retire(ra)
    solo.close()

    mix = Batch(20000, 16, cooling=cool)
    rb = mix.admit(b, 2e-4, cooling=True)
    mix.step(2)
    ra = mix.admit(a, 1e-4)
    mix.step(2)
    rc = mix.admit(c, 1e-4)
    mix.step(1)
    mix.retire(rb)                 # a's segment moves down when b leaves
    mix.step(2)
    in_mix = mix.retire(ra)
    mix.retire(rc)
    mix.close()
    _same(alone, in_mix)


def test_graph_replays_match_eager(gpu, runs):
    """Every replay checked against the eager twin, with eager work and
    synchronisations between replays."""
    from sph_2_amd.serve import Batch

    torch = gpu
    a, b, c, cool = runs
    eager = Batch(20000, 16, cooling=cool)
    graph = Batch(20000, 16, cooling=cool, graph=True)
    ids = []
    for bt in (eager, graph):
        ids.append((bt.admit(a, 1e-4), bt.admit(b, 2e-4, cooling=True)))
    x = torch.randn(256, 256, device="cuda")
    for step in range(6):
        eager.step(1)
        graph.step(1)
        y = (x @ x).sum()          # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(y)
        for k in range(2):
            _same(eager.store(ids[0][k]), graph.store(ids[1][k]))
        if step == 2:              # the run set changes: the graph is recaptured
            for bt, idx in ((eager, ids[0]), (graph, ids[1])):
                bt.admit(c, 1e-4)
    eager.close()
    graph.close()


def test_admission_refused_beyond_budget(gpu, runs):
    from sph_2_amd.serve import Batch

    a, b, c, cool = runs
    bt = Batch(8000, 2, cooling=cool)
    ra = bt.admit(a, 1e-4)
    assert ra is not None
    assert bt.admit(b, 1e-4) is not None
    assert bt.admit(c, 1e-4) is None          # run slots exhausted
    bt.retire(ra)
    big = ic.evrard(6000, seed=1)
    assert bt.admit(big, 1e-4) is None         # particle capacity exhausted
    assert bt.admit(c, 1e-4) is not None
    bt.close()


def test_plan_graph_matches_eager(gpu):
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(6000, seed=5)
    outs = []
    for g in (False, True):
        pl = SphPlan(6000)
        if g:
            pl.set_graph(True)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        for _ in range(3):
            pl.step(2, 1e-4)
        outs.append(pl.state())
        pl.close()
    for f in outs[0]:
        assert np.array_equal(outs[0][f], outs[1][f]), f


def test_serving_loop_small(gpu):
    from sph_2_amd.serve import Batch, Server, make_stream

    reqs = make_stream(12, n_min=2048, n_max=16384, rate=50.0, steps=20, seed=3)
    srv = Server(hbm_budget_bytes=40000 * Batch.bytes_per_particle(), max_runs=8, graph=True)
    rep = srv.run(reqs, keep_results=True)
    assert rep["runs"] == 12
    assert rep["p99_latency_s"] >= rep["p50_latency_s"] > 0
    # every result equals the same request run alone
    from sph_2_amd.serve import Batch, build_payload, default_cooling

    for r in reqs[:3]:
        solo = Batch(20000, 8, cooling=default_cooling())
        solo.set_gravity_symmetric(srv.symmetric[0], srv.sym_budget)
        rid = solo.admit(build_payload(r), r.dt, cooling=r.kind == "cloud")
        solo.step(r.steps)
        alone = solo.retire(rid)
        solo.close()
        _same(alone, rep["results"][r.rid])


def test_symmetric_gravity_alone_equals_in_the_mix(gpu, runs):
    """With the symmetric gravity sum a run's result still does not depend on
    the other runs; its partial sums stream through the reserved memory in row
    bands, so the result does not depend on the reservation either (one band
    or one row per band), and only a run whose largest row does not fit is
    refused."""
    from sph_2_amd.serve import Batch

    a, b, c, cool = runs
    need = Batch.symmetric_min_bytes(a.n, 2)
    assert 0 < need < Batch.symmetric_bytes(a.n, 2)
    outs = []
    for mixed in (False, True):
        bt = Batch(60000, 8, cooling=cool)
        bt.set_gravity_symmetric(2, 64 * 2 ** 20)
        if mixed:
            rb = bt.admit(b, 1e-4)
            bt.step(2)
        ra = bt.admit(a, 2e-4)
        bt.step(2)
        if mixed:
            bt.retire(rb)
            rc = bt.admit(c, 1e-4, cooling=True)
        bt.step(2)
        outs.append(bt.retire(ra))
        bt.close()
    _same(outs[0], outs[1])
    tiny = Batch(60000, 8)
    tiny.set_gravity_symmetric(2, need - 1)
    assert tiny.admit(a, 1e-4) is None             # its first row does not fit
    tiny.close()
    banded = []
    for budget in (need, 64 * 2 ** 20):            # a band per row / one band
        bt = Batch(60000, 8)
        bt.set_gravity_symmetric(2, budget)
        ra = bt.admit(a, 2e-4)
        assert ra is not None
        assert bt.admit(c, 1e-4) is not None       # another run streams through too
        bt.step(3)
        banded.append(bt.retire(ra))
        bt.close()
    _same(banded[0], banded[1])


@pytest.mark.parametrize("slots", [8, 512])
def test_batch_cuts_buckets_like_a_plan(gpu, slots):
    """A batch's keys give each axis fewer Morton bits than a plan's (the run
    id takes the top bits: 20 per axis for 8 run slots, 18 for 512), and the
    bucket cuts must read the level of a key difference in that layout: then
    a run's neighbour walks visit the same candidate buckets in a batch as in
    a plan at the same cut level.  A uniform cloud (no two particles share a
    cell at 18 bits) and the bootstrap's walks on the initial positions, so
    both sort alike."""
    from sph_2_amd.plan import SphPlan
    from sph_2_amd.serve import Batch

    e = ic.turbulent_cloud(16384, seed=3, ng=32)
    pl = SphPlan(e.n)
    pl.set_cell_level(0)                           # a batch's level (no tuning)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    cp = pl.counters()
    pl.close()
    bt = Batch(2 * e.n, slots)
    assert bt.admit(e, 1e-4) is not None
    cb = bt.counters()
    bt.close()
    assert cb["buckets"] == cp["buckets"] == 256
    assert cb["h_iters_sum"] == cp["h_iters_sum"]
    assert cb["cand_buckets"] == cp["cand_buckets"]


def test_batch_on_a_side_stream(gpu, runs):
    """The batch calls on a non-default stream (the caller's current stream):
    admitting and retiring runs rewrites the run tables while steps queued on
    that stream may still read them, so the upload is ordered on the caller's
    stream.  The results equal the same sequence on the default stream."""
    import torch

    from sph_2_amd.serve import Batch

    a, b, c, cool = runs
    outs = []
    for side in (False, True):
        s = torch.cuda.Stream() if side else torch.cuda.current_stream()
        with torch.cuda.stream(s):
            bt = Batch(20000, 8, cooling=cool, graph=True)
            ra = bt.admit(a, 1e-4)
            bt.step(2)                             # queued, not waited for
            rb = bt.admit(b, 2e-4, cooling=True)   # new tables behind the queued steps
            bt.step(2)
            bt.retire(rb)
            rc = bt.admit(c, 1e-4)
            bt.step(2)
            outs.append(bt.retire(ra))
            bt.close()
        torch.cuda.synchronize()
    _same(outs[0], outs[1])
