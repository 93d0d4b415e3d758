"""2-D (u, rho) cooling table on the GPU (myCppSPHLibs_v3.h:794-859 applyCooling,
createCoolingGridX2.cu table fill) against the oracle."""
import numpy as np
import pytest

import oracle
from sph_2_amd import ic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table():
    from sph_2_amd.libsx import CoolingGrid
    from sph_2_amd.serve import default_cooling

    cool = default_cooling()
    # the cloud (1e4 Msun in 20 pc) has a mean n_H ~ 9 cm^-3: cover it and its
    # compressions
    grid = CoolingGrid.from_solve(cool, T_min=1e3, T_max=1e7, n_u=120, nH_min=1e-1, nH_max=1e5,
                                  n_rho=90)
    kw = cool.kw
    cp = oracle.CoolParams(kw["unit_u_cgs"], kw["unit_rho_cgs"], kw["unit_time_s"], XH=kw["XH"],
                           mu=kw["mu"], gamma=kw["gamma"], heat=kw["heat"], tol=kw["tol"],
                           T=cool.T, L=cool.L)
    un = grid.u_nodes.cpu().numpy()
    rn = grid.rho_nodes.cpu().numpy()
    args = dict(u_nodes=un, rho_nodes=rn, du=grid.du_host(), dt_ref_s=grid.dt_ref_s,
                unit_u_cgs=kw["unit_u_cgs"], unit_rho_cgs=kw["unit_rho_cgs"],
                unit_time_s=kw["unit_time_s"])
    return cool, grid, cp, args


def test_table_fill_bit_exact(gpu, table):
    cool, grid, cp, args = table
    ref = oracle.cool_grid_build(args["u_nodes"], args["rho_nodes"], grid.dt_ref_s, cp)
    assert np.array_equal(grid.du_host(), ref)


def test_apply_bit_exact(gpu, table):
    from sph_2_amd.libsx import applyCooling

    cool, grid, cp, args = table
    rng = np.random.default_rng(4)
    n = 20011
    uc = 10.0 ** rng.uniform(np.log10(args["u_nodes"][0]) - 0.3,
                             np.log10(args["u_nodes"][-1]) + 0.3, n)
    rc = 10.0 ** rng.uniform(np.log10(args["rho_nodes"][0]) - 0.5,
                             np.log10(args["rho_nodes"][-1]) + 0.5, n)
    u = uc / args["unit_u_cgs"]
    rho = rc / args["unit_rho_cgs"]
    u[:50] = args["u_nodes"][10:60] / args["unit_u_cgs"]        # on nodes
    for dt in (1e-4, 3.7e-3):
        got = applyCooling(rho, u, dt, grid)
        ref = oracle.cool_grid_apply(rho, u, dt, **args)
        assert np.array_equal(got, ref)


def test_plan_step_with_table_cooling(gpu, table):
    """FP64 plan step with table cooling against the oracle's composed step."""
    from sph_2_amd.plan import SphPlan

    cool, grid, cp, args = table
    c = ic.turbulent_cloud(2048, seed=6, ng=16)
    c.u = c.u * 30.0                                 # ~3e5 K: inside the table, cooling fast
    p = oracle.Params()
    s0 = oracle.bootstrap(oracle.State(c.pos, c.vel, c.u, c.m, c.h), p)
    s1 = oracle.step(s0, 2e-4, p, grid=args)
    pl = SphPlan(2048, mode="fp64", cooling=cool)
    pl.set_cooling_grid(grid)
    pl.load(c.pos, c.vel, c.u, c.m, c.h)
    pl.init()
    pl.step(1, 2e-4, cooling="table")
    st = pl.state()
    pl.close()
    assert np.max(np.abs(st["u"] / s1.u - 1)) < 1e-9
    assert not np.allclose(st["u"], s0.u)          # the table did cool


def test_batch_table_run_alone_equals_in_the_mix(gpu, table):
    from sph_2_amd.serve import Batch

    cool, grid, cp, args = table
    a = ic.turbulent_cloud(3000, seed=21, ng=16)
    a.u = a.u * 30.0
    b = ic.evrard(2500, seed=22)
    outs = []
    for mixed in (False, True):
        bt = Batch(20000, 8, cooling=cool)
        bt.set_cooling_grid(grid)
        if mixed:
            rb = bt.admit(b, 1e-4, cooling=True)
            bt.step(1)
        ra = bt.admit(a, 2e-4, cooling="table")
        bt.step(3)
        outs.append(bt.retire(ra))
        bt.close()
    for f in outs[0]:
        assert np.array_equal(outs[0][f], outs[1][f]), f
