"""GPU parity of the drop-in functions against the FP64 oracle.

Workload: Evrard sphere N = 4096, seed 0 (the reference's one-step CPU
case), plus a small random cube.  Bars:
  neighbour counts / sets / smoothing lengths  bit-exact
  density                                       1e-6 relative (max)
  hydro and gravity accelerations               1e-5 relative RMS
  FP64 baseline kernels                         1e-12 relative
"""
import warnings

import numpy as np
import pytest

import oracle
from sph_2_amd import ic

pytestmark = pytest.mark.gpu

N = 4096


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def rel_max(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.fixture(scope="module")
def evr():
    e = ic.evrard(N, seed=0)
    h, k, it, st = oracle.smoothing(e.pos, e.h)
    assert st.sum() == 0
    rho, nn = oracle.density(e.pos, e.m, h)
    rng = np.random.default_rng(1)
    v = 0.3 * rng.standard_normal((N, 3))       # non-trivial velocities for AV paths
    u = e.u * (1.0 + 0.2 * rng.random(N))
    P, c = oracle.eos(rho, u, 5.0 / 3.0)
    dv, cv = oracle.divcurl(e.pos, v, rho, e.m, h)
    return dict(pos=e.pos, m=e.m, h0=e.h, h=h, k=k, rho=rho, nn=nn, v=v, u=u, P=P, c=c, dv=dv,
                cv=cv)


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_smoothing_bit_exact(gpu, evr, mode):
    from sph_2_amd import libsx

    h, k, it = libsx.smoothing_h(evr["pos"], evr["h0"], mode=mode)
    assert np.array_equal(k, evr["k"])
    assert np.array_equal(h, evr["h"])


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
@pytest.mark.parametrize("h_mode", [2, 0, 4])
def test_smoothing_predictor_rows(gpu, evr, mode, h_mode):
    """smoothing_fast_new_mpi3 on rows [100, 900) from the eq. 31 start: by
    default the reference's own rule (steps of 0.002 h), with h_mode=0 the
    count-band iteration, with h_mode=4 the typed driver's rule (steps of
    0.01 h halved on every jump across the band, LIBV7:255-328); h and counts
    bit-exact against the oracle's rule."""
    from sph_2_amd import libsx

    kprev = np.clip(evr["k"] + np.arange(N) % 7 - 3, 1, None)
    dt = 1e-3
    ho, ko, _, _ = oracle.smoothing(evr["pos"], evr["h"], kprev, evr["dv"], dt,
                                    p=oracle.Params(h_mode=h_mode), rows=np.arange(100, 900))
    h, kk = libsx.smoothing_fast_new_mpi3(100, 900, evr["pos"], evr["h"], kprev, evr["dv"], dt,
                                          mode=mode, h_mode=h_mode)
    assert np.array_equal(kk.astype(np.int32), ko)
    assert np.array_equal(h, ho)


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_reference_h_rule_drop_ins(gpu, evr, mode):
    """smoothing_fast_new_mpi2 (steps of 0.002 h from 2h, LIBSX:940-990) and
    do_smoothingX4 (steps of 0.02 h) from the IC's guess, where many rows run
    into the reference's limit of 101 counts and take their largest radius:
    every row bit-exact against the oracle's restatement, with the warning
    the reference prints as a RuntimeWarning."""
    from sph_2_amd import libsx

    for fn, h_mode in ((lambda: libsx.smoothing_fast_new_mpi2(0, N, evr["pos"], evr["h0"],
                                                              mode=mode), 2),
                       (lambda: libsx.do_smoothingX4(evr["pos"], evr["h0"], mode=mode), 3)):
        ho, ko, it, st = oracle.smoothing(evr["pos"], evr["h0"], p=oracle.Params(h_mode=h_mode))
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            h = fn()
        assert np.array_equal(h, ho), h_mode
        assert st.any() == any(issubclass(x.category, RuntimeWarning) for x in w)
    assert st.sum() < N  # most rows of do_smoothingX4 land in the band


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_halving_h_rule_from_the_guess(gpu, evr, mode):
    """The typed driver's rule (LIBV7:255-328: steps of 0.01 h from 2h, halved
    on every jump across the band, 501 counts at most, counting stopped past
    69) from the IC's guess: h, counts and pass counts bit-exact against the
    oracle on every row."""
    from sph_2_amd import libsx

    ho, ko, ito, st = oracle.smoothing(evr["pos"], evr["h0"], p=oracle.Params(h_mode=4))
    h, k, it = libsx.smoothing_h(evr["pos"], evr["h0"], mode=mode, h_mode=4)
    assert np.array_equal(h, ho)
    assert np.array_equal(k, ko) and np.array_equal(it, ito)
    assert k.max() <= 70


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_h_smooth_fast_rule(gpu, evr, mode):
    """h_smooth_fast (LIBSX:1133-1184: h -+= 0.003 h on a strict count, no
    pass limit) from the IC's guess, all rows and a row range through
    h_smooth_fast_mpi, and with augsphx6.5x.py's 50 +- 10 neighbours: h
    bit-exact against the oracle."""
    from sph_2_amd import libsx

    ho, ko, ito, st = oracle.smoothing(evr["pos"], evr["h0"], p=oracle.Params(h_mode=5))
    assert not st.any() and ito.max() > 20
    assert np.array_equal(libsx.h_smooth_fast(evr["pos"], evr["h0"], mode=mode), ho)
    h, k, it = libsx.smoothing_h(evr["pos"], evr["h0"], mode=mode, h_mode=5)
    assert np.array_equal(k, ko) and np.array_equal(it, ito)
    assert np.array_equal(libsx.h_smooth_fast_mpi(100, 900, evr["pos"], evr["h0"], mode=mode),
                          ho[100:900])
    ho, _, _, _ = oracle.smoothing(evr["pos"], evr["h0"],
                                   p=oracle.Params(h_mode=5, nngb=50, ndown=40, nup=60))
    assert np.array_equal(libsx.h_smooth_fast(evr["pos"], evr["h0"], mode=mode, nngb=50, ndown=40,
                                              nup=60), ho)


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_h_smooth_fast_floor(gpu, evr, mode):
    """h_smooth_fast_h_minimum_set and its row-range twin (LIBSX:1192-1370)
    with a floor at the 30th percentile of the start h, which many rows hit:
    h bit-exact against the oracle."""
    from sph_2_amd import libsx

    hmin = float(np.quantile(evr["h0"], 0.3))
    ho, ko, ito, st = oracle.smoothing(evr["pos"], evr["h0"], p=oracle.Params(h_mode=5, h_min=hmin))
    assert (ho == hmin).sum() > 50
    assert np.array_equal(libsx.h_smooth_fast_h_minimum_set(evr["pos"], evr["h0"], hmin, mode=mode),
                          ho)
    assert np.array_equal(libsx.h_smooth_fast_mpi_min_h_set(300, 1300, evr["pos"], evr["h0"], hmin,
                                                            mode=mode), ho[300:1300])


@pytest.mark.parametrize("h_mode,tol", [(2, 2.1e-3), (4, 1.1e-2), (5, 6.1e-3)])
def test_plan_reference_h_rule(gpu, h_mode, tol):
    """A plan with h_mode=2 (the reference driver's smoothing_h rule,
    LIBV2:7-64, DRV2:419), h_mode=4 (the typed driver's, LIBV7:255-328) or
    h_mode=5 (h_smooth_fast each step, augsphx6.5x.py:590): after the
    bootstrap h and counts bit-exact against the oracle's composition; after
    a step, whose start carries the previous step's state, within two steps
    of the rule (0.001 h, 0.005 h or 0.3 %)."""
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(N, seed=0)
    p = oracle.Params(h_mode=h_mode)
    s0 = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    s1 = oracle.step(s0, 1e-4, p)
    for mode in ("mixed", "fp64"):
        pl = SphPlan(N, mode=mode, h_mode=h_mode)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        st0 = pl.state()
        assert np.array_equal(st0["h"], s0.h) and np.array_equal(st0["k"], s0.k)
        pl.step(1, 1e-4)
        st1 = pl.state()
        assert np.mean(st1["h"] != s1.h) < 0.01
        assert rel_max(st1["h"], s1.h) < tol
        pl.close()


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_density(gpu, evr, mode):
    from sph_2_amd import libsx

    rho, cnt = libsx.getDensity(evr["pos"], evr["m"], evr["h"], mode=mode, return_count=True)
    assert np.array_equal(cnt, evr["nn"])
    tol = 1e-6 if mode == "mixed" else 1e-12
    assert rel_max(rho, evr["rho"]) < tol


def test_density_rows(gpu, evr):
    from sph_2_amd import libsx

    rho = libsx.getDensity_mpi(1000, 1500, evr["pos"], evr["m"], evr["h"])
    assert rel_max(rho, evr["rho"][1000:1500]) < 1e-6


def test_neighbor_sets_bit_exact(gpu, evr):
    from sph_2_amd import libsx

    rows = (0, 600)
    off, idx = libsx.neighbors(evr["pos"], evr["h"], nbeg=rows[0], nend=rows[1])
    ooff, oidx = oracle.neighbors(evr["pos"], evr["h"], rows=np.arange(*rows))
    assert np.array_equal(off, ooff)
    assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_divcurl(gpu, evr, mode):
    from sph_2_amd import libsx

    dv, cv = libsx.div_curlVel(evr["pos"], evr["v"], evr["rho"], evr["m"], evr["h"], mode=mode)
    tol = 1e-5 if mode == "mixed" else 1e-12
    assert rel_rms(dv, evr["dv"]) < tol
    assert rel_rms(cv, evr["cv"]) < tol


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
@pytest.mark.parametrize("av_form,balsara", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_hydro(gpu, evr, mode, av_form, balsara):
    from sph_2_amd import libsx

    p = oracle.Params(av_form=av_form, balsara=balsara)
    acc_o, du_o = oracle.hydro(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"],
                               evr["m"], evr["dv"], evr["cv"], p)
    acc, du = libsx.getAcc_sph_dU(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"],
                                  evr["m"], av_form=av_form, balsara=balsara, divV=evr["dv"],
                                  curlV=evr["cv"], mode=mode)
    tol = 1e-5 if mode == "mixed" else 1e-12
    assert rel_rms(acc, acc_o) < tol
    assert rel_rms(du, du_o) < tol


def test_reference_signatures(gpu, evr):
    from sph_2_amd import libsx

    g = 5.0 / 3.0
    a = libsx.getAcc_sph(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"], evr["m"],
                         g, 0.1, 1.0, 2.0)
    ao, duo = oracle.hydro(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"],
                           evr["m"], p=oracle.Params(av_form=1))
    assert a.shape == (N, 3)
    assert rel_rms(a, ao) < 1e-5
    du = libsx.get_dU(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"], evr["m"],
                      g, 0.1, 1.0, 2.0)
    assert rel_rms(du, duo) < 1e-5
    P = libsx.getPressure(evr["rho"], evr["u"], g)
    assert rel_max(P, evr["P"]) < 1e-15
    ash = libsx.getAcc_sph_shear(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"],
                                 evr["m"], evr["dv"], evr["cv"], 1.0)
    ao2, _ = oracle.hydro(evr["pos"], evr["v"], evr["rho"], evr["P"], evr["c"], evr["h"],
                          evr["m"], evr["dv"], evr["cv"], oracle.Params(av_form=1, balsara=1))
    assert rel_rms(ash, ao2) < 1e-5


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_gravity(gpu, evr, mode):
    from sph_2_amd import libsx

    acc_o, pot_o = oracle.gravity(evr["pos"], evr["m"], evr["h"], 1.0, potential=True)
    acc, pot = libsx.getAcc_g_smth(evr["pos"], evr["m"], 1.0, evr["h"], mode=mode,
                                   return_potential=True)
    tol = 1e-5 if mode == "mixed" else 1e-12
    assert rel_rms(acc, acc_o) < tol
    assert rel_rms(pot, pot_o) < tol


@pytest.mark.parametrize("cfg", [(1, 256, 1), (1, 512, 3), (2, 256, 16), (2, 512, 8),
                                 (2, 1024, 2), (4, 256, 4), (4, 512, 1)])
def test_gravity_launch_shapes(gpu, evr, cfg):
    """Every launch shape the tuner may pick (particles per lane, j-tile, j
    split with partial sums) gives the same forces and potential."""
    from sph_2_amd import libsx

    acc_o, pot_o = oracle.gravity(evr["pos"], evr["m"], evr["h"], 1.0, potential=True)
    libsx.set_gravity_config(*cfg)
    try:
        acc, pot = libsx.getAcc_g_smth(evr["pos"], evr["m"], 1.0, evr["h"], return_potential=True)
    finally:
        libsx.set_gravity_config()
    assert rel_rms(acc, acc_o) < 1e-5
    assert rel_rms(pot, pot_o) < 1e-5


@pytest.mark.parametrize("n", [4096, 3001, 2200, 300])
@pytest.mark.parametrize("masses", ["equal", "random"])
@pytest.mark.parametrize("eps", ["h", 0.002])
def test_gravity_symmetric(gpu, evr, n, masses, eps):
    """The symmetric sum (each pair once, applied to both particles) against
    the oracle: far block pairs with one common mass, with mixed masses, near
    and diagonal block pairs, partial blocks (n = 3001: a last block of 441,
    whose second half of 256 is partial; 2200: a last block of 152, with no
    second half; 300: one partial block), segment lengths 1, 3 and 64, and
    block pairs split into parts of 256 and 128 j particles."""
    from sph_2_amd import libsx

    rng = np.random.default_rng(n)
    pos = evr["pos"][:n]
    m = evr["m"][:n] if masses == "equal" else evr["m"][:n] * rng.uniform(0.5, 1.5, n)
    e = evr["h"][:n] if eps == "h" else np.full(n, eps)
    acc_o = oracle.gravity(pos, m, e, 1.0)
    try:
        for tj, S in ((512, 1), (512, 3), (512, 64), (256, 1), (128, 1)):
            libsx.set_gravity_config(8, tj, S)
            acc = libsx.getAcc_g_smth(pos, m, 1.0, e)
            assert rel_rms(acc, acc_o) < 1e-5, (tj, S)
            assert np.max(np.linalg.norm(acc - acc_o, axis=1) /
                          np.linalg.norm(acc_o, axis=1)) < 1e-4, (tj, S)
            # the potential call falls back to the tiled sum
            acc2, pot = libsx.getAcc_g_smth(pos, m, 1.0, e, return_potential=True)
            assert rel_rms(acc2, acc_o) < 1e-5
    finally:
        libsx.set_gravity_config()


def test_plan_step_symmetric_gravity(gpu):
    """Plan bootstrap + one step with the symmetric sum vs the oracle's composed
    step; graph replays equal eager steps bit for bit."""
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(N, seed=0)
    p = oracle.Params()
    s0 = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    s1 = oracle.step(s0, 1e-4, p)
    outs = []
    for graph in (False, True):
        pl = SphPlan(N)
        pl.set_gravity_config(8, 512, 2)
        pl.set_graph(graph)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        pl.step(1, 1e-4)
        st1 = pl.state()
        assert rel_rms(st1["acc_g"], s1.acc_g) < 1e-5
        assert rel_rms(st1["acc"], s1.acc) < 1e-5
        assert rel_max(st1["rho"], s1.rho) < 1e-6
        pl.step(3, 1e-4)
        outs.append(pl.state())
        pl.close()
    for f in outs[0]:
        assert np.array_equal(outs[0][f], outs[1][f]), f


def test_gravity_rows_and_fixed_eps(gpu, evr):
    from sph_2_amd import libsx

    rows = np.arange(200, 1200)
    acc_o = oracle.gravity(evr["pos"], evr["m"], 0.02, 1.0, rows=rows)
    acc = libsx.getAcc_g_smth_mpi(200, 1200, evr["pos"], evr["m"], 1.0, 0.02)
    assert rel_rms(acc, acc_o) < 1e-5
    acc64 = libsx.getAcc_g_smth_mpi(200, 1200, evr["pos"], evr["m"], 1.0, 0.02, mode="fp64")
    assert rel_rms(acc64, acc_o) < 1e-12


@pytest.mark.parametrize("n", [50_000, 1_500_001])
def test_morton_keys_and_sort(gpu, n):
    """Keys and the stable sort permutation against a numpy restatement; at
    1,500,001 particles the radix sort's 367 tile histograms give its
    per-digit scan two entries per thread and a partial last tile."""
    from sph_2_amd import libsx

    rng = np.random.default_rng(3)
    pos = rng.random((n, 3))
    pos[::7] = pos[3]              # ties: equal keys must keep their order
    keys, perm = libsx.morton_sort(pos)
    lo = pos.min(0)
    ext = (pos.max(0) - lo).max()
    scale = float(2 ** 21) / (ext * (1.0 + 2.0 ** -30))
    q = np.clip(np.floor((pos - lo) * scale), 0, 2 ** 21 - 1).astype(np.uint64)
    ref = np.zeros(n, np.uint64)
    for b in range(21):
        for a, sh in ((0, 2), (1, 1), (2, 0)):
            ref |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + sh)
    order = np.argsort(ref, kind="stable")
    assert np.array_equal(perm, order)
    assert np.array_equal(keys, ref[order])


def test_cooling_bit_exact(gpu):
    from sph_2_amd import libsx

    T, L = ic.cooling_table(seed=0)
    units = ic.cloud_units()
    tab = libsx.CoolingTable(T, L, unit_u_cgs=units["unit_u_cgs"],
                             unit_rho_cgs=units["unit_rho_cgs"], unit_time_s=units["unit_time_s"],
                             heat=2e-26)
    cp = oracle.CoolParams(units["unit_u_cgs"], units["unit_rho_cgs"], units["unit_time_s"],
                           mu=1.22, heat=2e-26, T=T, L=L)
    rng = np.random.default_rng(5)
    n = 20000
    rho = 10 ** rng.uniform(-3, 3, n)
    u = 10 ** rng.uniform(-4, 2, n)
    out = libsx.DoCooling(rho, u, 1e-4, tab)
    ref, _ = oracle.cool(rho, u, 1e-4, cp)
    assert np.array_equal(out, ref)


def test_plan_one_step(gpu):
    """Resident plan: bootstrap + one KDK step vs the oracle's composition."""
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(N, seed=0)
    p = oracle.Params()
    s0 = oracle.bootstrap(oracle.State(e.pos, e.vel, e.u, e.m, e.h), p)
    s1 = oracle.step(s0, 1e-4, p)
    for mode, tol in (("mixed", 1e-5), ("fp64", 1e-11)):
        pl = SphPlan(N, mode=mode)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        st0 = pl.state()
        assert np.array_equal(st0["h"], s0.h)
        assert rel_max(st0["rho"], s0.rho) < (1e-6 if mode == "mixed" else 1e-12)
        assert rel_rms(st0["acc"], s0.acc) < tol
        pl.step(1, 1e-4)
        st1 = pl.state()
        # After a step, h starts from the Gadget-2 predictor, which uses div v
        # from the previous step: bit-exact only with bit-exact inputs (the
        # stateless smoothing tests), so here h and k carry tolerances.
        assert np.mean(st1["k"] != s1.k) < 1e-3
        assert rel_max(st1["h"], s1.h) < 1e-8
        assert rel_max(st1["rho"], s1.rho) < (1e-6 if mode == "mixed" else 1e-12)
        assert rel_rms(st1["acc"], s1.acc) < tol
        assert rel_rms(st1["acc_g"], s1.acc_g) < tol
        assert rel_rms(st1["dudt"], s1.dudt) < tol
        assert rel_max(st1["u"], s1.u) < 1e-9
        assert rel_rms(st1["pos"], s1.pos) < 1e-12
        pl.close()


def test_plan_bind_keeps_the_callers_buffers(gpu):
    """sphx_plan_bind: the caller's device tensors hold the state after every
    call, in the caller's order; the run is bit-identical to a loaded plan's;
    an edit of a bound buffer between calls is what the next step starts from."""
    torch = gpu
    from sph_2_amd.plan import SphPlan

    e = ic.evrard(6000, seed=8)
    bufs = {k: torch.tensor(getattr(e, k), dtype=torch.float64, device="cuda")
            for k in ("pos", "vel", "u", "m", "h")}
    pb = SphPlan(e.n)
    pb.bind(bufs["pos"], bufs["vel"], bufs["u"], bufs["m"], bufs["h"])
    pb.init()
    assert np.array_equal(bufs["h"].cpu().numpy(), pb.state(("h",))["h"])
    pb.step(2, 1e-4)
    pl = SphPlan(e.n)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(2, 1e-4)
    a, b = pb.state(), pl.state()
    for f in a:
        assert np.array_equal(a[f], b[f]), f
    for f in ("pos", "vel", "u", "h"):
        assert np.array_equal(bufs[f].cpu().numpy(), a[f]), f
    bufs["u"][:100] *= 2.0                       # edited between calls
    pb.step(1, 1e-4)
    pl.step(1, 1e-4)
    a2, b2 = pb.state(), pl.state()
    assert np.all(a2["u"][:100] > 1.5 * b2["u"][:100])
    assert np.array_equal(bufs["u"].cpu().numpy(), a2["u"])
    pb.close()
    pl.close()


def test_diagnostics(gpu, evr):
    from sph_2_amd import libsx

    d = libsx.diagnostics(evr["pos"], evr["v"], evr["m"], evr["u"], evr["h"], 1.0)
    ke = oracle.ke(evr["v"], evr["m"])
    pe = oracle.pe(evr["pos"], evr["m"], evr["h"], 1.0)
    assert abs(d["ke"] - ke) < 1e-13 * ke
    assert abs(d["pe"] - pe) < 1e-12 * abs(pe)
    assert abs(d["ut"] - float((evr["m"] * evr["u"]).sum())) < 1e-13
    assert abs(libsx.getKE(evr["v"], evr["m"]) - ke) < 1e-13 * ke
    assert abs(libsx.getPE(evr["pos"], evr["m"], 1.0, evr["h"]) - pe) < 1e-12 * abs(pe)


def test_cell_level_does_not_change_results(gpu):
    """The bucket cut level only changes which candidate buckets each segment
    walks; every kept pair is met in the same order, so a plan gives the same
    bits at any level."""
    from sph_2_amd.plan import SphPlan

    n = 32768
    e = ic.evrard(n, seed=4)
    outs = []
    for b in (-1, 0, 2):                      # cut levels 1, 2 and 4 at this size
        pl = SphPlan(n)
        pl.set_cell_level(b)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        pl.step(2, 1e-4)
        outs.append(pl.state())
        if b == 0:
            pl.set_cell_level(2)              # also when changed between steps
            pl.step(1, 1e-4)
            later = pl.state()
        pl.close()
    for o in outs[1:]:
        for f in o:
            assert np.array_equal(outs[0][f], o[f]), f
    pl = SphPlan(n)
    pl.set_cell_level(0)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    pl.step(3, 1e-4)
    ref = pl.state()
    pl.close()
    for f in ref:
        assert np.array_equal(ref[f], later[f]), f


@pytest.mark.parametrize("case", ["evrard", "cloud_balsara", "overflow"])
def test_neighbor_lists_do_not_change_results(gpu, case):
    """div/curl and hydro over the density pass's neighbour lists (one fused
    pass that also adds gravity, without the Balsara switch; two passes with
    it) give the same bits as the candidate walks: Evrard; a turbulent cloud with the Gadget
    viscosity and the Balsara switch (per-particle factors for the lists); and
    buckets whose lists overflow (h held at three times its converged value for
    the particles with x > 0.3 by an iteration cap of one count, so their
    neighbour counts pass 128): those buckets walk, the others iterate their
    lists, inside the same step."""
    from sph_2_amd.plan import SphPlan

    kw = {}
    if case == "evrard":
        e = ic.evrard(32768, seed=4)
    elif case == "cloud_balsara":
        e = ic.turbulent_cloud(16384, seed=2, ng=32)
        kw = dict(av_form=1, balsara=1)
    else:
        e = ic.evrard(16384, seed=6)
        e.h = np.where(e.pos[:, 0] > 0.3, 3.0 * e.h, e.h)
        kw = dict(max_iter=1)
    outs, cnts = [], []
    for lists in (True, False):
        pl = SphPlan(e.n, **kw)
        pl.set_neighbor_lists(lists)
        pl.load(e.pos, e.vel, e.u, e.m, e.h)
        pl.init()
        pl.step(3, 1e-4)
        outs.append(pl.state())
        cnts.append(pl.counters())
        pl.close()
    for f in outs[0]:
        assert np.array_equal(outs[0][f], outs[1][f]), f
    if case == "overflow":
        assert 0 < cnts[0]["nbr_overflow"] < cnts[0]["buckets"], cnts[0]
    else:
        assert cnts[0]["nbr_overflow"] == 0, cnts[0]


def test_symmetric_gravity_conserves_momentum_at_scale(gpu):
    """Size-independent check at N = 262,144 (beyond the oracle's reach): the
    symmetric sum applies each pair's force to both particles, so the total
    gravitational force sum_i m_i a_i vanishes to round-off, and the result
    agrees with the tiled sum on the same input."""
    from sph_2_amd import libsx

    n = 262144
    e = ic.evrard(n, seed=9)
    try:
        libsx.set_gravity_config(8, 512, 2)
        a_sym = libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.h)
        libsx.set_gravity_config(2, 512, 4)
        a_til = libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.h)
    finally:
        libsx.set_gravity_config()
    scale = float((e.m[:, None] * np.abs(a_sym)).sum())
    net = np.abs((e.m[:, None] * a_sym).sum(0)).max()
    assert net / scale < 1e-6
    assert rel_rms(a_sym, a_til) < 1e-6


def test_pair_issue_rate(gpu):
    """The far-pair issue-rate reference runs, returns a rate in a physical range
    for one MI355X (10^11..10^14 pair evaluations per second) and rejects
    iteration counts outside 1..65536."""
    from sph_2_amd import libsx
    from sph_2_amd._lib import SphxError

    r = libsx.pair_issue_rate(64)
    assert 1e11 < r < 1e14, r
    for bad in (0, 70000):
        with pytest.raises(SphxError, match="iters"):
            libsx.pair_issue_rate(bad)


def test_library_first_in_a_fresh_process(gpu):
    """A caller whose first step is loading the library (before anything has
    imported torch) still reaches the device: the process holds one HIP
    runtime.  Runs in a child process so that nothing is imported beforehand."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import ctypes\n"
            "from sph_2_amd import _lib\n"
            "L = _lib.load()\n"
            "c = ctypes.c_void_p()\n"
            "rc = L.sphx_ctx_create(0, ctypes.byref(c))\n"
            "assert rc == 0, L.sphx_last_error()\n"
            "assert L.sphx_ctx_destroy(c) == 0\n"
            "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=150)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
