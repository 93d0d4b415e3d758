"""Pins of the oracle's 5-D CHIMES table lookup (hcoolingx, hfvCLibs_v7.h:1077-1240).

The tables of chimes_synth encode every node in their values, so each lookup
has a closed-form answer that does not restate the search: the energy after
the lookup is u * c(k, n, z, j) / c(k, n, z, 0) for the nodes chosen by
construction of the query, the species fractions and mu are flat indices.
"""
import numpy as np

import oracle
from chimes_synth import coded_table, dt_for_time_node, node_code, queries


def _expect(tab, u, jk, jn, ju, below, jt):
    n_kpc, n_t, n_nh, n_z, n_m, n_time = tab.metalz.shape
    jz = tab.z_index
    c0 = node_code(tab, jk, jn, jz, 0)
    cj = node_code(tab, jk, jn, jz, jt)
    u_cgs = u * tab.unit_u_cgs
    want = u_cgs * cj / c0
    col0 = tab.u_evol[jk, 0, jn, jz, 0].astype(np.float64)
    want = np.where(below, col0 * cj / c0, want)
    want = np.minimum(want, tab.u_max_cgs)
    ion = np.ravel_multi_index((jk[:, None], ju[:, None], jn[:, None], jz, np.arange(n_m)[None, :],
                                jt), tab.metalz.shape).astype(np.float64)
    mu = np.ravel_multi_index((jk, ju, jn, jz, jt), tab.u_evol.shape) + 0.5
    return want / tab.unit_u_cgs, ion, mu


def test_chimes_lookup_closed_form():
    tab = coded_table(oracle)
    for jt in (1, 3, 5):
        pos, rho, u, jk, jn, ju, below = queries(tab, 4000, seed=jt)
        dt = dt_for_time_node(tab, jt, 0.37)
        got, ion, mu = oracle.chimes_apply(pos, rho, u, dt, tab)
        want, ion_w, mu_w = _expect(tab, u, jk, jn, ju, below, jt)
        assert np.max(np.abs(got / want - 1)) < 1e-13
        assert np.array_equal(ion.astype(np.float64), ion_w)
        assert np.array_equal(mu, mu_w)


def test_chimes_time_node_edges():
    """dt exactly on a node picks that node; dt past the last node picks the last."""
    tab = coded_table(oracle)
    pos, rho, u, jk, jn, ju, below = queries(tab, 500, seed=9)
    ts = tab.time_s.astype(np.float64)
    for dt_s, jt in ((ts[2], 2), (ts[-1] * 4.0, len(ts) - 1), (0.0, 0)):
        got, ion, mu = oracle.chimes_apply(pos, rho, u, dt_s / tab.unit_time_s, tab)
        want, _, mu_w = _expect(tab, u, jk, jn, ju, below, jt)
        if dt_s == ts[2]:
            # the division by unit_time_s may round dt_s just above the node
            dt_back = (dt_s / tab.unit_time_s) * tab.unit_time_s
            if dt_back > ts[2]:
                continue
        assert np.max(np.abs(got / want - 1)) < 1e-13
        assert np.array_equal(mu, mu_w)


def test_chimes_nearest_node_tie_goes_up():
    """A radius exactly between two kpc nodes takes the upper node (the reference
    moves down only when the lower node is strictly closer); beyond the last
    node the last, below the first the first."""
    tab = coded_table(oracle, unit_length_cm=1.0)
    tab.kpc_cm = 1.0
    kpc = tab.kpc.astype(np.float64)
    r = np.array([0.5 * (kpc[1] + kpc[2]), kpc[-1] * 3.0, kpc[0] * 0.5, kpc[3]])
    want_k = np.array([2, len(kpc) - 1, 0, 3])
    pos = np.zeros((4, 3))
    pos[:, 0] = r
    nh0 = float(tab.nh[7])
    rho = np.full(4, nh0 * tab.m_h / (tab.xh * tab.unit_rho_cgs))
    u = np.full(4, 0.5 * float(tab.u_evol[0, 0, 7, 1, 0]) / tab.unit_u_cgs)   # below every column
    got, ion, mu = oracle.chimes_apply(pos, rho, u, dt_for_time_node(tab, 1), tab)
    jn = np.array([int(np.argmin(np.abs(tab.nh.astype(np.float64) - nh0)))] * 4)
    assert jn[0] == 7
    mu_w = np.ravel_multi_index((want_k, np.zeros(4, int), jn, 1, np.ones(4, int)),
                                tab.u_evol.shape) + 0.5
    assert np.array_equal(mu, mu_w)


def test_chimes_cap():
    """Energies above u_max_cgs (T = 1e10 K in the reference) are capped."""
    tab = coded_table(oracle, u_max_cgs=3.0e8)
    pos, rho, u, jk, jn, ju, below = queries(tab, 300, seed=4)
    got, _, _ = oracle.chimes_apply(pos, rho, u, dt_for_time_node(tab, 2), tab)
    cap = 3.0e8 / tab.unit_u_cgs
    assert np.all(got <= cap)
    assert 10 < np.sum(got == cap) < len(got)
