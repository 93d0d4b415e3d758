"""Synthetic 5-D CHIMES tables with closed-form contents (test data).

The real table (coolHeatGridZ.bin, written by the reference's
GPU_sph_Type_cooling_II/cooling_grids_use_EC2/step_2x.py from CHIMES runs) is
not part of the reference repository.  These tables have the same layout and
values chosen so that every lookup has a known answer:

* node energies U[k, t, n, z] = (1024 + 37 t + 3 k + n) * 2^20 (increasing in
  the T index, distinct per column), stored at time index 0 multiplied by the
  node code c(k, n, z, 0) / 4096;
* u_evol[k, t, n, z, j] = U[k, t, n, z] * c(k, n, z, j) / 4096 with a code c
  unique to each (kpc, nH, Z, time) node, so the interpolated energy of a
  particle with u between two T nodes is exactly u * c(j) / c(0): the node the
  lookup picked can be read off the result;
* mu and the species fractions hold their own flat index (+0.5 for mu).
Every product is exact in float32 (at most 12 + 12 significant bits).
"""
import numpy as np


def coded_table(oracle_mod, n_kpc=5, n_t=40, n_nh=25, n_z=3, n_m=4, n_time=6,
                unit_length_cm=3.086e21, unit_rho_cgs=6.77e-23, unit_u_cgs=1.0e12,
                unit_time_s=3.0e13, u_max_cgs=2.0287e18, z_index=1):
    kpc = np.array([0.1, 0.3, 0.6, 1.0, 2.0, 4.0, 8.0], np.float32)[:n_kpc]
    nh = (10.0 ** np.linspace(-2.0, 4.0, n_nh)).astype(np.float32)
    z = np.linspace(-2.0, 0.0, n_z).astype(np.float32)
    temp = (10.0 ** np.linspace(1.0, 8.0, n_t)).astype(np.float32)
    time_s = np.array([0.0, 1e7, 3e7, 1e8, 3e8, 1e9, 3e9, 1e10], np.float32)[:n_time]
    K, T, N, Z, J = np.meshgrid(np.arange(n_kpc), np.arange(n_t), np.arange(n_nh),
                                np.arange(n_z), np.arange(n_time), indexing="ij")
    U = (1024.0 + 37.0 * T + 3.0 * K + N) * 2.0 ** 20
    code = 1.0 + J + n_time * (N + n_nh * (K + n_kpc * Z))
    assert code.max() < 4096
    u_evol = (U * code / 4096.0).astype(np.float32)
    assert np.array_equal(u_evol.astype(np.float64), U * code / 4096.0)
    mu = (np.arange(u_evol.size, dtype=np.float64).reshape(u_evol.shape) + 0.5).astype(np.float32)
    metalz = np.arange(n_kpc * n_t * n_nh * n_z * n_m * n_time, dtype=np.float64)
    assert metalz.size < 2 ** 24
    metalz = metalz.reshape(n_kpc, n_t, n_nh, n_z, n_m, n_time).astype(np.float32)
    tab = oracle_mod.ChimesTable(kpc=kpc, nh=nh, z=z, temp=temp, time_s=time_s, u_evol=u_evol,
                                 mu=mu, metalz=metalz, unit_length_cm=unit_length_cm,
                                 unit_rho_cgs=unit_rho_cgs, unit_u_cgs=unit_u_cgs,
                                 unit_time_s=unit_time_s, z_index=z_index, u_max_cgs=u_max_cgs)
    return tab


def node_code(tab, k, n, z, j):
    n_kpc, n_t, n_nh, n_z, n_m, n_time = tab.metalz.shape
    return 1.0 + j + n_time * (n + n_nh * (k + n_kpc * z))


def queries(tab, m, seed):
    """m particles whose nearest kpc and nH nodes, T bracket and time node are
    known by construction.  Returns (pos, rho, u, dt, expected dict)."""
    rng = np.random.default_rng(seed)
    n_kpc, n_t, n_nh, n_z, n_m, n_time = tab.metalz.shape
    kpc = tab.kpc.astype(np.float64)
    nh = tab.nh.astype(np.float64)
    ts = tab.time_s.astype(np.float64)

    def near(grid, idx):
        f = rng.uniform(-0.45, 0.45, len(idx))
        lo = np.where(idx > 0, grid[np.maximum(idx - 1, 0)], grid[0] - (grid[1] - grid[0]))
        hi = np.where(idx < len(grid) - 1, grid[np.minimum(idx + 1, len(grid) - 1)],
                      grid[-1] + (grid[-1] - grid[-2]))
        gap = np.where(f < 0, grid[idx] - lo, hi - grid[idx])
        return grid[idx] + f * gap

    jk = rng.integers(0, n_kpc, m)
    jn = rng.integers(0, n_nh, m)
    r_kpc = np.maximum(near(kpc, jk), 1e-3)
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = d * (r_kpc * tab.kpc_cm / tab.unit_length_cm)[:, None]
    nh_q = np.maximum(near(nh, jn), 1e-6)
    rho = nh_q * tab.m_h / (tab.xh * tab.unit_rho_cgs)
    jz = tab.z_index
    col = tab.u_evol[jk, :, jn, jz, 0].astype(np.float64)          # (m, n_t) node energies
    # T bracket: mostly inside, a few below the first and above the last node
    ju = rng.integers(1, n_t, m)
    f = rng.uniform(0.05, 0.95, m)
    u_cgs = col[np.arange(m), ju - 1] + f * (col[np.arange(m), ju] - col[np.arange(m), ju - 1])
    below = rng.random(m) < 0.05
    above = (rng.random(m) < 0.05) & ~below
    u_cgs[below] = col[below, 0] * 0.5
    ju[below] = 0
    u_cgs[above] = col[above, -1] * 1.3
    ju[above] = n_t - 1
    u = u_cgs / tab.unit_u_cgs
    return pos, rho, u, jk, jn, ju, below


def dt_for_time_node(tab, j, frac=0.5):
    """A dt whose first time node >= dt (seconds) is j (j > 0)."""
    ts = tab.time_s.astype(np.float64)
    return (ts[j - 1] + frac * (ts[j] - ts[j - 1])) / tab.unit_time_s
