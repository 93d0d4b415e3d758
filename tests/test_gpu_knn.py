"""Exact kNN smoothing lengths on the GPU (do_smoothingX, LIBSX:762-783) against
the oracle's brute-force selection: bit-equal, including ties and duplicates."""
import numpy as np
import pytest

import oracle
from sph_2_amd import ic

pytestmark = pytest.mark.gpu


def test_evrard_and_uniform_bit_exact(gpu):
    from sph_2_amd import libsx

    e = ic.evrard(4096, seed=0)
    assert np.array_equal(libsx.do_smoothingX((e.pos, e.pos)), oracle.knn_h(e.pos))
    rng = np.random.default_rng(3)
    pos = rng.random((20000, 3))
    assert np.array_equal(libsx.do_smoothingX((pos, pos)), oracle.knn_h(pos))


def test_row_block_and_other_rank(gpu):
    from sph_2_amd import libsx

    e = ic.plummer(6000, seed=5)
    pos = e.pos
    got = libsx.do_smoothingX((pos, pos[1000:1700]))
    assert np.array_equal(got, oracle.knn_h(pos, rows=np.arange(1000, 1700)))
    assert np.array_equal(libsx.do_smoothingX((pos, pos), kth=20), oracle.knn_h(pos, kth=20))
    with pytest.raises(ValueError):
        libsx.do_smoothingX((pos, pos[::2].copy()))


def test_smoothing_length_mpi_rows(gpu):
    """smoothing_length_mpi / _mpix (LIBSX:1070-1126): rows [nbeg, nend) of
    the exact kNN h, bit-exact against the oracle; an empty range is empty."""
    from sph_2_amd import libsx

    e = ic.plummer(6000, seed=6)
    want = oracle.knn_h(e.pos, rows=np.arange(2500, 3100))
    assert np.array_equal(libsx.smoothing_length_mpi(2500, 3100, e.pos), want)
    assert np.array_equal(libsx.smoothing_length_mpix(2500, 3100, e.pos, e.h), want)
    assert libsx.smoothing_length_mpi(7, 7, e.pos).shape == (0,)
    with pytest.raises(ValueError):
        libsx.smoothing_length_mpi(10, 6001, e.pos)


def test_lattice_ties_and_duplicates(gpu):
    """Shell degeneracy on a lattice, and 300 coincident particles whose 65th
    distance is 0 (the bracket must be bisected down to the tie)."""
    from sph_2_amd import libsx

    g = np.arange(16) * 0.1
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    assert np.array_equal(libsx.do_smoothingX((lat, lat)), oracle.knn_h(lat))
    rng = np.random.default_rng(8)
    pos = rng.random((5000, 3))
    pos[:300] = pos[7]
    got = libsx.do_smoothingX((pos, pos))
    assert np.array_equal(got, oracle.knn_h(pos))
    assert np.all(got[:300] == 0.0)


def test_planar_and_collinear_input(gpu):
    """Flat (z = 0) and collinear sets: the search starts from the area or
    length density, not a vanishing volume, and still selects the exact rank."""
    from sph_2_amd import libsx

    rng = np.random.default_rng(11)
    flat = rng.random((6000, 3))
    flat[:, 2] = 0.0
    got = libsx.do_smoothingX((flat, flat))
    assert np.all(np.isfinite(got)) and np.all(got > 0)
    assert np.array_equal(got, oracle.knn_h(flat))
    line = np.zeros((3000, 3))
    line[:, 0] = rng.random(3000)
    got = libsx.do_smoothingX((line, line))
    assert np.array_equal(got, oracle.knn_h(line))
