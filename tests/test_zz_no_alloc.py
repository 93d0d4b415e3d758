"""The allocation-free global time step (sphx_global_dt_noalloc) through the C
ABI against sphx_global_dt: the same five values (dt, dt_f, dt_kin, dt_cfl,
dt_dens) bit for bit on two calls in a row, the second with fewer particles
and milder extrema than the first, so that reading anything the first call
left in the shared scratch would change its result; sizes past the fixed
scratch and unknown devices are refused."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _inputs(torch, n, seed, scale):
    rng = np.random.default_rng(seed)
    h = 0.01 + rng.random(n)
    c = 0.1 + rng.random(n)
    acc = scale * rng.standard_normal((n, 3))
    acc_g = scale * rng.standard_normal((n, 3))
    dv = rng.standard_normal(n)
    return [torch.tensor(a, dtype=torch.float64, device="cuda") for a in (h, c, acc, acc_g, dv)]


def test_noalloc_global_dt_matches_the_context_call(gpu):
    torch = gpu
    from sph_2_amd import _lib, libsx

    L = _lib.load()
    ctx = libsx._context(0)
    st = _lib.stream_handle()
    for n, seed, scale in ((300_001, 1, 100.0), (4099, 2, 1.0)):
        bufs = _inputs(torch, n, seed, scale)
        ptrs = [b.data_ptr() for b in bufs]
        ref = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
        got = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
        print(f"global dt: n={n} seed={seed}", flush=True)
        assert L.sphx_global_dt(ctx, n, *ptrs, 0.25, 1e-9, 1.0, ref.data_ptr(), st) == 0, \
            L.sphx_last_error()
        assert L.sphx_global_dt_noalloc(0, n, *ptrs, 0.25, 1e-9, 1.0, got.data_ptr(), st) == 0, \
            L.sphx_last_error()
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(r)) and np.array_equal(r, g), (r, g)
        # and the FP64 oracle's criterion on the same inputs
        dt, parts = oracle.global_dt(*[b.cpu().numpy() for b in bufs], 0.25, 1e-9, 1.0)
        assert dt == g[0] and parts == tuple(g[1:])


def test_noalloc_global_dt_refuses_bad_sizes_and_devices(gpu):
    torch = gpu
    from sph_2_amd import _lib

    L = _lib.load()
    x = torch.zeros(16, dtype=torch.float64, device="cuda").data_ptr()
    st = _lib.stream_handle()
    assert L.sphx_global_dt_noalloc(0, (1 << 23) + 1, x, x, x, x, x, 0.25, 1e-9, 1.0, x, st) == -1
    assert b"8388608" in L.sphx_last_error()
    assert L.sphx_global_dt_noalloc(0, 0, x, x, x, x, x, 0.25, 1e-9, 1.0, x, st) == -1
    assert L.sphx_global_dt_noalloc(4096, 16, x, x, x, x, x, 0.25, 1e-9, 1.0, x, st) == -1
    assert b"device" in L.sphx_last_error()
    assert L.sphx_global_dt_noalloc(0, 16, x, x, x, x, x, 0.0, 1e-9, 1.0, x, st) == -1
