"""C-ABI boundary and host-logic checks that need no GPU."""
import os
import re
import subprocess

import numpy as np

from sph_2_amd import _lib, ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    hdr = open(os.path.join(ROOT, "include", "sphx.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return set(re.findall(r"\b(sphx_[A-Za-z0-9_]+)\s*\(", hdr))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    assert lib.sphx_version().decode().startswith("sphx")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = header_functions()
    assert declared, "no declarations parsed"
    assert declared <= exported, declared - exported
    assert {s for s in exported if s.startswith("sphx_")} <= declared


def test_ctypes_table_matches_header():
    assert set(_lib.SIGNATURES) == header_functions()


def test_default_params_and_errors_without_gpu():
    p = _lib.default_params()
    assert (p.nngb, p.ndown, p.nup, p.max_iter) == (64, 59, 69, 100)
    assert abs(p.gamma - 5 / 3) < 1e-15 and p.alpha == 1.0 and p.beta == 2.0 and p.eta == 0.1
    lib = _lib.load()
    # argument errors are reported, not crashes (no device needed to reject them)
    rc = lib.sphx_plan_create(0, 7, None, None, 10, None)
    assert rc == -1
    assert "output" in lib.sphx_last_error().decode() or "mode" in lib.sphx_last_error().decode()


def test_null_context_and_bad_rows_are_rejected_without_gpu():
    """Every stateless entry point checks its context first, and the row
    range before any device work: both come back as SPHX_ERR_ARG with a
    message, on a machine with no GPU at all."""
    lib = _lib.load()
    assert lib.sphx_getDensity(None, 0, 10, None, None, None, 0, 10, None, None, None) == -1
    assert "context" in lib.sphx_last_error().decode()
    assert lib.sphx_getAcc_g_smth(None, 0, 10, None, None, None, 1.0, 0, 10, None, None,
                                  None) == -1
    assert lib.sphx_pair_issue_rate(None, 16, None, None) == -1
    assert "context" in lib.sphx_last_error().decode()


def test_wrappers_refuse_arrays_of_the_wrong_shape():
    """The kernels read n rows of every array; the Python layer refuses an
    array of another length before any of it reaches the device (where it
    would be read past its end), as numpy indexing would have in the
    reference's functions."""
    import pytest

    from sph_2_amd import libsx

    e = ic.evrard(64, seed=0)
    g = 5.0 / 3.0
    bad = [
        lambda: libsx.getDensity(e.pos, e.m[:-1], e.h),
        lambda: libsx.getDensity(e.pos[:, :2], e.m, e.h),
        lambda: libsx.getAcc_sph(e.pos, e.vel[:-1], e.h, e.h, e.h, e.h, e.m, g, 0.1, 1.0, 2.0),
        lambda: libsx.get_dU(e.pos, e.vel, e.h, e.h, e.h, e.h, e.m[:9], g, 0.1, 1.0, 2.0),
        lambda: libsx.getAcc_sph_shear(e.pos, e.vel, e.h, e.h, e.h, e.h, e.m, e.h[:3], e.h, 1.0),
        lambda: libsx.div_curlVel(e.pos, e.vel, e.h[:10], e.m, e.h),
        lambda: libsx.getAcc_g_smth(e.pos, e.m[:3], 1.0, e.h),
        lambda: libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.h[:3]),
        lambda: libsx.getAcc_g_bh(e.pos, e.m[:3], 1.0, e.h),
        lambda: libsx.getKE(e.vel, e.m[:5]),
        lambda: libsx.getPE(e.pos[:7], e.m, 1.0, e.h),
        lambda: libsx.global_dt(e.h, e.h, e.vel, e.vel[:4], e.h),
        lambda: libsx.smoothing_h(e.pos, e.h[:5]),
        lambda: libsx.smoothing_fast_new_mpi3(0, 64, e.pos, e.h, np.zeros(3), e.h, 1e-3),
        lambda: libsx.neighbors(e.pos, e.h[:5]),
        lambda: libsx.getPressure_csound(e.h, e.h[:3], g),
        lambda: libsx.getDensity_mpi(10, 70, e.pos, e.m, e.h),
    ]
    for k, f in enumerate(bad):
        with pytest.raises(ValueError):
            f()
    with pytest.raises(ValueError, match="vel"):
        _lib.check_run_arrays(e.pos, e.vel[:-1], e.u, e.m, e.h)
    with pytest.raises(ValueError, match="eps"):
        _lib.check_run_arrays(e.pos, e.vel, e.u, e.m, e.h, np.ones(3))
    assert _lib.check_run_arrays(e.pos, e.vel, e.u, e.m, e.h, np.ones(64)) == 64


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "sph_2_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), f
                assert "liboracle" not in src, f


def test_splitmix_is_deterministic_and_uniform():
    a = ic.uniform(0, 0, 5)
    b = ic.uniform(0, 0, 5)
    assert np.array_equal(a, b)
    assert np.array_equal(ic.splitmix64(0, 0, 3, offset=2), ic.splitmix64(0, 0, 5)[2:])
    u = ic.uniform(123, 4, 200000)
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002


def test_evrard_ic_profile():
    """M(<r) = r^2 for rho ~ 1/r with R = M = 1."""
    e = ic.evrard(100000, seed=0)
    r = np.sort(np.linalg.norm(e.pos, axis=1))
    frac = np.arange(1, len(r) + 1) / len(r)
    for q in (0.1, 0.3, 0.5, 0.9):
        k = int(q * len(r))
        assert abs(r[k] ** 2 - frac[k]) < 0.01
    assert abs(e.m.sum() - 1.0) < 1e-12


def test_turbulent_cloud_mach():
    c = ic.turbulent_cloud(20000, seed=0)
    units = c.units
    cs = np.sqrt(ic.KB * 1e4 / (1.22 * ic.MH)) / units["unit_velocity_cms"]
    assert abs(np.sqrt((c.vel ** 2).sum(1).mean()) / cs - 2.0) < 1e-9
    assert np.abs(c.vel.mean(0)).max() < 1e-12


def test_autotune_pick_prefers_least_partial_memory():
    """Among shapes within 1 % of the fastest the tuner keeps the one with the
    least extra memory: a tiled shape, else the most block pairs per wave,
    whole block pairs before split ones; a clearly faster shape still wins."""
    from sph_2_amd.autotune import pick

    assert pick({(8, 512, 1): 1.0, (8, 256, 1): 0.995}) == (8, 512, 1)
    assert pick({(8, 512, 1): 1.0, (8, 256, 1): 0.95}) == (8, 256, 1)
    assert pick({(8, 512, 1): 1.0, (8, 512, 4): 1.005, (8, 128, 1): 0.999}) == (8, 512, 4)
    assert pick({(8, 512, 16): 1.0, (2, 512, 1): 1.008}) == (2, 512, 1)
