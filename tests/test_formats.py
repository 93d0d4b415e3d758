"""The reference's file formats through the library's C++ I/O (host only, no GPU).

Layouts are pinned against the reference's own writer recipes restated with
numpy/pandas (IC/3.py:51-73 pandas CSV and positional params file;
IC_evrard_binary.py:60-82 numpy tobytes; hfvCppLibs.h:68-96 snapshot), and
against the params.GPU files the reference ships (tests/golden/)."""
import os

import numpy as np
import pandas as pd
import pytest

from sph_2_amd import _lib, formats, ic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture
def evr():
    e = ic.evrard(500, seed=3)
    rng = np.random.default_rng(3)
    vel = rng.normal(size=(500, 3)) * 0.1
    return dict(pos=e.pos, vel=vel, m=e.m, h=e.h, eps=np.full(500, 0.05), u=e.u)


def test_params_gpu_reference_files():
    p = formats.read_params_gpu(os.path.join(GOLDEN, "params_GPU_evrard.txt"))
    assert p["N"] == 212472 and p["gamma"] == 1.6666666666666667 and p["Mach"] == 0.0
    assert p["muu" if "muu" in p else "mu"] == 3.0
    assert p["grav_const_in_cgs"] == 6.67259e-08 and p["G"] == 1.0
    q = formats.read_params_gpu(os.path.join(GOLDEN, "params_GPU_restart_cloud.txt"))
    assert q["N"] == 50000 and q["c_0"] == 822571.5807307774 and q["Rcld_in_cm"] == 6.172e19
    assert q["Mcld_in_g"] == 1.0837144556928e38 and q["mu"] == 1.22 and q["Mach"] == 20.0


def test_params_gpu_round_trip_is_positional(tmp_path):
    src = formats.read_params_gpu(os.path.join(GOLDEN, "params_GPU_restart_cloud.txt"))
    f = tmp_path / "params.GPU"
    formats.write_params_gpu(f, **src)
    assert formats.read_params_gpu(f) == src
    # one value per line, in the driver's read order (hfv_sph_gpu_v2_evrard.cu:47-56)
    vals = [float(v) for v in open(f).read().split()]
    order = ["N", "c_0", "gamma", "Rcld_in_pc", "Rcld_in_cm", "Mcld_in_g", "mu", "Mach",
             "grav_const_in_cgs", "G"]
    assert vals == [float(src[k]) for k in order]


def test_ic_csv_reads_the_pandas_layout(tmp_path, evr):
    """IC/3.py:51-60 writes the dict x,y,z,vx,vy,vz,m,h,eps,u with pandas,
    no index, no header."""
    d = {"x": evr["pos"][:, 0], "y": evr["pos"][:, 1], "z": evr["pos"][:, 2],
         "vx": evr["vel"][:, 0], "vy": evr["vel"][:, 1], "vz": evr["vel"][:, 2],
         "m": evr["m"], "h": evr["h"], "eps": evr["eps"], "u": evr["u"]}
    f = tmp_path / "GPU_IC.csv"
    pd.DataFrame(d).to_csv(f, index=False, header=False)
    got = formats.read_ic_csv(f)
    for k in ("m", "h", "eps", "u"):
        assert np.array_equal(got[k], evr[k]), k
    assert np.array_equal(got["pos"], evr["pos"]) and np.array_equal(got["vel"], evr["vel"])


def test_ic_and_restart_csv_round_trip_exactly(tmp_path, evr):
    f = tmp_path / "ic.csv"
    formats.write_ic_csv(f, **evr)
    back = formats.read_ic_csv(f, N=500)
    for k in evr:
        assert np.array_equal(back[k], evr[k]), k
    dudt = np.random.default_rng(1).normal(size=500)
    r = tmp_path / "RESTART_0.015751.csv"
    formats.write_restart_csv(r, dudt=dudt, **evr)
    assert np.loadtxt(r, delimiter=",").shape == (500, 11)
    back = formats.read_ic_csv(r, restart=True)
    assert np.array_equal(back["dudt"], dudt) and np.array_equal(back["pos"], evr["pos"])
    with pytest.raises(_lib.SphxError, match="11th column"):
        formats.read_ic_csv(f, restart=True)


def test_snapshot_csv_header_and_columns(tmp_path, evr):
    f = tmp_path / "G-0.01.csv"
    rho = np.random.default_rng(2).random(500)
    formats.write_snapshot_csv(f, evr["pos"], evr["vel"], evr["h"], rho, evr["u"])
    df = pd.read_csv(f, float_precision="round_trip")
    assert list(df.columns) == ["x", "y", "z", "vx", "vy", "vz", "h", "rho", "u"]
    assert np.array_equal(df["rho"].to_numpy(), rho)
    assert np.array_equal(df[["x", "y", "z"]].to_numpy(), evr["pos"])


def test_binary_ic_bytes_match_the_reference_recipe(tmp_path, evr):
    """IC_evrard_binary.py:60-82: Typ int32, then float32 x,y,z,vx,vy,vz,m,h,eps,u."""
    f = tmp_path / "Evrard_GPU_IC.bin"
    typ = np.arange(500, dtype=np.int32) % 3
    formats.write_ic_binary(f, typ=typ, **evr)
    want = typ.tobytes()
    cols = [evr["pos"][:, 0], evr["pos"][:, 1], evr["pos"][:, 2], evr["vel"][:, 0],
            evr["vel"][:, 1], evr["vel"][:, 2], evr["m"], evr["h"], evr["eps"], evr["u"]]
    for c in cols:
        want += c.astype(np.float32).tobytes()
    assert open(f, "rb").read() == want
    back = formats.read_ic_binary(f, 500)
    assert np.array_equal(back["typ"], typ)
    assert np.array_equal(back["pos"], evr["pos"].astype(np.float32).astype(np.float64))
    assert np.array_equal(back["u"], evr["u"].astype(np.float32).astype(np.float64))
    with pytest.raises(_lib.SphxError, match="44 bytes"):
        formats.read_ic_binary(f, 499)


def test_binary_snapshot_layout(tmp_path, evr):
    """hfvCppLibs.h:68-96: int32 N, Typ, float32 x,y,z,vx,vy,vz,rho,h,u,mass."""
    f = tmp_path / "snap.bin"
    rho = np.random.default_rng(4).random(500)
    formats.write_snapshot_binary(f, evr["pos"], evr["vel"], rho, evr["h"], evr["u"], evr["m"])
    raw = open(f, "rb").read()
    assert np.frombuffer(raw[:4], np.int32)[0] == 500
    assert np.array_equal(np.frombuffer(raw[4:2004], np.int32), np.zeros(500, np.int32))
    body = np.frombuffer(raw[2004:], np.float32).reshape(10, 500)
    assert np.array_equal(body[6], rho.astype(np.float32))
    assert np.array_equal(body[9], evr["m"].astype(np.float32))
    back = formats.read_snapshot_binary(f)
    assert np.array_equal(back["rho"], rho.astype(np.float32).astype(np.float64))
    assert np.array_equal(back["vel"], evr["vel"].astype(np.float32).astype(np.float64))


def test_errors_name_the_file(tmp_path):
    with pytest.raises(_lib.SphxError, match="cannot open"):
        formats.read_ic_csv(tmp_path / "missing.csv")
    bad = tmp_path / "bad.csv"
    bad.write_text("1,2,3\n")
    with pytest.raises(_lib.SphxError, match="row 1"):
        formats.read_ic_csv(bad)
    short = tmp_path / "short.txt"
    short.write_text("5\n1.0\n")
    with pytest.raises(_lib.SphxError, match="fewer than 10"):
        formats.read_params_gpu(short)


def _chimes_arrays(rng, nk=3, nt=7, nn=5, nz=3, nm=2, ntime=4):
    return dict(kpc=np.sort(rng.uniform(0, 5, nk)).astype(np.float32),
                nh=np.sort(rng.uniform(0, 1e3, nn)).astype(np.float32),
                z=np.array([-2, -1, 0], np.float32)[:nz],
                temp=np.sort(rng.uniform(1, 1e8, nt)).astype(np.float32),
                time_s=np.sort(rng.uniform(0, 1e10, ntime)).astype(np.float32),
                u_evol=rng.random((nk, nt, nn, nz, ntime), np.float32),
                mu=rng.random((nk, nt, nn, nz, ntime), np.float32),
                metalz=rng.random((nk, nt, nn, nz, nm, ntime), np.float32))


def test_chimes_table_bytes_match_the_reference_recipe(tmp_path):
    """coolHeatGridZ.bin as step_2x.py:101-140 writes it: struct.pack('i') of
    N_kpc, N_nH, N_Z, N_T, N_M, N_time, then 'f' values of the axes and of the
    three arrays in nested-loop (C) order."""
    import struct

    a = _chimes_arrays(np.random.default_rng(5))
    nk, nt, nn, nz, nm, ntime = a["metalz"].shape
    want = b"".join(struct.pack("i", v) for v in (nk, nn, nz, nt, nm, ntime))
    for k in ("kpc", "nh", "z", "temp", "time_s", "u_evol", "mu", "metalz"):
        want += b"".join(struct.pack("f", v) for v in a[k].ravel())
    f = tmp_path / "coolHeatGridZ.bin"
    formats.write_chimes_table(f, **a)
    assert f.read_bytes() == want
    back = formats.read_chimes_table(f)
    for k, v in a.items():
        assert back[k].shape == v.shape and np.array_equal(back[k], v), k


def test_chimes_table_errors(tmp_path):
    a = _chimes_arrays(np.random.default_rng(6))
    f = tmp_path / "t.bin"
    formats.write_chimes_table(f, **a)
    data = f.read_bytes()
    (tmp_path / "short.bin").write_bytes(data[:-4])
    with pytest.raises(_lib.SphxError, match="size does not match"):
        formats.read_chimes_table(tmp_path / "short.bin")
    with pytest.raises(_lib.SphxError, match="cannot open"):
        formats.read_chimes_table(tmp_path / "missing.bin")
    with pytest.raises(ValueError):
        formats.write_chimes_table(f, **{**a, "mu": a["mu"][:, :-1]})
