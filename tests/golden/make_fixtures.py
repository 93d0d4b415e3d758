"""Extract the reference's own data files into small npz fixtures.

Run in a checkout that has the reference mounted (it is not needed later):
    python tests/golden/make_fixtures.py /root/reference

The reference stores its Evrard initial conditions, its energy history and
Gadget-4 comparison profiles as pickles of dicts of numpy arrays.  They are
NOT unpickled here: pickletools.genops only disassembles the opcode stream,
and the raw array payloads (BINBYTES) are read with numpy.frombuffer, so
nothing in the files is executed.

Outputs (tests/golden/):
  evrard_1472_ic.npz      x, y, z float32 (Evrard_collapse/Evrard_1472.pkl,
                          loaded by Evrard_collapse/augsphx6.5.py:406-417)
  evrard_1472_energy.npz  U, PE, KE, t float64 (Evrard_collapse/Energy.pkl,
                          written by augsphx6.5.py:569-571: U = sum m u,
                          PE = getPE(r, m, G, 0.005), KE = getKE(v, m) with
                          N = 1472, m = 1/N, u = 0.05, dt = 1e-3)
  gadget4_profiles.npz    r, rho / P / vr arrays of Evrard_collapse/
                          {rho,P,vr}Gadget.pkl (Gadget-4 run, t ~ 0.8)
  params_GPU_evrard.txt   the ten numbers of GPU_sph/Important_Test_Simulations/
                          Evrard_collapse/IC/params.GPU (a data file, copied)
  params_GPU_restart_cloud.txt
                          the same for GPU_sph/With_metal_cooling/
                          Check_Restart_version/params.GPU
"""
import os
import pickletools
import shutil
import sys

import numpy as np

SKIP = {"numpy.core.multiarray", "_reconstruct", "numpy", "ndarray", "dtype", "<", ">", "|",
        "f4", "f8", "i4", "i8", "b"}


def arrays(path):
    """{key: ndarray} of a pickled dict of 1-D numpy arrays, without unpickling."""
    data = open(path, "rb").read()
    out = {}
    key = None
    dtype = None
    for op, arg, _ in pickletools.genops(data):
        if op.name in ("SHORT_BINUNICODE", "BINUNICODE", "UNICODE"):
            if arg in ("f4", "f8", "i4", "i8"):
                dtype = arg
            elif arg not in SKIP:
                key = arg
        elif op.name in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8") and len(arg) > 8:
            if key is None or dtype is None:
                raise ValueError(f"{path}: payload without key/dtype")
            # later arrays reuse the memoised dtype object, so the last
            # dtype string seen stays in force
            out[key] = np.frombuffer(arg, dtype="<" + dtype).copy()
    return out


def main(ref_root):
    base = os.path.join(ref_root, "11_Dec_2022", "Evrard_collapse")
    here = os.path.dirname(os.path.abspath(__file__))
    ic = arrays(os.path.join(base, "Evrard_1472.pkl"))
    np.savez(os.path.join(here, "evrard_1472_ic.npz"), **ic)
    en = arrays(os.path.join(base, "Energy.pkl"))
    np.savez(os.path.join(here, "evrard_1472_energy.npz"), **en)
    prof = {}
    for name in ("rho", "P", "vr"):
        a = arrays(os.path.join(base, f"{name}Gadget.pkl"))
        for k, v in a.items():
            prof[f"{name}_{k}"] = v
    np.savez(os.path.join(here, "gadget4_profiles.npz"), **prof)
    for f, d in (("ic", ic), ("energy", en), ("profiles", prof)):
        print(f, {k: (v.dtype.str, v.shape) for k, v in d.items()})
    gpu = os.path.join(ref_root, "11_Dec_2022", "GPU_sph")
    for src, dst in (("Important_Test_Simulations/Evrard_collapse/IC/params.GPU",
                      "params_GPU_evrard.txt"),
                     ("With_metal_cooling/Check_Restart_version/params.GPU",
                      "params_GPU_restart_cloud.txt")):
        shutil.copyfile(os.path.join(gpu, src), os.path.join(here, dst))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
