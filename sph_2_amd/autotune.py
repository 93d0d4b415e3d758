"""Launch-shape selection for the gravity kernel.

choose_gravity_config(n) returns (ipl, tj, js): particles per lane, j-tile
size and j split of the tiled sum, or (8, 512, block pairs per wave) for the
symmetric sum, or (8, tj, 1) for the symmetric sum with every block pair split
into parts of tj j particles (tj = 256 or 128: more, shorter waves).  The defaults below keep >= 4 workgroups per CU on the
256-CU MI355X; ``tune_gravity`` measures the candidates on the device and
caches the winner per size class in a JSON file.
"""
from __future__ import annotations

import json
import os
import time

CUS = 256
CANDIDATES = [(1, 256), (1, 512), (2, 256), (2, 512), (2, 1024), (4, 256), (4, 512)]
# Symmetric sum (ipl 8, blocks of 512): js = block pairs per wave.  Its
# partial sums (12 B x n^2 / 1024 in all) are summed in row bands of bounded
# memory, so it is offered at every size the tuner sweeps (up to 8M).
SYM_SEGMENTS = (1, 2, 4, 8, 16, 32, 64)
# ... and block pairs split into parts of tj j particles (one part per wave)
SYM_SPLITS = (256, 128)
SYM_MAX_N = 1 << 23
_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gravity_tuning.json")


def _size_class(n: int) -> str:
    """The power of two nearest to n in ratio (the tuned size whose shape a run
    of n particles takes): n < sqrt(2) k for the largest power of two k <= n."""
    k = 1
    while 2 * k <= n:
        k <<= 1
    return str(k if n * n < 2 * k * k else 2 * k)


def _default(n: int):
    """The shape of a size the table holds no entry for: from 1M particles the
    symmetric sum with 16 block pairs per wave (its partial sums are summed in
    row bands of bounded memory), below that the tiled kernel with enough j
    split for four workgroups per CU."""
    if (1 << 20) <= n <= SYM_MAX_N:
        return 8, 512, 16
    ipl, tj = 2, 512
    blocks = max(1, (n + 256 * ipl - 1) // (256 * ipl))
    js = 1
    while blocks * js < 4 * CUS and js < 32 and (n // tj) // (js * 2) >= 4:
        js *= 2
    return ipl, tj, js


def choose_gravity_config(n: int):
    try:
        with open(_CACHE) as f:
            tab = json.load(f)
        v = tab.get(_size_class(n))
        if v:
            return tuple(v)
    except (OSError, ValueError):
        pass
    return _default(n)


_CELL_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cell_tuning.json")
CELL_BIASES = (-2, -1, 0, 1, 2)


def choose_cell_level(n: int) -> int:
    """Bias of the bucket cut level for a run of n particles (0 when untuned)."""
    try:
        with open(_CELL_CACHE) as f:
            return int(json.load(f).get(_size_class(n), 0))
    except (OSError, ValueError):
        return 0


def tune_cells(n: int, seed: int = 0, steps: int = 3):
    """Time the four neighbour passes of an Evrard step of n particles at every
    cut-level bias (results are the same at every bias; only the candidate
    walks change).  Returns (best bias, {bias: ms per step})."""
    import torch

    from . import ic
    from .plan import SphPlan

    e = ic.evrard(n, seed=seed)
    pl = SphPlan(n)
    pl.load(e.pos, e.vel, e.u, e.m, e.h)
    pl.init()
    res = {}
    for b in CELL_BIASES:
        pl.set_cell_level(b)
        pl.step(1, 1e-4)
        torch.cuda.synchronize()
        pl.set_timing(True)
        pl.step(steps, 1e-4)
        ph = pl.timing()
        pl.set_timing(False)
        res[b] = sum(ph[k][0] / max(ph[k][1], 1)
                     for k in ("smoothing", "density", "divcurl", "hydro"))
    pl.close()
    return min(res, key=res.get), res


def pick(res: dict):
    """The fastest shape, except that among shapes within 1 % of it the one
    needing the least extra memory wins: a tiled shape (none), else the
    symmetric sum with the most block pairs per wave (fewest partial sums),
    whole block pairs before split ones."""
    t_best = min(res.values())
    near = [k for k, v in res.items() if v <= 1.01 * t_best]
    tiled = [k for k in near if k[0] != 8]
    if tiled:
        return min(tiled, key=res.get)
    return max(near, key=lambda k: (k[2], k[1]))


def _sweep(libsx, e, n: int, reps: int, res: dict):
    """Time every candidate shape (stateless call, mean of reps after one
    warm-up) into res."""
    import torch

    def timed(cfg):
        libsx.set_gravity_config(*cfg)
        libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.h)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            libsx.getAcc_g_smth(e.pos, e.m, 1.0, e.h)
        torch.cuda.synchronize()
        res[cfg] = (time.perf_counter() - t0) / reps

    for ipl, tj in CANDIDATES:
        blocks = (n + 256 * ipl - 1) // (256 * ipl)
        for js in (1, 2, 4, 8, 16):
            if js > 1 and blocks * js > 16 * CUS:
                break
            timed((ipl, tj, js))
    if n <= SYM_MAX_N:
        for S in SYM_SEGMENTS:
            timed((8, 512, S))
        for tj in SYM_SPLITS:
            timed((8, tj, 1))


def tune_gravity(n: int, seed: int = 0, reps: int = 2, write: bool = True):
    """Time every candidate shape on an Evrard sphere of n particles."""
    from . import ic, libsx

    e = ic.evrard(n, seed=seed)
    res = {}
    try:
        _sweep(libsx, e, n, reps, res)
    finally:
        libsx.set_gravity_config()  # back to the tuned shapes
    best = pick(res)
    if write:
        try:
            with open(_CACHE) as f:
                tab = json.load(f)
        except (OSError, ValueError):
            tab = {}
        tab[_size_class(n)] = list(best)
        with open(_CACHE, "w") as f:
            json.dump(tab, f, indent=1)
    return best, {f"{k[0]},{k[1]},{k[2]}": v for k, v in res.items()}


def main(argv=None) -> int:
    """Sweep the gravity shapes for several run sizes on the device and write
    the winners (per size class) to --out; copy that file next to this module
    to make them the defaults."""
    import argparse
    import sys

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--sizes", type=int, nargs="+", default=[32768, 131072, 1048576])
    ap.add_argument("--out", required=True)
    ap.add_argument("--cells-out", default=None, help="also tune the bucket cut level")
    a = ap.parse_args(argv)
    if a.cells_out:
        ctab = {}
        for n in a.sizes:
            best, res = tune_cells(n)
            ctab[_size_class(n)] = best
            print(f"[tune] cells n={n} best={best} " +
                  " ".join(f"{k}:{v:.3f}ms" for k, v in sorted(res.items())),
                  file=sys.stderr, flush=True)
        with open(a.cells_out, "w") as f:
            json.dump(ctab, f, indent=1)
    tab = {}
    for n in a.sizes:
        best, res = tune_gravity(n, write=False)
        tab[_size_class(n)] = list(best)
        ranked = sorted(res.items(), key=lambda kv: kv[1])[:4]
        print(f"[tune] n={n} best={best} " +
              " ".join(f"{k}:{1e3 * v:.2f}ms" for k, v in ranked), file=sys.stderr, flush=True)
    with open(a.out, "w") as f:
        json.dump(tab, f, indent=1)
    print(json.dumps(tab))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
