"""FP64 GPU step against the CPU oracle over many steps (not collected by
pytest: it takes minutes on the host's cores).

Workload: the torus on the single-level 256x128x128 grid as 8x4x4 blocks of
32^3 (r 1.2-1000 M log-spaced, a = 0.9375, Gamma = 13/9, FM torus r_in = 6,
r_max = 12, min beta = 100, seeded +-4% perturbation of u, floors 1e-5 r^-1.5
and 1e-7 r^-2.5), both sides stepped with the oracle's dt sequence from the
same initial state.  Prints the relative L1 error (sum |d| / sum |ref| over the
interior) and the max error (relative to the variable's largest magnitude) of
rho, u and the field every `--every` steps.

    python tests/fp64_long_run.py --steps 100
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import hamr_gamma_wind_amd as H  # noqa: E402
from tests import gridtools as GT  # noqa: E402
from tests import oracle_ctypes as O  # noqa: E402


def errors(desc, ctx, P):
    p = GT.blocks_to_global(desc, ctx.get(H.HAMR_P))
    I = (slice(3, -3),) * 3
    out = {}
    for name, k in (("rho", 0), ("u", 1), ("B1", 5), ("B2", 6), ("B3", 7)):
        d = np.abs(p[k][I] - P[k][I])
        ref = np.abs(P[k][I])
        out[name] = (float(d.sum() / ref.sum()), float(d.max() / ref.max()))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--every", type=int, default=10)
    args = ap.parse_args()
    desc = H.make_desc((32, 32, 32), (8, 4, 4))
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    grid = GT.oracle_grid(O.load(), desc)
    P = GT.blocks_to_global(desc, ctx.get(H.HAMR_P))
    PS = GT.blocks_to_global(desc, ctx.get(H.HAMR_PS))
    grid.bound(P, PS)
    dt = ctx.estimate_dt()
    t0 = time.time()
    for s in range(1, args.steps + 1):
        ndt, _ = grid.step(P, PS, dt)
        ctx.step(1, dt=dt)
        dt = ndt
        if s % args.every == 0 or s == args.steps:
            e = errors(desc, ctx, P)
            print("step %d (%.0f s): " % (s, time.time() - t0)
                  + ", ".join("%s L1 %.2e max %.2e" % (k, a, b) for k, (a, b) in e.items()), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
