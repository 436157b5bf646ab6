"""Error of the mixed-precision step against the FP64 step over many steps
(not collected by pytest).

Workload: the torus (r 1.2-1000 M log-spaced, a = 0.9375, Gamma = 13/9, FM
torus r_in = 6, r_max = 12, min beta = 100, seeded +-4% perturbation of u,
floors 1e-5 r^-1.5 and 1e-7 r^-2.5) on 384x192x192 cells as 12x6x6 blocks of
32^3 (or 256x128x128 as 8x4x4).  Both arithmetics start from the same state
and advance with the FP64 run's dt sequence; the FP64 GPU step is the oracle's
arithmetic bit for bit (tests/fp64_long_run.py), so the figures are the mixed
step's error against the FP64 oracle.  Printed at the requested steps: per
variable, the max error relative to the variable's largest magnitude and the
L1 error (sum |d| / sum |ref|) over the interior; for the velocity and the
field also the max error relative to the largest component of the vector.

    python tests/mixed_error_run.py --config c3 --report 1,100
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import hamr_gamma_wind_amd as H  # noqa: E402

CONFIGS = {"c2": ((32, 32, 32), (8, 4, 4)), "c3": ((32, 32, 32), (12, 6, 6))}
NAMES = ["rho", "u", "U1", "U2", "U3", "B1", "B2", "B3"]


def report(step, a, b):
    I = (slice(None), slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    a, b = a[I], b[I]
    parts = []
    for k, name in enumerate(NAMES):
        d = np.abs(a[:, k] - b[:, k])
        r = np.abs(b[:, k])
        parts.append("%s max %.2e L1 %.2e" % (name, d.max() / max(r.max(), 1e-300), d.sum() / max(r.sum(), 1e-300)))
    for name, k0, k1 in (("velocity", 2, 5), ("field", 5, 8)):
        d = np.abs(a[:, k0:k1] - b[:, k0:k1]).max()
        parts.append("%s max/|vector| %.2e" % (name, d / np.abs(b[:, k0:k1]).max()))
    print("step %d: " % step + "; ".join(parts), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--report", default="1,100", help="steps at which to print the errors")
    args = ap.parse_args()
    steps = sorted(int(x) for x in args.report.split(","))
    bs, nb = CONFIGS[args.config]
    desc = H.make_desc(bs, nb)
    f64 = H.Context(desc)
    f64.set_grid()
    f64.init_torus(H.torus_desc())
    f64.boundprim(1)
    f64.set_precision(H.HAMR_PREC_FP64)
    mix = H.Context(desc)
    mix.set_grid()
    mix.set(H.HAMR_P, f64.get(H.HAMR_P))
    mix.set(H.HAMR_PS, f64.get(H.HAMR_PS))
    mix.set_precision(H.HAMR_PREC_MIXED)
    dt = f64.estimate_dt()
    t0 = time.time()
    done = 0
    for s in steps:
        for _ in range(s - done):
            ndt = f64.step(1, dt)
            mix.step(1, dt)
            dt = ndt
        done = s
        report(s, mix.get(H.HAMR_P), f64.get(H.HAMR_P))
    print("%s: %d steps in %.0f s" % (args.config, done, time.time() - t0), flush=True)
    f64.close()
    mix.close()


if __name__ == "__main__":
    main()
