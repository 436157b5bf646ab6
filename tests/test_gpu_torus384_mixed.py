"""The mixed-precision step on the 384x192x192 torus (12x6x6 blocks of 32^3,
14.2 M zones) against the FP64 step, after 1 and 100 steps.

Workload: the torus (r 1.2-1000 M log-spaced, Kerr a = 0.9375, Gamma = 13/9,
FM torus r_in = 6, r_max = 12, min beta = 100, seeded +-4% perturbation of u,
floors 1e-5 r^-1.5 and 1e-7 r^-2.5).  The FP64 step is the oracle's arithmetic
bit for bit (tests/test_gpu_torus256.py holds it to the oracle over 100 steps),
so its state stands for the oracle's here.  All runs advance with the FP64
run's dt sequence from the same initial state.

Printed at steps 1 and 100, per variable: the max error relative to the
variable's largest magnitude, the L1 error (sum |d| / sum |ref| over the
interior), and for the field the max error relative to the largest field
component.  Beside the mixed step, a second FP64 run starts from the initial
state rounded to FP32 (every primitive and face field): its deviation from the
FP64 run is the flow's own response to a perturbation of FP32 size, the error
floor of any arithmetic that stores the state or its fluxes in FP32.

Bars (asserted): after one step every variable within 1e-6 of its largest
magnitude; after 100 steps rho within 1e-5 (max) and the field within 1e-5 in
L1 and within 1e-4 of its largest component (max).  The max error of the field
after 100 steps is of the size of the FP32-rounding floor above (the printed
ratio): limiter and upwind selections that flip between nearly equal
candidates move a few cells by O(1) of their local gradient.
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from hamr_gamma_wind_amd import workloads as W

pytestmark = pytest.mark.gpu

NG = 3
NAMES = ["rho", "u", "U1", "U2", "U3", "B1", "B2", "B3"]


def errors(a, b):
    I = (slice(None), slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
    a, b = a[I], b[I]
    out = {}
    for k, name in enumerate(NAMES):
        d = np.abs(a[:, k] - b[:, k])
        r = np.abs(b[:, k])
        out[name] = (float(d.max() / max(r.max(), 1e-300)), float(d.sum() / max(r.sum(), 1e-300)))
    d = np.abs(a[:, 5:8] - b[:, 5:8])
    out["field"] = (float(d.max() / np.abs(b[:, 5:8]).max()),
                    float(d.sum() / np.abs(b[:, 5:8]).sum()))
    return out


def fmt(e):
    return "; ".join("%s max %.2e L1 %.2e" % (k, v[0], v[1]) for k, v in e.items())


@pytest.fixture(scope="module")
def runs():
    bs, nb = W.TORUS_384
    desc = H.make_desc(bs, nb)
    f64 = H.Context(desc)
    f64.set_grid()
    f64.init_torus(H.torus_desc())
    f64.boundprim(1)
    p0, ps0 = f64.get(H.HAMR_P), f64.get(H.HAMR_PS)
    mix = H.Context(desc)
    mix.set_grid()
    mix.set(H.HAMR_P, p0)
    mix.set(H.HAMR_PS, ps0)
    mix.set_precision(H.HAMR_PREC_MIXED)
    pert = H.Context(desc)
    pert.set_grid()
    pert.set(H.HAMR_P, p0.astype(np.float32).astype(np.float64))
    pert.set(H.HAMR_PS, ps0.astype(np.float32).astype(np.float64))
    pert.boundprim(1)
    del p0, ps0
    st = dict(f64=f64, mix=mix, pert=pert, step=0, dt=f64.estimate_dt())
    yield st
    for c in (f64, mix, pert):
        c.close()


def _advance(st, upto):
    while st["step"] < upto:
        dt = st["dt"]
        ndt = st["f64"].step(1, dt)
        st["mix"].step(1, dt)
        st["pert"].step(1, dt)
        st["step"], st["dt"] = st["step"] + 1, ndt


def _report(st):
    ref = st["f64"].get(H.HAMR_P)
    em = errors(st["mix"].get(H.HAMR_P), ref)
    ep = errors(st["pert"].get(H.HAMR_P), ref)
    print("step %d mixed vs FP64: %s" % (st["step"], fmt(em)), flush=True)
    print("step %d FP64 from the FP32-rounded state vs FP64: %s" % (st["step"], fmt(ep)), flush=True)
    print("step %d field max error, mixed / FP32-rounding floor: %.2f" % (st["step"], em["field"][0] /
                                                                         max(ep["field"][0], 1e-300)), flush=True)
    return em, ep


def test_one_step(runs):
    _advance(runs, 1)
    em, ep = _report(runs)
    assert all(em[k][0] <= 1e-6 for k in ("rho", "u", "field")), em


def test_hundred_steps(runs):
    _advance(runs, 100)
    em, ep = _report(runs)
    assert em["rho"][0] <= 1e-5, em["rho"]
    assert em["field"][1] <= 1e-5, em["field"]
    assert em["field"][0] <= 1e-4, em["field"]
