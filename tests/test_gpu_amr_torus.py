"""The multi-level torus of the benchmark: 722 blocks of 64x32x32, 32x16x16 and
16^3 on three levels (ratio 2; level 1: r <= 80 M, 30-150 deg; level 2:
r <= 34.6 M, 60-120 deg; 5.21 M zones), against the oracle's multi-level step.

Workload: the torus (Kerr a = 0.9375, Gamma = 13/9, FM torus r_in = 6, r_max =
12, min beta = 100, seeded +-4% perturbation of u, floors 1e-5 r^-1.5 and
1e-7 r^-2.5) on that block set, every block resident.

Bars: the ghost fill (exchange tables, radial outflow, polar mirror, final
outflow pass) bit-exact; one FP64 step: every primitive and face field within
1e-12 of the oracle cell by cell (|a - b| <= 1e-12 |b|), flags bit-exact, dt
within 1e-12; the divergence of the field in every interior cell at roundoff
(1e-12 of the block's scale) after three steps in each arithmetic.
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from hamr_gamma_wind_amd import workloads as W
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

NG = 3


@pytest.fixture(scope="module")
def setup(oracle):
    A = W.AMR_TORUS
    desc = H.make_desc(A["bs"], A["nb"])
    amr = H.make_amr_desc(A["boxes"], seed=A["seed"])
    ctx = H.AmrContext(desc, amr)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    g, rows = GT.amr_oracle(oracle, desc, amr)
    assert ctx.nblk == len(rows) == 722
    yield desc, amr, ctx, g, rows
    ctx.close()


def _copy(arrs):
    return [np.ascontiguousarray(a).copy() for a in arrs]


def test_ghost_fill_bit_exact(setup):
    desc, amr, ctx, g, rows = setup
    ctx.init_torus(H.torus_desc())
    p, ps = _copy(ctx.get(H.HAMR_P)), _copy(ctx.get(H.HAMR_PS))
    flags = [np.ascontiguousarray(ctx.flags(n)[0]) for n in range(ctx.nblk)]
    ctx.boundprim(1)
    g.bound(p, ps, flags)
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    bad = [b for b in range(ctx.nblk) if not (np.array_equal(gp[b], p[b]) and np.array_equal(gps[b], ps[b]))]
    print("blocks with differing ghosts:", bad[:10], flush=True)
    assert not bad
    assert all(np.array_equal(ctx.flags(n)[0], flags[n]) for n in range(ctx.nblk))


def test_one_fp64_step_matches_oracle(setup):
    desc, amr, ctx, g, rows = setup
    ctx.set_precision(H.HAMR_PREC_FP64)
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    ctx.clear_flags()
    p, ps = _copy(ctx.get(H.HAMR_P)), _copy(ctx.get(H.HAMR_PS))
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
    dt = ctx.estimate_dt()
    ndt = ctx.step(1, dt=dt)
    want_dt = g.step(p, ps, flags, dt)
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    J = (slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
    ep = GT.pointwise_relerr_blocks([a[J] for a in gp], [b[J] for b in p])
    es = GT.pointwise_relerr_blocks([a[J] for a in gps], [b[J] for b in ps])
    print("dt %.17g oracle %.17g" % (ndt, want_dt), flush=True)
    print("P pointwise rel err per var", np.array2string(ep, precision=2), "face B", np.array2string(es, precision=2),
          flush=True)
    assert abs(ndt - want_dt) <= 1e-12 * want_dt
    assert np.all(ep <= 1e-12) and np.all(es <= 1e-12)
    assert all(np.array_equal(ctx.flags(n)[0], flags[n]) for n in range(ctx.nblk))


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_div_b_stays_at_roundoff(setup, mode):
    desc, amr, ctx, g, rows = setup
    _, dx0 = H.desc_coords(desc)
    ctx.set_precision(H.HAMR_PREC_FP64 if mode == "fp64" else H.HAMR_PREC_MIXED)
    try:
        ctx.init_torus(H.torus_desc())
        ctx.boundprim(1)
        w0 = GT.amr_divb(g._m, rows, dx0, ctx.get(H.HAMR_PS))
        dt = ctx.estimate_dt()
        w = []
        for _ in range(3):
            dt = ctx.step(1, dt)
            w.append(GT.amr_divb(g._m, rows, dx0, ctx.get(H.HAMR_PS)))
        print("%s: div B relative to the block scale: initial %.2e, steps 1-3 %s" % (mode, w0, w), flush=True)
        assert max(w) < 1e-12
    finally:
        ctx.set_precision(H.HAMR_PREC_FP64)
