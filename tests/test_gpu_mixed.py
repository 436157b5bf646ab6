"""The mixed-precision step (FP32 reconstruction, fluxes, speeds and EMFs; FP64
metric and connection, face-field update, compensated conserved update and
Noble inversions) against the FP64 step on the same state.

Workload: the 64x32x32 torus (r 1.2-1000 M, a = 0.9375, Gamma = 13/9, FM torus
r_in = 6, r_max = 12, seeded +-4% perturbation of u, floors 1e-5 r^-1.5 and
1e-7 r^-2.5) as 4x4x4 blocks of 16x8x8.

Bars (tolerances of an FP32 flux path, relative to each variable's largest
magnitude): fluxes within 1e-4 of the FP64 fluxes (the flux of the normal
field component, zero but for roundoff in both, is left out); one step within
1e-5; after 100 steps the L1 error of rho and of every field component within
1e-5.  The mixed step's own sequencing invariants hold bitwise (per-block
calls == batched step == graph replay).
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

NG = 3
I = slice(NG, -NG)
RHO, B1, B2, B3 = 0, 5, 6, 7


@pytest.fixture(scope="module")
def setup():
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    p0 = ctx.get(H.HAMR_P)
    ps0 = ctx.get(H.HAMR_PS)
    yield desc, ctx, p0, ps0
    ctx.close()


def _reset(ctx, p0, ps0, mode):
    ctx.set_precision(mode)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)


def _l1(a, b):
    return np.abs(a - b).sum() / np.abs(b).sum()


def test_mixed_fluxes_round_the_fp64_fluxes(setup):
    desc, ctx, p0, ps0 = setup
    F = {}
    for mode in (H.HAMR_PREC_FP64, H.HAMR_PREC_MIXED):
        _reset(ctx, p0, ps0, mode)
        for d in (1, 2, 3):
            for n in range(ctx.nblk):
                ctx.fluxcalc(d, 0, n)
        F[mode] = [ctx.get(s) for s in (H.HAMR_F1, H.HAMR_F2, H.HAMR_F3)]
    for d in range(3):
        sl = [slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)]
        sl[d] = slice(NG, NG + desc.bs[d] + 1)
        a = F[H.HAMR_PREC_MIXED][d][:, :, sl[0], sl[1], sl[2]]
        b = F[H.HAMR_PREC_FP64][d][:, :, sl[0], sl[1], sl[2]]
        err = GT.relerr_by_var(a, b)
        print("F%d mixed vs FP64: %s" % (d + 1, np.array2string(err, precision=2)), flush=True)
        err[B1 + d] = 0.0
        assert np.all(err <= 1e-4), err
        # and they are FP32 values
        assert np.array_equal(a, a.astype(np.float32).astype(np.float64))
    ctx.set_precision(H.HAMR_PREC_FP64)


def test_mixed_one_step_close_to_fp64(setup):
    desc, ctx, p0, ps0 = setup
    out = {}
    for mode in (H.HAMR_PREC_FP64, H.HAMR_PREC_MIXED):
        _reset(ctx, p0, ps0, mode)
        dt = ctx.estimate_dt()
        ndt = ctx.step(1, dt)
        out[mode] = (dt, ndt, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
    d64, n64, p64, s64 = out[H.HAMR_PREC_FP64]
    d32, n32, p32, s32 = out[H.HAMR_PREC_MIXED]
    assert abs(d32 - d64) <= 1e-5 * d64 and abs(n32 - n64) <= 1e-5 * n64
    err = GT.relerr_by_var(p32[:, :, I, I, I], p64[:, :, I, I, I])
    print("one step, mixed vs FP64, max rel err per var:", np.array2string(err, precision=2), flush=True)
    assert np.all(err <= 1e-5), err
    ctx.set_precision(H.HAMR_PREC_FP64)


def test_mixed_per_block_sequence_equals_batched_and_graph(setup):
    desc, ctx, p0, ps0 = setup
    _reset(ctx, p0, ps0, H.HAMR_PREC_MIXED)
    dt = ctx.estimate_dt()
    ndt_a = H.advance_step(ctx, dt)
    pa, psa = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    res = []
    for graph in (False, True):
        _reset(ctx, p0, ps0, H.HAMR_PREC_MIXED)
        res.append((ctx.step(1, dt, graph=graph), ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)))
    for ndt, pb, psb in res:
        assert ndt == ndt_a
        assert np.array_equal(pa, pb) and np.array_equal(psa, psb)
    ctx.set_precision(H.HAMR_PREC_FP64)


def test_mixed_100_steps_rho_and_field_error(setup):
    desc, ctx, p0, ps0 = setup
    out = {}
    for mode in (H.HAMR_PREC_FP64, H.HAMR_PREC_MIXED):
        _reset(ctx, p0, ps0, mode)
        dt = ctx.estimate_dt()
        ndt = ctx.step(100, dt, graph=True)
        out[mode] = (ndt, ctx.get(H.HAMR_P))
    n64, p64 = out[H.HAMR_PREC_FP64]
    n32, p32 = out[H.HAMR_PREC_MIXED]
    a, b = p32[:, :, I, I, I], p64[:, :, I, I, I]
    maxerr = GT.relerr_by_var(a, b)
    l1 = np.array([_l1(a[:, k], b[:, k]) for k in range(9)])
    print("100 steps, mixed vs FP64: max rel err per var", np.array2string(maxerr, precision=2),
          "L1 rel err per var", np.array2string(l1, precision=2), "dt", n32, n64, flush=True)
    for k in (RHO, B1, B2, B3):
        assert l1[k] <= 1e-5, (k, l1[k])
    ctx.set_precision(H.HAMR_PREC_FP64)
