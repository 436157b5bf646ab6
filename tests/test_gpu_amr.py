"""Multi-level block set of uneven block sizes on the GPU (hamr_create_amr)
against the oracle's multi-level step.

Workload: the torus (a = 0.9375, Gamma = 13/9, FM torus r_in = 6, r_max = 12,
seeded +-4% perturbation of u, floors 1e-5 r^-1.5 and 1e-7 r^-2.5) on a base
grid of 64x64x32 cells with level 1 over x1 [0,64) x x2 [32,96) of its
128x128x64 cells and level 2 over x1 [0,64) x x2 [96,160) of its 256x256x128
cells (3 levels, blocks of 64x32x32, 32x16x16 and 16^3 by a seeded choice;
level jumps along x1 and x2, refined blocks on the inner radial edge, polar
rings on level 0); a second set refines only x3 [0,64) of level 2's 128
cells, which adds level jumps along x3, one across the periodic seam.  The corrector's fluxes are corrected at every level jump (the coarse
face takes the mean of the four finer faces) on both sides.

Bars: the ghost fill (exchange tables + outflow + polar mirror + final
outflow pass) and the flags bit-exact; after one FP64 step every variable
within 1e-12 of the oracle cell by cell (|a - b| <= 1e-12 |b|) and dt within
1e-12; the divergence of the field in every interior cell at roundoff over
several steps; graph replay, tuned launches and the per-block stage calls give the
batched step's state bit for bit; after one mixed step, rho and u within 1e-5
of the FP64 step (relative to their largest magnitude), velocity and field
within 1e-5 of the largest component of their vector.
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

NG = 3
DESC = dict(bs=(64, 32, 32), nb=(1, 2, 1))
BOXES = [((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 128))]
SEED = 5


def _new_ctx():
    desc = H.make_desc(DESC["bs"], DESC["nb"])
    amr = H.make_amr_desc(BOXES, seed=SEED)
    ctx = H.AmrContext(desc, amr)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    return desc, amr, ctx


@pytest.fixture(scope="module")
def setup(oracle):
    desc, amr, ctx = _new_ctx()
    g, rows = GT.amr_oracle(oracle, desc, amr)
    assert ctx.nblk == len(rows)
    yield desc, amr, ctx, g, rows
    ctx.close()


def _pointwise(got, want):
    """per-variable max over all blocks' cells of |a - b| / |b| (floor: the
    smallest normal double)"""
    return GT.pointwise_relerr_blocks(got, want)


def _relerr(got, want):
    err = np.zeros(got[0].shape[0])
    scale = np.zeros_like(err)
    for a, b in zip(got, want):
        err = np.maximum(err, np.abs(a - b).reshape(a.shape[0], -1).max(axis=1))
        scale = np.maximum(scale, np.abs(b).reshape(b.shape[0], -1).max(axis=1))
    scale[scale == 0] = 1.0
    return err / scale


def test_block_rows_match_the_host_builder(setup):
    desc, amr, ctx, g, rows = setup
    assert np.array_equal(ctx.layout, rows)
    print("blocks", ctx.info(), flush=True)


def test_ghost_fill_bit_exact(setup):
    desc, amr, ctx, g, rows = setup
    ctx.init_torus(H.torus_desc())
    p = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_P)]
    ps = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_PS)]
    flags = [np.ascontiguousarray(ctx.flags(n)[0]) for n in range(ctx.nblk)]
    ctx.boundprim(1)
    g.bound(p, ps, flags)
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    bad = [b for b in range(ctx.nblk) if not (np.array_equal(gp[b], p[b]) and np.array_equal(gps[b], ps[b]))]
    print("blocks with differing ghosts:", bad[:10], flush=True)
    assert not bad
    for n in range(ctx.nblk):
        assert np.array_equal(ctx.flags(n)[0], flags[n]), n


def test_one_fp64_step_matches_oracle(setup):
    desc, amr, ctx, g, rows = setup
    ctx.set_precision(H.HAMR_PREC_FP64)
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    ctx.clear_flags()
    p = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_P)]
    ps = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_PS)]
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
    dt = ctx.estimate_dt()
    ndt = ctx.step(1, dt=dt)
    want_dt = g.step(p, ps, flags, dt)
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    ep, es = _pointwise(gp, p), _pointwise(gps, ps)
    print("dt %.17g oracle %.17g" % (ndt, want_dt), flush=True)
    print("P rel err per var", np.array2string(ep, precision=2), "face B", np.array2string(es, precision=2), flush=True)
    assert abs(ndt - want_dt) <= 1e-12 * want_dt
    assert np.all(ep <= 1e-12) and np.all(es <= 1e-12)
    for n in range(ctx.nblk):
        assert np.array_equal(ctx.flags(n)[0], flags[n]), n


# level jumps along x3 as well (level 2 over x3 [0,64) of its 128 cells)
BOXES_Z = [((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 64))]


def test_x3_level_jumps_step_matches_oracle(oracle):
    desc = H.make_desc(DESC["bs"], DESC["nb"])
    amr = H.make_amr_desc(BOXES_Z, seed=SEED)
    ctx = H.AmrContext(desc, amr)
    try:
        info = ctx.info()
        print("blocks", info, flush=True)
        assert info["reflux3"] > 0
        g, rows = GT.amr_oracle(oracle, desc, amr)
        ctx.set_precision(H.HAMR_PREC_FP64)
        ctx.set_grid()
        ctx.init_torus(H.torus_desc())
        ctx.boundprim(1)
        start = _state(ctx)
        p = [a.copy() for a in start[0]]  # the oracle steps these in place
        ps = [a.copy() for a in start[1]]
        flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
        dt = ctx.estimate_dt()
        ndt = ctx.step(1, dt=dt)
        want_dt = g.step(p, ps, flags, dt)
        gp, gps = _state(ctx)
        ep, es = _pointwise(gp, p), _pointwise(gps, ps)
        print("P rel err per var", np.array2string(ep, precision=2), "face B", np.array2string(es, precision=2),
              flush=True)
        assert abs(ndt - want_dt) <= 1e-12 * want_dt
        assert np.all(ep <= 1e-12) and np.all(es <= 1e-12)
        # the per-block sequence (flux correction through hamr_flux_correct) gives the same step
        one = (gp, gps)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        H.advance_step(ctx, dt)
        got = _state(ctx)
        assert _same(got[0], one[0]) and _same(got[1], one[1])
    finally:
        ctx.close()


def _state(ctx):
    return ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_graph_tuned_and_per_block_calls_equal_batched_step():
    desc, amr, ctx = _new_ctx()
    try:
        ctx.boundprim(1)
        dt = ctx.estimate_dt()
        start = _state(ctx)
        ctx.step(2, dt=dt)
        ref = _state(ctx)
        # graph replay
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.step(2, dt=dt, graph=True)
        got = _state(ctx)
        assert _same(got[0], ref[0]) and _same(got[1], ref[1])
        # tuned launch shapes
        table = ctx.autotune(1)
        print("tuned:", {k: (v["threads"], v["grid"], v["xcd_order"]) for k, v in table.items()}, flush=True)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.step(2, dt=dt, graph=True)
        got = _state(ctx)
        assert _same(got[0], ref[0]) and _same(got[1], ref[1])
        # persistent grid with XCD order on every kernel
        for slot in range(len(H.HAMR_TUNE_NAMES)):
            ctx.tune_set(slot, 128, 512, 1)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.step(2, dt=dt)
        got = _state(ctx)
        assert _same(got[0], ref[0]) and _same(got[1], ref[1])
        # the fluxes' line-staged tiles (blocks of 16, 32 and 64 cells along each
        # direction: one and two segments per line) and the update in two kernels
        for f in ("flux1", "flux2", "flux3", "cons"):
            ctx.tune_set(H.HAMR_TUNE_NAMES.index(f), 256, 0, 0, 1)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.step(2, dt=dt)
        got = _state(ctx)
        assert _same(got[0], ref[0]) and _same(got[1], ref[1])
        # the update's x1-marching form (blocks of three sizes)
        ctx.tune_set(H.HAMR_TUNE_NAMES.index("cons"), 128, -1, 0)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.step(2, dt=dt)
        got = _state(ctx)
        assert _same(got[0], ref[0]) and _same(got[1], ref[1])
        # the reference-shaped per-block sequence (one step)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.step(1, dt=dt)
        one = _state(ctx)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        H.advance_step(ctx, dt)
        got = _state(ctx)
        assert _same(got[0], one[0]) and _same(got[1], one[1])
        # the same sequence driven by the C host (libhamr_host.so), flux and EMF
        # corrections at the level jumps included
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        H.host_advance(ctx, dt)
        got = _state(ctx)
        assert _same(got[0], one[0]) and _same(got[1], one[1])
    finally:
        ctx.close()


def test_mixed_step_close_to_fp64():
    desc, amr, ctx = _new_ctx()
    try:
        ctx.boundprim(1)
        dt = ctx.estimate_dt()
        start = _state(ctx)
        ctx.set_precision(H.HAMR_PREC_FP64)
        ctx.step(1, dt=dt)
        ref = _state(ctx)
        ctx.set(H.HAMR_P, start[0])
        ctx.set(H.HAMR_PS, start[1])
        ctx.set_precision(H.HAMR_PREC_MIXED)
        ctx.step(1, dt=dt)
        got = _state(ctx)
        ep = _relerr(got[0], ref[0])
        # velocity and field components relative to the largest component of
        # their vector (U2 and B3 start at zero: their own maxima after one
        # step are the small generated values)
        err = np.zeros(9)
        sc = np.zeros(9)
        for a, b in zip(got[0], ref[0]):
            err = np.maximum(err, np.abs(a - b).reshape(9, -1).max(axis=1))
            sc = np.maximum(sc, np.abs(b).reshape(9, -1).max(axis=1))
        ev = err[2:5].max() / sc[2:5].max()
        eb = err[5:8].max() / sc[5:8].max()
        print("mixed vs FP64 after one step, per var:", np.array2string(ep, precision=2),
              "velocity %.2e field %.2e (of the vector's largest component)" % (ev, eb), flush=True)
        assert ep[0] <= 1e-5 and ep[1] <= 1e-5 and eb <= 1e-5 and ev <= 1e-5
    finally:
        ctx.close()
