"""Local adaptive time stepping on the GPU (hamr_set_timelevels,
hamr_step_las, hamr_timelevels_from_dt) against the oracle's cycle
(oracle_amr_las_cycle) and the standard step.

Workload: the three-level torus set of test_gpu_amr (base 64x64x32, level 1
over x1 [0,64) x x2 [32,96), level 2 over x1 [0,64) x x2 [96,160), blocks of
64x32x32, 32x16x16 and 16^3), FM torus with a uniform toroidal field B^3 =
0.005 added (divergence-free on the axisymmetric grid and regular on the
axis) so that every interface sees nonzero EMFs; dt a quarter of the Courant
dt.

Bars:
- every block at time level 1, maxtl 1: hamr_step_las equals hamr_step bit for
  bit (state, flags, next dt); every block at time level 2, maxtl 2, with dt:
  equal to hamr_step with 2 dt bit for bit;
- mixed time levels (level-0 blocks 4 or 2, the polar rings 4, level-1 blocks
  2 or 1, level-2 blocks 1; maxtl 4): after one FP64 cycle every variable
  within 1e-12 of its largest magnitude in every cell, and within 1e-10 of
  itself in every cell where it is at least 1e-2 of that magnitude; flags
  equal; the next dt and every block's Courant dt within 1e-12.  (The finest
  blocks take four steps in the cycle.  One step agrees cell by cell, as
  test_gpu_amr checks; over several steps the inversion in atmosphere cells,
  where u is 1e-9 of its maximum and comes from a difference of conserved
  quantities a million times larger, magnifies the last-digit differences of
  the two updates, so that there the agreement is relative to the variable's
  scale); the mixed-precision cycle's difference from the FP64 cycle (rho
  and u relative to their largest magnitude, velocity and field to the
  largest component of their vector) within ten times that of four standard
  steps, rho within 1e-6 and the field within 1e-4;
- three replays of the captured cycle equal three launched cycles bit for bit;
- the time levels from the block dts: powers of 2 up to maxtl, none above
  set_timelevel's rule for its block, equal over each polar ring, accepted by
  hamr_set_timelevels, and a cycle with them runs;
- bad time levels (not a power of 2, unequal over a ring, a finer block at a
  level jump slower than its coarser neighbour), a cycle before any time
  levels or with maxtl below the largest, and a uniform context are refused
  with a message.
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

DESC = dict(bs=(64, 32, 32), nb=(1, 2, 1))
BOXES = [((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 128))]
BG = np.array([0.0, 0.0, 0.005])


def _new_ctx():
    desc = H.make_desc(DESC["bs"], DESC["nb"])
    amr = H.make_amr_desc(BOXES, seed=5)
    ctx = H.AmrContext(desc, amr)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    p, ps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    for a, c in zip(p, ps):
        a[5:8] += BG[:, None, None, None]
        c += BG[:, None, None, None]
    ctx.set(H.HAMR_P, p)
    ctx.set(H.HAMR_PS, ps)
    ctx.boundprim(1)
    return desc, amr, ctx


@pytest.fixture(scope="module")
def las(oracle):
    desc, amr, ctx = _new_ctx()
    g, rows = GT.amr_oracle(oracle, desc, amr)
    assert np.array_equal(ctx.layout, rows)
    start = (ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
    dt = 0.25 * ctx.estimate_dt()
    yield ctx, g, rows, start, dt
    ctx.close()


def _reset(ctx, start):
    ctx.set(H.HAMR_P, start[0])
    ctx.set(H.HAMR_PS, start[1])
    ctx.clear_flags()


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_unit_and_uniform_time_levels_equal_standard_steps(las):
    ctx, g, rows, start, dt = las
    ctx.set_precision(H.HAMR_PREC_FP64)
    nb = ctx.nblk
    print("unit time levels, maxtl 1", flush=True)
    _reset(ctx, start)
    want_dt = ctx.step(1, dt=dt)
    want = (ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS), [f for f, _ in ctx.all_flags()])
    _reset(ctx, start)
    ctx.set_timelevels(np.ones(nb, dtype=np.int32))
    got_dt = ctx.step_las(1, 1, dt=dt)
    got = (ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS), [f for f, _ in ctx.all_flags()])
    print("dt: step %.17g cycle %.17g" % (want_dt, got_dt), flush=True)
    assert got_dt == want_dt
    assert _same(got[0], want[0]) and _same(got[1], want[1]) and _same(got[2], want[2])
    print("time level 2 everywhere, maxtl 2, against a step of 2 dt", flush=True)
    _reset(ctx, start)
    ctx.step(1, dt=2 * dt)
    want = (ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
    _reset(ctx, start)
    ctx.set_timelevels(np.full(nb, 2, dtype=np.int32))
    ctx.step_las(2, 1, dt=dt)
    got = (ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
    assert _same(got[0], want[0]) and _same(got[1], want[1])


def test_graph_replay_equals_launches(las):
    """Three cycles launched one by one against three replays of the captured
    cycle, each replay checked against the launched state after the same
    number of cycles, with state reads in between."""
    ctx, g, rows, start, dt = las
    ctx.set_precision(H.HAMR_PREC_FP64)
    ctx.set_timelevels(GT.las_levels(rows))
    _reset(ctx, start)
    want = []
    d = dt
    for k in range(3):
        d = ctx.step_las(4, 1, dt=d)
        want.append((ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS), d))
    _reset(ctx, start)
    d = dt
    for k in range(3):
        d = ctx.step_las(4, 1, dt=d, graph=True)
        got = (ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
        print("replay %d: dt %.17g launched %.17g" % (k, d, want[k][2]), flush=True)
        assert d == want[k][2] and _same(got[0], want[k][0]) and _same(got[1], want[k][1]), k


def test_mixed_time_levels_match_the_oracle(las):
    ctx, g, rows, start, dt = las
    ctx.set_precision(H.HAMR_PREC_FP64)
    tl = GT.las_levels(rows)
    _reset(ctx, start)
    ctx.set_timelevels(tl)
    info = ctx.las_info()
    print("time levels per block count", np.bincount(tl).tolist(), "corrections", info, flush=True)
    assert info["tlmax"] == 4 and info["flux1"] + info["flux2"] + info["flux3"] > 0 and info["emf_same"] > 0
    ndt = ctx.step_las(4, 1, dt=dt)
    bdt = ctx.block_dts()
    gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    p = [np.array(a, copy=True) for a in start[0]]
    ps = [np.array(a, copy=True) for a in start[1]]
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
    print("oracle cycle", flush=True)
    want_dt, want_bdt = g.las_cycle(tl, 4, p, ps, flags, dt)
    print("dt %.17g oracle %.17g; block dt max rel %.2e" % (ndt, want_dt, np.max(np.abs(bdt - want_bdt) / want_bdt)),
          flush=True)
    ea, _ = _two_tier(gp, p)
    fa, _ = _two_tier(gps, ps)
    print("P: |diff| / max per var", np.array2string(ea, precision=2), "face B", np.array2string(fa, precision=2),
          flush=True)
    for above in (1e-6, 1e-4, 1e-2):
        print("   pointwise where |value| >= %.0e of max: P %s, face B %s" % (
            above, np.array2string(_two_tier(gp, p, above)[1], precision=2),
            np.array2string(_two_tier(gps, ps, above)[1], precision=2)), flush=True)
    er, fr = _two_tier(gp, p, 1e-2)[1], _two_tier(gps, ps, 1e-2)[1]
    assert abs(ndt - want_dt) <= 1e-12 * want_dt
    assert np.all(np.abs(bdt - want_bdt) <= 1e-12 * want_bdt)
    assert np.all(ea <= 1e-12) and np.all(fa <= 1e-12)
    assert np.all(er <= 1e-10) and np.all(fr <= 1e-10)
    for n in range(ctx.nblk):
        assert np.array_equal(ctx.flags(n)[0], flags[n]), n


def _two_tier(got, want, above=1e-6):
    """per component: max |a - b| / max |b| over every cell of every block, and
    max |a - b| / |b| over the cells with |b| >= above * max |b|"""
    nc = want[0].shape[0]
    big = np.zeros(nc)
    for b in want:
        big = np.maximum(big, np.abs(b).reshape(nc, -1).max(axis=1))
    big[big == 0] = 1.0
    ea, er = np.zeros(nc), np.zeros(nc)
    for a, b in zip(got, want):
        d = np.abs(a - b).reshape(nc, -1)
        bb = np.abs(b).reshape(nc, -1)
        ea = np.maximum(ea, d.max(axis=1) / big)
        m = bb >= above * big[:, None]
        er = np.maximum(er, np.where(m, d / np.maximum(bb, 1e-300), 0.0).max(axis=1))
    return ea, er


def _mixed_error(ctx, start, run):
    out = {}
    for prec in (H.HAMR_PREC_FP64, H.HAMR_PREC_MIXED):
        ctx.set_precision(prec)
        _reset(ctx, start)
        run()
        out[prec] = ctx.get(H.HAMR_P)
    ctx.set_precision(H.HAMR_PREC_FP64)
    a, b = out[H.HAMR_PREC_MIXED], out[H.HAMR_PREC_FP64]
    assert all(np.isfinite(x).all() for x in a)
    err = {}
    for name, sl in (("rho", [0]), ("u", [1]), ("v", [2, 3, 4]), ("B", [5, 6, 7])):
        e = max(float(np.abs(x[sl] - y[sl]).max()) for x, y in zip(a, b))
        s = max(float(np.abs(y[sl]).max()) for y in b)
        err[name] = e / s
    return err


def test_mixed_precision_cycle_close_to_fp64(las):
    """The mixed-precision cycle against the FP64 cycle, next to the same
    measure for four standard steps (what the finest blocks take in the
    cycle): the mixed arithmetic's error over several steps is set by a few
    atmosphere cells near the floors, 1e-3 of the largest velocity after four
    steps either way; the cycle may not add to it beyond a factor of ten."""
    ctx, g, rows, start, dt = las
    tl = GT.las_levels(rows)
    ctx.set_timelevels(tl)
    e_las = _mixed_error(ctx, start, lambda: ctx.step_las(4, 1, dt=dt))
    e_std = _mixed_error(ctx, start, lambda: ctx.step(4, dt=dt))
    print("mixed vs FP64, cycle:", {k: "%.2e" % v for k, v in e_las.items()}, flush=True)
    print("mixed vs FP64, four standard steps:", {k: "%.2e" % v for k, v in e_std.items()}, flush=True)
    for k in e_las:
        assert e_las[k] <= max(10 * e_std[k], 1e-8), (k, e_las[k], e_std[k])
    assert e_las["rho"] < 1e-6 and e_las["B"] < 1e-4


def test_time_levels_from_block_dts(las):
    ctx, g, rows, start, dt = las
    ctx.set_precision(H.HAMR_PREC_FP64)
    _reset(ctx, start)
    ndt = ctx.step(1, dt=dt)
    bdt = ctx.block_dts()
    tl = ctx.timelevels_from_dt(4)
    raw = GT.las_rule_py(bdt, ndt, 4, [], [])
    print("time levels from dt:", np.bincount(tl).tolist(), "unconstrained", np.bincount(raw).tolist(), flush=True)
    assert np.array_equal(ctx.timelevels(), tl)
    assert set(np.unique(tl).tolist()) <= {1, 2, 4} and np.all(tl <= raw)
    # the polar rings (blocks on the axis: same level, side, radial range and shape)
    rings = {}
    for n, r in enumerate(rows):
        if int(r[8]):
            rings.setdefault((int(r[0]), int(r[8]), int(r[1]), int(r[4]), int(r[5]), int(r[6])), []).append(n)
    assert rings
    for members in rings.values():
        assert len(set(tl[members].tolist())) == 1, members
    ctx.step_las(4, 1)


def test_bad_time_levels_are_refused(las):
    ctx, g, rows, start, dt = las
    nb = ctx.nblk
    tl = GT.las_levels(rows)
    bad = tl.copy()
    bad[0] = 3
    with pytest.raises(H.HamrError, match="powers of 2"):
        ctx.set_timelevels(bad)
    ring = [n for n, r in enumerate(rows) if int(r[8])]
    bad = tl.copy()
    bad[ring[0]] = 2
    with pytest.raises(H.HamrError, match="ring"):
        ctx.set_timelevels(bad)
    bad = np.ones(nb, dtype=np.int32)
    bad[[n for n, r in enumerate(rows) if int(r[0]) == 2][0]] = 2
    with pytest.raises(H.HamrError, match="level jump"):
        ctx.set_timelevels(bad)
    ctx.set_timelevels(tl)
    with pytest.raises(H.HamrError, match="maxtl"):
        ctx.step_las(2, 1)
    desc, amr, fresh = _new_ctx()
    try:
        with pytest.raises(H.HamrError, match="set the time levels"):
            fresh.step_las(1, 1)
    finally:
        fresh.close()
    uni = H.Context(H.make_desc((16, 8, 8), (2, 2, 2)))
    try:
        v = np.ones(8, dtype=np.int32)
        assert uni.L.hamr_set_timelevels(uni.h, H._ptr(v, H._ip)) != 0
        assert "multi-level" in uni.L.hamr_last_error(uni.h).decode()
    finally:
        uni.close()
