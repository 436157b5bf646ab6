"""The HLLC fluxes (hamr_set_riemann(ctx, HAMR_RIEMANN_HLLC): fluxcalc2D_FT,
GPU_hllc.h:3-247, with the tetrads of set_Mud, hllc.c:20-178) on the GPU
against the oracle's restatement.

Workloads: the 64x32x32 torus grid in 4x4x4 blocks of 16x8x8 (a = 0.9375,
Gamma = 13/9, FM torus r_in = 6, r_max = 12, seeded +-4% perturbation of u,
floors 1e-5 r^-1.5 and 1e-7 r^-2.5; ghosts filled), compared with the oracle
on the same state as one global grid; and the three-level set of the
multi-level tests.

Bars: the device tetrads equal the oracle's from the same metric bit for bit;
fluxes and the per-block Courant minima within 1e-12 of the oracle cell by
cell (|a - b| <= 1e-12 |b|); after one FP64 step every variable and dt within
1e-12; the per-block stage calls, the batched step and the graph
replay give the same state bit for bit.  The mixed-precision HLLC (FP32
fluxes) against the FP64 HLLC, relative to each variable's largest magnitude:
x1 and x3 fluxes within 5e-4 and 1e-5, x2 fluxes (near zero in the torus at
rest, where the solver switches at an exact-zero contact speed) in L1 within
1e-2; one step within 2.5e-4 (face fields 1e-5); after 100 steps the L1 error
of rho within 1e-6 and of B within 2e-5 (measured 1.12e-5 in B3).
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

NG = 3
TOL = 1e-12
# Fluxes: 1e-12 cell by cell, but 1e-10 for the entropy flux (KTOT), whose
# factor P rho^-Gamma comes from the device's pow and the oracle's glibc pow,
# which may differ in the last ulp; where the HLL (or HLLC) combination of the
# two sides nearly cancels, that ulp is a larger part of the face's flux
# (observed 1.5e-12 and 2.3e-11; every other component agrees bit for bit).
FLUX_TOL = np.array([TOL] * 8 + [1e-10])


def _check(name, got, want, tol=TOL):
    err = GT.pointwise_relerr(got, want)
    print("%s: max rel err per var %s" % (name, np.array2string(err, precision=2)), flush=True)
    assert np.all(err <= tol), (name, err)


@pytest.fixture(scope="module")
def setup(oracle):
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    ctx.set_riemann(H.HAMR_RIEMANN_HLLC)
    grid = GT.oracle_grid(oracle, desc, riemann=1)
    p0 = ctx.get(H.HAMR_P)
    ps0 = ctx.get(H.HAMR_PS)
    yield desc, ctx, grid, p0, ps0
    ctx.close()


def _global_state(desc, grid, p0, ps0):
    P = GT.blocks_to_global(desc, p0)
    PS = GT.blocks_to_global(desc, ps0)
    grid.bound(P, PS)
    return P, PS


def test_device_tetrads_match_oracle(setup, oracle):
    desc, ctx, grid, p0, ps0 = setup
    D1, D2 = ctx.D[0], ctx.D[1]
    for geo in range(ctx.ngeo):
        tet = ctx.download_tetrads(geo)
        m = ctx.download_metric(geo)
        for d in (1, 2, 3):
            pos = (H.FACE1, H.FACE2, H.FACE3)[d - 1]
            for col in range(0, D1 * D2, 7):
                mud, mudi = oracle.tetrad(m["gcov"][pos, :, col], m["gcon"][pos, :, col], d)
                assert np.array_equal(tet[d - 1, :16, col], mud.reshape(-1)), (geo, d, col)
                assert np.array_equal(tet[d - 1, 16:, col], mudi.reshape(-1)), (geo, d, col)
    print("tetrads of %d geometry slabs bit-exact" % ctx.ngeo, flush=True)


def test_hllc_fluxes_match_oracle(setup):
    desc, ctx, grid, p0, ps0 = setup
    P, PS = _global_state(desc, grid, p0, ps0)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    Fo, dto = {}, {}
    for d in (1, 2, 3):
        Fo[d], dto[d] = grid.flux(d, P, PS)
        for n in range(ctx.nblk):
            ctx.fluxcalc(d, 0, n)
    bdt = np.array([[ctx.block_dt(n, d) for d in (1, 2, 3)] for n in range(ctx.nblk)])
    ctx.sync()
    for d, sel in ((1, H.HAMR_F1), (2, H.HAMR_F2), (3, H.HAMR_F3)):
        got = ctx.get(sel)
        want = GT.global_to_blocks(desc, Fo[d])
        sl = [slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)]
        sl[d - 1] = slice(NG, NG + desc.bs[d - 1] + 1)
        _check("HLLC F%d" % d, got[:, :, sl[0], sl[1], sl[2]], want[:, :, sl[0], sl[1], sl[2]], tol=FLUX_TOL)
    want_dt = np.zeros_like(bdt)
    dtb = {d: GT.block_view(desc, dto[d]) for d in (1, 2, 3)}
    for n in range(ctx.nblk):
        for d in (1, 2, 3):
            sl = [slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)]
            sl[d - 1] = slice(NG, NG + desc.bs[d - 1] + 1)
            v = dtb[d][n][sl[0], sl[1], sl[2]]
            v = v[(v < 1e9) & (v > 1e-6)]
            want_dt[n, d - 1] = v.min() if v.size else 1e9
    assert np.max(np.abs(bdt - want_dt) / want_dt) <= TOL


def test_hllc_differs_from_hll(setup):
    """The solver switch reaches the kernels: HLL and HLLC fluxes differ."""
    desc, ctx, grid, p0, ps0 = setup
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    for n in range(ctx.nblk):
        ctx.fluxcalc(1, 0, n)
    f_c = ctx.get(H.HAMR_F1)
    ctx.set_riemann(H.HAMR_RIEMANN_HLL)
    try:
        for n in range(ctx.nblk):
            ctx.fluxcalc(1, 0, n)
        f_l = ctx.get(H.HAMR_F1)
    finally:
        ctx.set_riemann(H.HAMR_RIEMANN_HLLC)
    I = slice(NG, -NG)
    assert not np.array_equal(f_c[:, 0, I, I, I], f_l[:, 0, I, I, I])


def test_one_step_hllc(setup):
    desc, ctx, grid, p0, ps0 = setup
    P, PS = _global_state(desc, grid, p0, ps0)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    dt = ctx.estimate_dt()
    F1, d1 = grid.flux(1, P, PS)
    F2, d2 = grid.flux(2, P, PS)
    F3, d3 = grid.flux(3, P, PS)
    dt_o = grid.block_dt(d1, d2, d3)
    assert abs(dt - dt_o) <= TOL * dt_o
    ctx.clear_flags()
    ndt = ctx.step(1, dt)
    ndt_o, pflag_o = grid.step(P, PS, dt)
    assert abs(ndt - ndt_o) <= TOL * ndt_o, (ndt, ndt_o)
    got_p, got_ps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    want_p, want_ps = GT.global_to_blocks(desc, P), GT.global_to_blocks(desc, PS)
    I = slice(NG, -NG)
    _check("P after 1 HLLC step", got_p[:, :, I, I, I], want_p[:, :, I, I, I])
    for d in range(3):
        sl = [slice(NG, NG + desc.bs[k]) for k in range(3)]
        sl[d] = slice(NG, NG + desc.bs[d] + 1)
        _check("B%d after 1 HLLC step" % (d + 1), got_ps[:, d:d + 1, sl[0], sl[1], sl[2]],
               want_ps[:, d:d + 1, sl[0], sl[1], sl[2]])
    pf, fi = ctx.all_flags()
    assert np.array_equal(pf[:, I, I, I], GT.block_view(desc, pflag_o)[:, I, I, I])


def test_per_block_batched_and_graph_agree(setup):
    desc, ctx, grid, p0, ps0 = setup
    outs = []
    for mode in ("blocks", "stream", "graph"):
        ctx.set(H.HAMR_P, p0)
        ctx.set(H.HAMR_PS, ps0)
        dt = ctx.estimate_dt()
        if mode == "blocks":
            ndt = H.advance_step(ctx, dt)
        else:
            ndt = ctx.step(1, dt, graph=(mode == "graph"))
        outs.append((ndt, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)))
    for o in outs[1:]:
        assert o[0] == outs[0][0]
        assert np.array_equal(o[1], outs[0][1]) and np.array_equal(o[2], outs[0][2])


def _mixed_reset(ctx, p0, ps0, mode):
    ctx.set_precision(mode)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)


def _outliers(a, b, tol):
    """Per variable, the number of entries whose error exceeds tol times the
    variable's largest FP64 magnitude."""
    scale = np.abs(b).max(axis=tuple(i for i in range(b.ndim) if i != 1), keepdims=True)
    return (np.abs(a - b) > tol * scale).sum(axis=tuple(i for i in range(b.ndim) if i != 1))


def test_mixed_hllc_fluxes_round_the_fp64_hllc_fluxes(setup):
    """The FP32 HLLC (energy carried with the rest mass taken out, so nothing
    cancels it) against the FP64 HLLC, relative to each variable's largest
    magnitude (the normal field's flux, zero but for roundoff in both, left
    out).  Measured: x1 within 2.6e-4, x3 within 1e-6.  The x2 fluxes of the
    torus at rest are near zero, and the contact speed there is zero or within
    roundoff of it: the solver takes the star state for a nonzero speed and
    HLL for an exact zero (GPU_hllc.h:277, :356), so a face where one
    arithmetic rounds to zero and the other does not switches solver; measured
    1.5% of the x2 faces beyond 1e-5, the L1 error of every x2 flux within
    7e-3.  Bars: x1 5e-4, x3 1e-5, x2 L1 1e-2 with at most 2% of faces beyond
    1e-5.  The mixed HLL fluxes differ from them."""
    desc, ctx, grid, p0, ps0 = setup
    F = {}
    try:
        for mode in (H.HAMR_PREC_FP64, H.HAMR_PREC_MIXED):
            _mixed_reset(ctx, p0, ps0, mode)
            for d in (1, 2, 3):
                for n in range(ctx.nblk):
                    ctx.fluxcalc(d, 0, n)
            F[mode] = [ctx.get(s) for s in (H.HAMR_F1, H.HAMR_F2, H.HAMR_F3)]
        ctx.set_riemann(H.HAMR_RIEMANN_HLL)
        for n in range(ctx.nblk):
            ctx.fluxcalc(1, 0, n)
        f_hll32 = ctx.get(H.HAMR_F1)
    finally:
        ctx.set_riemann(H.HAMR_RIEMANN_HLLC)
        ctx.set_precision(H.HAMR_PREC_FP64)
    for d in range(3):
        sl = [slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)]
        sl[d] = slice(NG, NG + desc.bs[d] + 1)
        a = F[H.HAMR_PREC_MIXED][d][:, :, sl[0], sl[1], sl[2]]
        b = F[H.HAMR_PREC_FP64][d][:, :, sl[0], sl[1], sl[2]]
        err = GT.relerr_by_var(a, b)
        nout = _outliers(a, b, 1e-5)
        nface = a[:, 0].size
        print("HLLC F%d mixed vs FP64: max %s; faces beyond 1e-5 per var %s of %d" %
              (d + 1, np.array2string(err, precision=2), nout, nface), flush=True)
        l1 = np.array([np.abs(a[:, k] - b[:, k]).sum() / max(np.abs(b[:, k]).sum(), 1e-300) for k in range(9)])
        print("   L1 rel err per var %s" % np.array2string(l1, precision=2), flush=True)
        err[5 + d] = 0.0
        l1[5 + d] = 0.0
        nout[5 + d] = 0
        if d == 1:
            assert np.all(l1 <= 1e-2), l1
            assert np.all(nout <= nface // 50), nout
        else:
            assert np.all(err <= (5e-4 if d == 0 else 1e-5)), err
        assert np.array_equal(a, a.astype(np.float32).astype(np.float64))
    I = slice(NG, -NG)
    assert not np.array_equal(F[H.HAMR_PREC_MIXED][0][:, 0, I, I, I], f_hll32[:, 0, I, I, I])


def test_mixed_hllc_steps_close_to_fp64_hllc(setup):
    """The mixed HLLC step against the FP64 HLLC step, relative to each
    variable's largest magnitude.  Measured: one step within 1.1e-4 (the
    x3 velocity in the cells beside the x2 faces that switch solver, above;
    every other variable within 5e-6), the face fields within 1e-5; after
    100 steps the L1 error of rho 1.1e-7, of B1, B2, B3 6.2e-6, 3.2e-6,
    1.12e-5 (the mixed HLL step keeps B within 1e-5; HLLC's B3 is 12% above).
    Bars: one step 2.5e-4, face fields 1e-5, dt 1e-5; after 100 steps L1 of
    rho 1e-6 and of each field component 2e-5.  Per-block calls, the batched
    step and the graph replay of the mixed HLLC step agree bit for bit."""
    desc, ctx, grid, p0, ps0 = setup
    out, out100 = {}, {}
    try:
        for mode in (H.HAMR_PREC_FP64, H.HAMR_PREC_MIXED):
            _mixed_reset(ctx, p0, ps0, mode)
            dt = ctx.estimate_dt()
            ndt = ctx.step(1, dt)
            out[mode] = (dt, ndt, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
            _mixed_reset(ctx, p0, ps0, mode)
            ndt = ctx.step(100, ctx.estimate_dt(), graph=True)
            out100[mode] = (ndt, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))
        d32 = out[H.HAMR_PREC_MIXED][0]
        _mixed_reset(ctx, p0, ps0, H.HAMR_PREC_MIXED)
        ndt_a = H.advance_step(ctx, d32)
        seq = [(ndt_a, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS))]
        for graph in (False, True):
            _mixed_reset(ctx, p0, ps0, H.HAMR_PREC_MIXED)
            seq.append((ctx.step(1, d32, graph=graph), ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)))
    finally:
        ctx.set_precision(H.HAMR_PREC_FP64)
    d64, n64, p64, s64 = out[H.HAMR_PREC_FP64]
    d32, n32, p32, s32 = out[H.HAMR_PREC_MIXED]
    assert abs(d32 - d64) <= 1e-5 * d64 and abs(n32 - n64) <= 1e-5 * n64
    I = slice(NG, -NG)
    a, b = p32[:, :, I, I, I], p64[:, :, I, I, I]
    err = GT.relerr_by_var(a, b)
    nout = _outliers(a, b, 1e-5)
    print("one HLLC step, mixed vs FP64: max rel err per var %s; cells beyond 1e-5 %s of %d" %
          (np.array2string(err, precision=2), nout, a[:, 0].size), flush=True)
    assert np.all(err <= 2.5e-4), err
    for d in range(3):
        e = GT.relerr_by_var(s32[:, d:d + 1, I, I, I], s64[:, d:d + 1, I, I, I])
        assert np.all(e <= 1e-5), (d, e)
    for ndt, pb, psb in seq[1:]:
        assert ndt == seq[0][0]
        assert np.array_equal(pb, seq[0][1]) and np.array_equal(psb, seq[0][2])
    a, b = out100[H.HAMR_PREC_MIXED][1][:, :, I, I, I], out100[H.HAMR_PREC_FP64][1][:, :, I, I, I]
    maxerr = GT.relerr_by_var(a, b)
    l1 = np.array([np.abs(a[:, k] - b[:, k]).sum() / np.abs(b[:, k]).sum() for k in range(9)])
    print("100 HLLC steps, mixed vs FP64: max rel err per var", np.array2string(maxerr, precision=2),
          "L1 rel err per var", np.array2string(l1, precision=2), flush=True)
    assert l1[0] <= 1e-6, l1[0]
    for k in (5, 6, 7):
        assert l1[k] <= 2e-5, (k, l1[k])


def test_multilevel_hllc_step_matches_oracle(oracle):
    desc = H.make_desc((64, 32, 32), (1, 2, 1))
    amr = H.make_amr_desc([((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 128))], seed=5)
    ctx = H.AmrContext(desc, amr)
    try:
        ctx.set_grid()
        ctx.init_torus(H.torus_desc())
        ctx.boundprim(1)
        ctx.set_riemann(H.HAMR_RIEMANN_HLLC)
        g, rows = GT.amr_oracle(oracle, desc, amr, riemann=1)
        p = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_P)]
        ps = [np.ascontiguousarray(a) for a in ctx.get(H.HAMR_PS)]
        flags = [np.ascontiguousarray(ctx.flags(n)[0]) for n in range(ctx.nblk)]
        dt = ctx.estimate_dt()
        ctx.clear_flags()
        ndt = ctx.step(1, dt)
        ndt_o = g.step(p, ps, flags, dt)
        assert abs(ndt - ndt_o) <= TOL * ndt_o, (ndt, ndt_o)
        gp, gps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
        I = (slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
        rel = GT.pointwise_relerr_blocks([a[I] for a in gp], [b[I] for b in p])
        print("multi-level HLLC step: pointwise rel err per var", np.array2string(rel, precision=2), flush=True)
        assert np.all(rel <= TOL)
    finally:
        ctx.close()
