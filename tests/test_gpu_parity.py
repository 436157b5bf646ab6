"""Parity of the HIP step (through the C ABI) with the FP64 oracle.

Workload: the 64x32x32 torus grid (r 1.2-1000 M log-spaced, a = 0.9375,
Gamma = 13/9, FM torus r_in = 6, r_max = 12, seeded +-4% perturbation of u,
floors 1e-5 r^-1.5 and 1e-7 r^-2.5) as 4x4x4 blocks of 16x8x8, compared with
the oracle on the same state as ONE global grid, so that block-boundary ghosts,
duplicated boundary faces and corners, the polar partner exchange and the
ring-wise polar EMF average are all exercised.

Bars: ghost fill and failure flags bit-exact; fluxes, EMFs, face fields and the
state after one FP64 step within 1e-12 of the oracle cell by cell
(|a - b| <= 1e-12 |b|, with the smallest normal double as the floor of |b|);
dt within 1e-12; after ten steps 1e-9 cell by cell.
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
    print("%s: pointwise rel err per var %s (relative to the grid maximum %s)"
          % (name, np.array2string(err, precision=2), np.array2string(GT.relerr_by_var(got, want), precision=2)),
          flush=True)
    assert np.all(err <= tol), (name, err)


@pytest.fixture(scope="module")
def setup(oracle):
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    grid = GT.oracle_grid(oracle, desc)
    p0 = ctx.get(H.HAMR_P)
    ps0 = ctx.get(H.HAMR_PS)
    yield desc, ctx, grid, p0, ps0
    ctx.close()


def _check_faces(name, desc, got, want):
    """Face fields on the faces a block updates: [NG, NG+BS] along the normal,
    interior across."""
    for d in range(3):
        sl = [slice(NG, NG + desc.bs[k]) for k in range(3)]
        sl[d] = slice(NG, NG + desc.bs[d] + 1)
        _check("%s B%d" % (name, d + 1), got[:, d:d + 1, sl[0], sl[1], sl[2]], want[:, d:d + 1, sl[0], sl[1], sl[2]])


def _global_state(desc, grid, p0, ps0):
    P = GT.blocks_to_global(desc, p0)
    PS = GT.blocks_to_global(desc, ps0)
    grid.bound(P, PS)
    return P, PS


def test_initial_state_matches_host_setup(setup):
    desc, ctx, grid, p0, ps0 = setup
    hp, hps = H.desc_torus(desc, H.torus_desc())
    I = slice(NG, -NG)
    assert np.array_equal(p0[:, :, I, I, I], hp[:, :, I, I, I])


def test_ghost_fill_bit_exact(setup):
    """GPU_boundprim: every block's ghosts equal the global grid's cells."""
    desc, ctx, grid, p0, ps0 = setup
    P, PS = _global_state(desc, grid, p0, ps0)
    want_p = GT.global_to_blocks(desc, P)
    want_ps = GT.global_to_blocks(desc, PS)
    assert np.array_equal(p0, want_p)
    # face fields: B^2 ghosts at the axis and B^1 ghosts at the radial ends are
    # never written (as in the reference); compare the rest
    for d in range(3):
        a, b = ps0[:, d], want_ps[:, d]
        if d == 1:
            for bi, (c1, c2, c3) in enumerate(GT.block_coords(desc)):
                if c2 == 0:
                    a[bi, :, :NG] = b[bi, :, :NG] = 0
                if c2 == desc.nb[1] - 1:
                    a[bi, :, -NG:] = b[bi, :, -NG:] = 0
        if d == 0:
            for bi, (c1, c2, c3) in enumerate(GT.block_coords(desc)):
                if c1 == 0:
                    a[bi, :NG] = b[bi, :NG] = 0
                if c1 == desc.nb[0] - 1:
                    a[bi, -NG:] = b[bi, -NG:] = 0
        assert np.array_equal(a, b), d


def test_fluxes_emf_and_ct(setup):
    desc, ctx, grid, p0, ps0 = setup
    P, PS = _global_state(desc, grid, p0, ps0)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    Fo = {}
    dto = {}
    for d in (1, 2, 3):
        Fo[d], dto[d] = grid.flux(d, P, PS)
        for n in range(ctx.nblk):
            ctx.fluxcalc(d, 0, n)
    nb = ctx.nblk
    bdt = np.array([[ctx.block_dt(n, d) for d in (1, 2, 3)] for n in range(nb)])
    ctx.sync()
    for d, sel in ((1, H.HAMR_F1), (2, H.HAMR_F2), (3, H.HAMR_F3)):
        got = ctx.get(sel)
        want = GT.global_to_blocks(desc, Fo[d])
        # compare the faces each block computes: [NG, NG+BS] along d, [NG-1, NG+BS] across
        sl = [slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)]
        sl[d - 1] = slice(NG, NG + desc.bs[d - 1] + 1)
        _check("F%d" % d, got[:, :, sl[0], sl[1], sl[2]], want[:, :, sl[0], sl[1], sl[2]], tol=FLUX_TOL)
    # per-block Courant minima
    want_dt = np.zeros_like(bdt)
    dtb = {d: GT.block_view(desc, dto[d]) for d in (1, 2, 3)}
    for n in range(nb):
        for d in (1, 2, 3):
            sl = [slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)]
            sl[d - 1] = slice(NG, NG + desc.bs[d - 1] + 1)
            v = dtb[d][n][sl[0], sl[1], sl[2]]
            v = v[(v < 1e9) & (v > 1e-6)]
            want_dt[n, d - 1] = v.min() if v.size else 1e9
    assert np.max(np.abs(bdt - want_dt) / want_dt) <= TOL
    # EMF (with polar average) and the face-B update
    dt = 0.05
    Eo = grid.emf(P, Fo[1], Fo[2], Fo[3], dt)
    for n in range(nb):
        ctx.ct_emf(0, n)
    ctx.ct_bound(dt)
    got = ctx.get(H.HAMR_EMF)
    want = GT.global_to_blocks(desc, Eo)
    sl = [slice(NG, NG + desc.bs[k] + 1) for k in range(3)]
    _check("EMF", got[:, :, sl[0], sl[1], sl[2]], want[:, :, sl[0], sl[1], sl[2]])
    PSH = PS.copy()
    grid.ct_update(0.5 * dt, PS, PSH, Eo)
    for n in range(nb):
        ctx.ct_update(0, 0.5 * dt, n)
    _check_faces("face", desc, ctx.get(H.HAMR_PSH), GT.global_to_blocks(desc, PSH))


def test_one_step_fp64(setup):
    """Predictor + corrector from the same state: P, face B, flags, dt."""
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
    got_p = ctx.get(H.HAMR_P)
    got_ps = ctx.get(H.HAMR_PS)
    want_p = GT.global_to_blocks(desc, P)
    want_ps = GT.global_to_blocks(desc, PS)
    I = slice(NG, -NG)
    _check("P after 1 step", got_p[:, :, I, I, I], want_p[:, :, I, I, I])
    _check_faces("face after 1 step", desc, got_ps, want_ps)
    pf, fi = ctx.all_flags()
    want_flag = GT.block_view(desc, pflag_o)
    assert np.array_equal(pf[:, I, I, I], want_flag[:, I, I, I])
    print("flags: codes", np.unique(pf[:, I, I, I], return_counts=True), "failimage", fi.sum(axis=0), flush=True)


def test_per_block_sequence_equals_batched_step(setup):
    """advance_step (per-block C-ABI calls in advance_GPU order) == hamr_step."""
    desc, ctx, grid, p0, ps0 = setup
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    dt = ctx.estimate_dt()
    ndt_a = H.advance_step(ctx, dt)
    pa, psa = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    ndt_b = ctx.step(1, dt)
    pb, psb = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    assert ndt_a == ndt_b
    assert np.array_equal(pa, pb) and np.array_equal(psa, psb)
    # the same sequence driven by the C host (libhamr_host.so)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    ndt_c = H.host_advance(ctx, dt)
    assert ndt_c == ndt_b
    assert np.array_equal(ctx.get(H.HAMR_P), pb) and np.array_equal(ctx.get(H.HAMR_PS), psb)


def test_graph_replay_equals_stream(setup):
    desc, ctx, grid, p0, ps0 = setup
    res = []
    for graph in (False, True):
        ctx.set(H.HAMR_P, p0)
        ctx.set(H.HAMR_PS, ps0)
        dt = ctx.estimate_dt()
        outs = []
        for k in range(3):
            dt = ctx.step(1, dt, graph=graph)
            outs.append((dt, ctx.get(H.HAMR_P)))
        res.append(outs)
    for (da, pa), (db, pb) in zip(res[0], res[1]):
        assert da == db and np.array_equal(pa, pb)


def test_ten_steps_track_oracle(setup):
    desc, ctx, grid, p0, ps0 = setup
    P, PS = _global_state(desc, grid, p0, ps0)
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    dt = ctx.estimate_dt()
    dto = dt
    for k in range(10):
        dt = ctx.step(1, dt)
        dto, _ = grid.step(P, PS, dto)
    want = GT.global_to_blocks(desc, P)
    I = slice(NG, -NG)
    got = ctx.get(H.HAMR_P)
    _check("10 steps", got[:, :, I, I, I], want[:, :, I, I, I], tol=1e-9)
    print("dt", dt, dto, flush=True)
    assert abs(dt - dto) <= 1e-9 * dto


def test_aos_staging_round_trip(setup):
    """GPU_write / GPU_read layout: p[cell][NPR], ps[cell][NDIM] (B^1..B^3 in
    components 1..3) in index_3D order, ghosts included <-> the device's
    per-block component slabs, bit-exact."""
    desc, ctx, grid, p0, ps0 = setup
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)
    for n in (0, ctx.nblk // 2, ctx.nblk - 1):
        pa, psa = ctx.download(n)
        assert np.array_equal(pa, np.moveaxis(p0[n], 0, -1))
        assert np.array_equal(psa[..., 1:], np.moveaxis(ps0[n], 0, -1)) and np.all(psa[..., 0] == 0)
        rng = np.random.default_rng(n)
        qa = rng.normal(size=pa.shape)
        qsa = rng.normal(size=psa.shape)
        ctx.upload(n, qa, qsa)
        assert np.array_equal(ctx.get(H.HAMR_P)[n], np.moveaxis(qa, -1, 0))
        assert np.array_equal(ctx.get(H.HAMR_PS)[n], np.moveaxis(qsa[..., 1:], -1, 0))
        back = ctx.download(n)
        assert np.array_equal(back[0], qa) and np.array_equal(back[1][..., 1:], qsa[..., 1:])
    ctx.set(H.HAMR_P, p0)
    ctx.set(H.HAMR_PS, ps0)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_marching_update_equals_per_cell_update(setup, mode):
    """The x1-marching form of the conserved update (HAMR_TUNE_CONS with
    workgroups -1) gives the per-cell form's step bit for bit."""
    desc, ctx, grid, p0, ps0 = setup
    slot = H.HAMR_TUNE_NAMES.index("cons")
    ctx.set_precision(H.HAMR_PREC_FP64 if mode == "fp64" else H.HAMR_PREC_MIXED)
    try:
        outs = []
        for cfg in ((256, 0, 0), (128, -1, 0)):
            ctx.tune_set(slot, *cfg)
            ctx.set(H.HAMR_P, p0)
            ctx.set(H.HAMR_PS, ps0)
            dt = ctx.estimate_dt()
            ndt = ctx.step(2, dt)
            outs.append((ndt, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)))
        assert outs[0][0] == outs[1][0]
        assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])
    finally:
        ctx.tune_set(slot, 256, 0, 0)
        ctx.set_precision(H.HAMR_PREC_FP64)


def _compare_flux_forms(desc, ctx, p0, ps0, mode):
    ctx.set_precision(H.HAMR_PREC_MIXED if mode == "mixed" else H.HAMR_PREC_FP64)
    if mode == "hllc":
        ctx.set_riemann(H.HAMR_RIEMANN_HLLC)
    slots = [H.HAMR_TUNE_NAMES.index(f) for f in ("flux1", "flux2", "flux3")]
    faces = [tuple([slice(None), slice(None)] + [slice(NG, NG + desc.bs[k] + 1) if k == d else
                                                 slice(NG - 1, NG + desc.bs[k] + 1) for k in range(3)])
             for d in range(3)]
    try:
        outs = []
        for line in (0, 1):
            for slot in slots:
                ctx.tune_set(slot, 256, 0, 0, line)
            ctx.set(H.HAMR_P, p0)
            ctx.set(H.HAMR_PS, ps0)
            dt = ctx.estimate_dt()
            fl = [ctx.get(a)[faces[d]] for d, a in enumerate((H.HAMR_F1, H.HAMR_F2, H.HAMR_F3))]
            ndt = ctx.step(2, dt)
            outs.append((dt, ndt, fl, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)))
        (a, b) = outs
        assert a[0] == b[0] and a[1] == b[1], (a[:2], b[:2])
        for x, y in zip(a[2], b[2]):
            assert np.array_equal(x, y)
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    finally:
        for slot in slots:
            ctx.tune_set(slot, 256, 0, 0)
        ctx.set_riemann(H.HAMR_RIEMANN_HLL)
        ctx.set_precision(H.HAMR_PREC_FP64)


@pytest.mark.parametrize("mode", ["fp64", "mixed", "hllc"])
def test_line_staged_fluxes_equal_per_face_fluxes(setup, mode):
    """The flux kernels' two forms, one thread per face (the default) and
    line-staged tiles (each cell's slope and limited parabola formed once per
    tile), give the same fluxes over every face they write, the same dt and
    the same state after two steps, bit for bit (16 x 8 x 8 blocks: 17 faces
    per line along x1, 9 along x2 and x3)."""
    desc, ctx, grid, p0, ps0 = setup
    _compare_flux_forms(desc, ctx, p0, ps0, mode)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_line_staged_fluxes_equal_per_face_fluxes_32(mode):
    """The same on 32^3 blocks (33 faces per line, the headline grid's tiles)."""
    desc = H.make_desc((32, 32, 32), (2, 2, 2))
    ctx = H.Context(desc)
    try:
        ctx.set_grid()
        ctx.init_torus(H.torus_desc())
        ctx.boundprim(1)
        _compare_flux_forms(desc, ctx, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS), mode)
    finally:
        ctx.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_fused_update_equals_two_kernels(setup, mode):
    """The update's one-kernel form (k_update, the default: conserved update and
    2D inversion with U' in registers) and its two-kernel form (k_cons, then
    k_invert through Unew) give the same steps, flags and failure counts bit
    for bit."""
    desc, ctx, grid, p0, ps0 = setup
    slot = H.HAMR_TUNE_NAMES.index("cons")
    ctx.set_precision(H.HAMR_PREC_FP64 if mode == "fp64" else H.HAMR_PREC_MIXED)
    try:
        outs = []
        for second in (0, 1):
            ctx.tune_set(slot, 256, 0, 0, second)
            ctx.set(H.HAMR_P, p0)
            ctx.set(H.HAMR_PS, ps0)
            ctx.clear_flags()
            dt = ctx.estimate_dt()
            ndt = ctx.step(2, dt)
            outs.append((ndt, ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)) + tuple(ctx.all_flags()))
        a, b = outs
        assert a[0] == b[0]
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y)
    finally:
        ctx.tune_set(slot, 256, 0, 0)
        ctx.set_precision(H.HAMR_PREC_FP64)
