"""The FP64 step on the 256x128x128 torus as 8x4x4 blocks of 32^3 (4.2 M
zones, 3 ghost zones per block) against the oracle on one global grid.

Workload: r 1.2-1000 M log-spaced, Kerr a = 0.9375, Gamma = 13/9, FM torus
r_in = 6, r_max = 12, poloidal loop normalised to min beta = 100, seeded +-4%
perturbation of u, floors 1e-5 r^-1.5 and 1e-7 r^-2.5, Courant 0.8.

Bars: the ghost fill bit-exact; one step from the same state: every primitive
and face field within 1e-12 of the oracle cell by cell (|a - b| <= 1e-12 |b|,
the smallest normal double as the floor of |b|), flags bit-exact, next dt
within 1e-12.  Then 100 steps, both sides advancing with the oracle's dt
sequence (the device's own next dt is checked against it at every step, 1e-12):
the relative L1 error of rho (sum |d| / sum |rho| over the interior) is printed
at steps 25, 50, 75 and 100 and must stay <= 1e-12; the state stays within
1e-9 of the oracle cell by cell.  The steps run in four tests of 25 so that
each one ends within about a minute on the host cores.
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from hamr_gamma_wind_amd import workloads as W
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

NG = 3
I = slice(NG, -NG)


@pytest.fixture(scope="module")
def run(oracle):
    bs, nb = W.TORUS_256
    desc = H.make_desc(bs, nb)
    ctx = H.Context(desc)
    ctx.set_grid()
    ctx.init_torus(H.torus_desc())
    ctx.boundprim(1)
    grid = GT.oracle_grid(oracle, desc)
    p0, ps0 = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    P = GT.blocks_to_global(desc, p0)
    PS = GT.blocks_to_global(desc, ps0)
    grid.bound(P, PS)
    st = dict(desc=desc, ctx=ctx, grid=grid, P=P, PS=PS, p0=p0, ps0=ps0, step=0, dt=ctx.estimate_dt())
    yield st
    ctx.close()


def test_ghost_fill_bit_exact(run):
    desc, p0, ps0 = run["desc"], run["p0"], run["ps0"]
    assert np.array_equal(p0, GT.global_to_blocks(desc, run["P"]))
    want = GT.global_to_blocks(desc, run["PS"])
    # B^1 ghosts at the radial ends and B^2 ghosts beyond the axis are never
    # written (as in the reference); compare the rest
    for d in range(3):
        a, b = ps0[:, d].copy(), want[:, d].copy()
        for bi, (c1, c2, c3) in enumerate(GT.block_coords(desc)):
            if d == 0 and c1 == 0:
                a[bi, :NG] = b[bi, :NG] = 0
            if d == 0 and c1 == desc.nb[0] - 1:
                a[bi, -NG:] = b[bi, -NG:] = 0
            if d == 1 and c2 == 0:
                a[bi, :, :NG] = b[bi, :, :NG] = 0
            if d == 1 and c2 == desc.nb[1] - 1:
                a[bi, :, -NG:] = b[bi, :, -NG:] = 0
        assert np.array_equal(a, b), d


def _faces(desc, ps):
    """Each block's own faces: [NG, NG+BS] along the normal, interior across."""
    out = []
    for d in range(3):
        sl = [slice(NG, NG + desc.bs[k]) for k in range(3)]
        sl[d] = slice(NG, NG + desc.bs[d] + 1)
        out.append(ps[:, d:d + 1, sl[0], sl[1], sl[2]])
    return out


def test_one_step_fp64(run):
    desc, ctx, grid = run["desc"], run["ctx"], run["grid"]
    assert run["step"] == 0
    dt = run["dt"]
    ctx.clear_flags()
    ndt = ctx.step(1, dt)
    ndt_o, pflag_o = grid.step(run["P"], run["PS"], dt)
    run["step"], run["dt"] = 1, ndt_o
    print("dt %.17g -> %.17g (oracle %.17g)" % (dt, ndt, ndt_o), flush=True)
    assert abs(ndt - ndt_o) <= 1e-12 * ndt_o
    got_p, got_ps = ctx.get(H.HAMR_P), ctx.get(H.HAMR_PS)
    want_p = GT.global_to_blocks(run["desc"], run["P"])
    want_ps = GT.global_to_blocks(run["desc"], run["PS"])
    err = GT.pointwise_relerr(got_p[:, :, I, I, I], want_p[:, :, I, I, I])
    print("P after 1 step: pointwise rel err per var", np.array2string(err, precision=2), flush=True)
    assert np.all(err <= 1e-12), err
    for d, (a, b) in enumerate(zip(_faces(desc, got_ps), _faces(desc, want_ps))):
        e = GT.pointwise_relerr(a, b)
        print("face B%d after 1 step: pointwise rel err %.2e" % (d + 1, e[0]), flush=True)
        assert e[0] <= 1e-12
    pf, fi = ctx.all_flags()
    assert np.array_equal(pf[:, I, I, I], GT.block_view(desc, pflag_o)[:, I, I, I])
    print("flags:", dict(zip(*[x.tolist() for x in np.unique(pf[:, I, I, I], return_counts=True)])),
          "failimage", fi.sum(axis=0).tolist(), flush=True)


@pytest.mark.parametrize("upto", [25, 50, 75, 100])
def test_hundred_steps_rho_l1(run, upto):
    desc, ctx, grid = run["desc"], run["ctx"], run["grid"]
    while run["step"] < upto:
        dt = run["dt"]
        ndt = ctx.step(1, dt)
        ndt_o, _ = grid.step(run["P"], run["PS"], dt)
        assert abs(ndt - ndt_o) <= 1e-12 * ndt_o, (run["step"], ndt, ndt_o)
        run["step"], run["dt"] = run["step"] + 1, ndt_o
    got = GT.blocks_to_global(desc, ctx.get(H.HAMR_P))
    P = run["P"]
    J = (slice(NG, -NG),) * 3
    l1 = {}
    for name, k in (("rho", 0), ("u", 1), ("B1", 5), ("B2", 6), ("B3", 7)):
        l1[name] = float(np.abs(got[k][J] - P[k][J]).sum() / np.abs(P[k][J]).sum())
    pw = GT.pointwise_relerr(got[None][:, :, NG:-NG, NG:-NG, NG:-NG], P[None][:, :, NG:-NG, NG:-NG, NG:-NG])
    print("step %d: L1 %s; pointwise max per var %s" % (run["step"], {k: "%.2e" % v for k, v in l1.items()},
                                                       np.array2string(pw, precision=2)), flush=True)
    assert l1["rho"] <= 1e-12
    assert np.all(pw <= 1e-9)
