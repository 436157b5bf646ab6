"""hamr_stage_adjoint_dev_noalloc (the adjoint with its work arrays in the
library's static device scratch, no allocation) against hamr_stage_adjoint_dev
(work arrays the context allocates), through the C ABI.

Two noalloc calls in a row: a 16x12x10 flat periodic box, then a smaller
12x10x8 box in another context, which finds the scratch as the first call
left it; each is compared with the allocating entry point on the same inputs.

Workload: the adjoint benchmark's random state (rho lognormal(0, 0.5),
u = rho lognormal(-1, 0.3), 3-velocity N(0,1) clipped to Lorentz factor 5,
divergence-free field from a seeded vector potential scaled to max
b^2/rho = 10), weights w ~ N(0,1), Dt = 0.3 min(dx).

Bar: dL/dP and dL/d(face field) equal the allocating entry point's to 1e-13
relative, with 1e-13 of the largest magnitude as the floor (the adjoint
accumulates with atomic adds, so two runs agree up to the order of the sums),
and no NaN is left from the sentinel the outputs start with.  A box whose work
arrays exceed the scratch is refused with an error that names the limit.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import hamr_gamma_wind_amd as H
from hamr_gamma_wind_amd import workloads as W

pytestmark = pytest.mark.gpu

GAM = 13.0 / 9.0
NG = 3


def _box(n, seed):
    desc = H.make_desc(n, (1, 1, 1), gam=GAM, minkowski=1, rho_floor=0.0, u_floor=0.0)
    ctx = H.Context(desc)
    ctx.set_grid()
    prims, faces = W.minkowski_box(n, gam=GAM, seed=seed)
    p, ps = W.to_block(prims, faces)
    _, dx = H.desc_coords(desc)
    return ctx, p, ps, dx


def _call(ctx, fn, Dt, w, p, ps):
    """One adjoint call through the C ABI on device buffers; the state is
    reset first (the stage fills the ghosts of p in place)."""
    ctx.set(H.HAMR_P, p)
    ctx.set(H.HAMR_PS, ps)
    gp = torch.full(p.shape, np.nan, dtype=torch.float64, device="cuda:0")
    gps = torch.full(ps.shape, np.nan, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = getattr(ctx.L, fn)(ctx.h, C.c_double(Dt), C.c_void_p(w.data_ptr()), C.c_void_p(gp.data_ptr()),
                            C.c_void_p(gps.data_ptr()))
    assert rc == 0, (fn, ctx.L.hamr_last_error(ctx.h))
    torch.cuda.synchronize()
    return gp.cpu().numpy(), gps.cpu().numpy()


def _close(a, b):
    return bool(np.allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max()))


def test_noalloc_adjoint_equals_allocating_adjoint_twice():
    for n, seed in (((16, 12, 10), 21), ((12, 10, 8), 22)):
        ctx, p, ps, dx = _box(n, seed)
        try:
            Dt = 0.3 * min(dx)
            rng = np.random.default_rng(seed + 100)
            w = torch.tensor(rng.normal(size=p.shape), device="cuda:0")
            print("box %s: noalloc adjoint" % (n,), flush=True)
            gp0, gps0 = _call(ctx, "hamr_stage_adjoint_dev_noalloc", Dt, w, p, ps)
            print("box %s: allocating adjoint" % (n,), flush=True)
            gp1, gps1 = _call(ctx, "hamr_stage_adjoint_dev", Dt, w, p, ps)
            assert np.all(np.isfinite(gp0)) and np.all(np.isfinite(gps0))
            d = float(np.abs(gp0 - gp1).max() / np.abs(gp1).max())
            ds = float(np.abs(gps0 - gps1).max() / np.abs(gps1).max())
            print("box %s: max |noalloc - alloc| / max: dL/dP %.2e, dL/dB_face %.2e" % (n, d, ds), flush=True)
            assert _close(gp0, gp1) and _close(gps0, gps1)
        finally:
            ctx.close()


def test_noalloc_refuses_boxes_above_the_scratch():
    # 72^3 interior: 78^3 ghost-inclusive cells need 57 * 8 B each, above 128 MiB
    n = (72, 72, 72)
    desc = H.make_desc(n, (1, 1, 1), gam=GAM, minkowski=1, rho_floor=0.0, u_floor=0.0)
    ctx = H.Context(desc)
    try:
        D = tuple(k + 2 * NG for k in n)
        w = torch.zeros((1, 9) + D, dtype=torch.float64, device="cuda:0")
        gp = torch.zeros_like(w)
        gps = torch.zeros((1, 3) + D, dtype=torch.float64, device="cuda:0")
        rc = ctx.L.hamr_stage_adjoint_dev_noalloc(ctx.h, C.c_double(1e-3), C.c_void_p(w.data_ptr()),
                                                  C.c_void_p(gp.data_ptr()), C.c_void_p(gps.data_ptr()))
        assert rc != 0
        msg = ctx.L.hamr_last_error(ctx.h).decode()
        print("refused:", msg, flush=True)
        assert "scratch" in msg and str(128 << 20) in msg
    finally:
        ctx.close()
