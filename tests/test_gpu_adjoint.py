"""Adjoint of one forward stage (hand-written reverse kernels, FP64) in a flat
periodic box, through the C ABI and through torch.autograd.

Workload: the adjoint benchmark's random state (rho lognormal(0, 0.5),
u = rho lognormal(-1, 0.3), 3-velocity N(0,1) clipped to Lorentz factor 5,
divergence-free field from a seeded vector potential scaled to max
b^2/rho = 10), loss L = sum w . U' with seeded N(0,1) weights, Dt = 0.3 min(dx).

Bars:
  - the GPU forward stage equals the oracle's to 1e-12 (relative to each
    variable's largest magnitude);
  - dL/dP and dL/d(face field) equal the FP64 reference gradient (torch
    autograd of the stage restated in tests/adjoint_reference.py, pinned to
    the oracle) to 1e-9 relative to each component's largest magnitude,
    on a 12x10x8 box;
  - on the 128^3 box, 64 random probes of dL/dP agree with central
    differences (h = 1e-6 of the larger of the probed value and the rms of its
    variable) to 1e-5 relative, plus the rounding bound of the difference
    quotient itself, 2^-52 sum |w| (|U'(+h)| + |U'(-h)|) / 2h over the cells
    the probe changes (a probe whose derivative is small next to the U' it
    moves is limited by that rounding, not by the adjoint);
  - torch.autograd through hamr_gamma_wind_amd.autograd gives the gradient of
    the direct C-ABI adjoint call (to 1e-13: the adjoint's accumulation order
    is that of its atomic adds).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import hamr_gamma_wind_amd as H
from hamr_gamma_wind_amd import workloads as W
from tests import adjoint_reference as R
from tests import gridtools as GT

pytestmark = pytest.mark.gpu

GAM = 13.0 / 9.0
NG = 3


def _box(n, seed):
    desc = H.make_desc(n, (1, 1, 1), gam=GAM, minkowski=1, rho_floor=0.0, u_floor=0.0)
    ctx = H.Context(desc)
    ctx.set_grid()
    prims, faces = W.minkowski_box(n, gam=GAM, seed=seed)
    p, ps = W.to_block(prims, faces)
    ctx.set(H.HAMR_P, p)
    ctx.set(H.HAMR_PS, ps)
    _, dx = H.desc_coords(desc)
    return desc, ctx, prims, faces, p, ps, dx


def _interior(n):
    return tuple(slice(NG, NG + k) for k in n)


@pytest.fixture(scope="module")
def small():
    n = (12, 10, 8)
    desc, ctx, prims, faces, p, ps, dx = _box(n, seed=5)
    yield n, desc, ctx, prims, faces, p, ps, dx
    ctx.close()


def test_stage_forward_matches_oracle(small, oracle):
    n, desc, ctx, prims, faces, p, ps, dx = small
    Dt = 0.3 * min(dx)
    ctx.set(H.HAMR_P, p)
    ctx.set(H.HAMR_PS, ps)
    ctx.stage_forward(Dt)
    got = ctx.get(H.HAMR_UNEW)[0]
    grid = GT.oracle_grid(oracle, desc)
    P, PS = np.ascontiguousarray(p[0]), np.ascontiguousarray(ps[0])
    grid.bound(P, PS)
    want, _ = grid.stage(Dt, P, PS)
    I = (slice(None),) + _interior(n)
    err = GT.relerr_by_var(got[I][None], want[I][None])
    print("stage forward vs oracle:", np.array2string(err, precision=2), flush=True)
    assert np.all(err <= 1e-12), err


def test_adjoint_matches_reference_gradient(small):
    n, desc, ctx, prims, faces, p, ps, dx = small
    Dt = 0.3 * min(dx)
    rng = np.random.default_rng(11)
    w = rng.normal(size=(1, 9) + tuple(k + 2 * NG for k in n))
    ctx.set(H.HAMR_P, p)
    ctx.set(H.HAMR_PS, ps)
    gp, gps = ctx.stage_adjoint(Dt, w)
    I = (slice(None),) + _interior(n)
    Pt = torch.tensor(prims, requires_grad=True)
    PSt = torch.tensor(faces, requires_grad=True)
    U = R.stage(Pt, PSt, Dt, dx, GAM)
    (U * torch.tensor(w[0][I])).sum().backward()
    eP = GT.relerr_by_var(gp[0][I][None], Pt.grad.numpy()[None])
    eS = GT.relerr_by_var(gps[0][I][None], PSt.grad.numpy()[None])
    print("dL/dP vs reference:", np.array2string(eP, precision=2), "dL/dB_face:", np.array2string(eS, precision=2),
          flush=True)
    assert np.all(eP[:8] <= 1e-9), eP
    assert np.abs(gp[0][8]).max() == 0.0  # the KTOT primitive does not enter the stage
    assert np.all(eS <= 1e-9), eS
    # ghosts were folded onto the interior
    mask = np.ones(gp.shape[2:], bool)
    mask[_interior(n)] = False
    assert np.abs(gp[0][:, mask]).max() == 0.0


def test_autograd_function_gives_the_adjoint(small):
    from hamr_gamma_wind_amd.autograd import stage

    n, desc, ctx, prims, faces, p, ps, dx = small
    Dt = 0.3 * min(dx)
    rng = np.random.default_rng(12)
    w = rng.normal(size=(1, 9) + tuple(k + 2 * NG for k in n))
    ctx.set(H.HAMR_P, p)
    ctx.set(H.HAMR_PS, ps)
    gp, gps = ctx.stage_adjoint(Dt, w)
    dev = torch.device("cuda:0")
    Pt = torch.tensor(p, device=dev, requires_grad=True)
    PSt = torch.tensor(ps, device=dev, requires_grad=True)
    U = stage(Pt, PSt, ctx, Dt)
    L = (U * torch.tensor(w, device=dev)).sum()
    L.backward()
    # the adjoint accumulates with atomic adds: equal up to the order of the sums
    assert np.allclose(Pt.grad.cpu().numpy(), gp, rtol=1e-13, atol=1e-13 * np.abs(gp).max())
    assert np.allclose(PSt.grad.cpu().numpy(), gps, rtol=1e-13, atol=1e-13 * np.abs(gps).max())


def test_adjoint_central_differences_128():
    from hamr_gamma_wind_amd.autograd import stage

    n = (128, 128, 128)
    desc, ctx, prims, faces, p, ps, dx = _box(n, seed=1)
    try:
        Dt = 0.3 * min(dx)
        dev = torch.device("cuda:0")
        rng = np.random.default_rng(2)
        D = tuple(k + 2 * NG for k in n)
        w = torch.tensor(rng.normal(size=(1, 9) + D), device=dev)
        Pt = torch.tensor(p, device=dev, requires_grad=True)
        PSt = torch.tensor(ps, device=dev)
        L = (w * stage(Pt, PSt, ctx, Dt)).sum()
        (gP,) = torch.autograd.grad(L, [Pt])
        gP = gP.cpu().numpy()[0]
        bad = []
        worst = 0.0
        for probe in range(64):
            k = int(rng.integers(0, 8))
            c = tuple(int(rng.integers(NG, NG + m)) for m in n)
            x = float(p[(0, k) + c])
            h = 1e-6 * max(abs(x), float(np.sqrt(np.mean(prims[k] ** 2))))
            Us = []
            for sgn in (1.0, -1.0):
                Pp = Pt.detach().clone()
                Pp[(0, k) + c] += sgn * h
                with torch.no_grad():
                    Us.append(stage(Pp, PSt, ctx, Dt))
            # U' changes only near the probe: difference first, so that the sum
            # over the box adds exact zeros elsewhere
            fd = float((w * (Us[0] - Us[1])).sum()) / (2 * h)
            moved = Us[0] != Us[1]
            noise = 2.0**-52 * float((w.abs() * (Us[0].abs() + Us[1].abs()))[moved].sum()) / (2 * h)
            ad = float(gP[(k,) + c])
            scale = max(abs(ad), abs(fd), 1e-8 * float(np.abs(gP[k]).max()))
            ratio = abs(fd - ad) / (1e-5 * scale + noise)
            worst = max(worst, ratio)
            if ratio > 1.0:
                bad.append((k, c, ad, fd, abs(fd - ad) / scale, noise))
        print("128^3 central differences: 64 probes, worst |fd - ad| / (1e-5 scale + rounding) = %.3g" % worst, bad,
              flush=True)
        assert not bad, bad
    finally:
        ctx.close()
