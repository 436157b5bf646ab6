"""The torch restatement of one forward stage (the FP64 reference gradient's
forward) agrees with the oracle's stage on a seeded flat periodic box.

Workload: a 10x8x12 periodic flat box with the adjoint benchmark's random
state (rho lognormal, u = rho lognormal, 3-velocities clipped to Lorentz
factor 5, divergence-free field scaled to max b^2/rho = 10), one stage with
Dt = 0.3 min(dx).  Bar: 1e-12 relative to each variable's largest magnitude.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import hamr_gamma_wind_amd as H
from hamr_gamma_wind_amd import workloads as W
from tests import adjoint_reference as R
from tests import gridtools as GT

N = (10, 8, 12)
GAM = 13.0 / 9.0


def test_torch_stage_matches_oracle_stage(oracle):
    desc = H.make_desc(N, (1, 1, 1), gam=GAM, minkowski=1, rho_floor=0.0, u_floor=0.0)
    grid = GT.oracle_grid(oracle, desc)
    prims, faces = W.minkowski_box(N, gam=GAM, seed=3)
    p, ps = W.to_block(prims, faces)
    P, PS = np.ascontiguousarray(p[0]), np.ascontiguousarray(ps[0])
    grid.bound(P, PS)
    _, dx = H.desc_coords(desc)
    Dt = 0.3 * min(dx)
    Un, _ = grid.stage(Dt, P, PS)
    I = (slice(None),) + tuple(slice(3, 3 + k) for k in N)
    want = Un[I]
    got = R.stage(torch.tensor(prims), torch.tensor(faces), Dt, dx, GAM).numpy()
    err = GT.relerr_by_var(got[None], want[None])
    print("torch stage vs oracle stage, max rel err per var:", err)
    assert np.all(err <= 1e-12), err
