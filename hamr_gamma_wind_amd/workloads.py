"""Synthetic initial states of the benchmark workloads (host side, seeded).

minkowski_box: the periodic flat box of the adjoint workload -- rho
lognormal(0, 0.5), u = rho lognormal(-1, 0.3), 3-velocity with N(0,1)
components clipped to Lorentz factor <= 5, a divergence-free face field curled
from a seeded random vector potential and scaled so that max b^2/rho = 10.
"""
import numpy as np

NG = 3

# The torus grids: (cells per block, blocks per direction).  r 1.2-1000 M
# log-spaced, theta over (0, pi), phi over 2 pi; NG = 3 ghost zones per block.
TORUS_64 = ((16, 8, 8), (4, 4, 4))        # 64x32x32 as 64 blocks of 16x8x8
TORUS_256 = ((32, 32, 32), (8, 4, 4))     # 256x128x128 as 128 blocks of 32^3 (4.2 M zones)
TORUS_384 = ((32, 32, 32), (12, 6, 6))    # 384x192x192 as 432 blocks of 32^3 (14.2 M zones)

# The multi-level torus: base 128x96x32 (64x32x32 superblocks); level 1 over
# x1 [0,160) x2 [32,160) of its 256x192x64 cells (r <= 80 M, 30-150 deg), level 2
# over x1 [0,256) x2 [128,256) of its 512x384x128 cells (r <= 34.6 M, 60-120 deg);
# blocks of 64x32x32, 32x16x16 and 16^3 by the seeded choice (722 blocks,
# 5.21 M zones).
AMR_TORUS = dict(bs=(64, 32, 32), nb=(2, 3, 1),
                 boxes=[((0, 160), (32, 160), (0, 64)), ((0, 256), (128, 256), (0, 128))], seed=2)


def curl_faces(A, dx):
    """Face fields (lower faces) B = curl A on a periodic grid with edge-centred
    A; their discrete divergence vanishes identically."""
    A1, A2, A3 = A
    r = np.roll
    b1 = (r(A3, -1, 1) - A3) / dx[1] - (r(A2, -1, 2) - A2) / dx[2]
    b2 = (r(A1, -1, 2) - A1) / dx[2] - (r(A3, -1, 0) - A3) / dx[0]
    b3 = (r(A2, -1, 0) - A2) / dx[0] - (r(A1, -1, 1) - A1) / dx[1]
    return np.array([b1, b2, b3])


def minkowski_box(n, gam=13.0 / 9.0, seed=1, b2rho_max=10.0, gamma_max=5.0):
    """Interior primitives (9, n1, n2, n3) and lower-face fields (3, n1, n2, n3)."""
    n = tuple(n)
    rng = np.random.default_rng(seed)
    dx = tuple(1.0 / k for k in n)
    rho = rng.lognormal(0.0, 0.5, n)
    u = rho * rng.lognormal(-1.0, 0.3, n)
    v = rng.normal(size=(3,) + n)
    vmag = np.sqrt((v * v).sum(axis=0))
    vmax = np.sqrt(1.0 - 1.0 / gamma_max ** 2)
    v *= np.where(vmag > vmax, vmax / np.maximum(vmag, 1e-300), 1.0)
    lor = 1.0 / np.sqrt(1.0 - (v * v).sum(axis=0))
    U = v * lor
    A = rng.normal(size=(3,) + n) * min(dx)
    faces = curl_faces(A, dx)
    Bc = np.array([0.5 * (faces[d] + np.roll(faces[d], -1, d)) for d in range(3)])
    # b^2 = (B^2 + (B.U)^2) / gamma^2 in flat space with U the relative 4-velocity
    BU = (Bc * U).sum(axis=0)
    b2 = ((Bc * Bc).sum(axis=0) + BU * BU) / (lor * lor)
    scale = np.sqrt(b2rho_max / np.max(b2 / rho))
    faces *= scale
    Bc *= scale
    prims = np.zeros((9,) + n)
    prims[0], prims[1], prims[2:5], prims[5:8] = rho, u, U, Bc
    prims[8] = (gam - 1.0) * u * rho ** (-gam)
    return prims, faces


def to_block(prims, faces):
    """A single block's ghost-inclusive arrays (1, ncomp, n+6, ...) with the
    interior set (ghosts zero; a ghost fill completes them)."""
    n = prims.shape[1:]
    D = tuple(k + 2 * NG for k in n)
    p = np.zeros((1, 9) + D)
    ps = np.zeros((1, 3) + D)
    I = tuple(slice(NG, NG + k) for k in n)
    p[(0, slice(None)) + I] = prims
    ps[(0, slice(None)) + I] = faces
    return p, ps
