"""Synthetic initial states of the benchmark workloads (host side, seeded).

minkowski_box: the periodic flat box of the adjoint workload -- rho
lognormal(0, 0.5), u = rho lognormal(-1, 0.3), 3-velocity with N(0,1)
components clipped to Lorentz factor <= 5, a divergence-free face field curled
from a seeded random vector potential and scaled so that max b^2/rho = 10.
"""
import numpy as np

NG = 3


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
