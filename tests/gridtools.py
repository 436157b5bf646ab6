"""Bridges between the product's block layout and the oracle's global grid.

Test infrastructure.  The product stores each block with its own ghost zones
([block][component][i][j][z], extents BS+6); the oracle works on one global
grid with ghosts.  Global cell index g (interior-relative) of block coordinate c
and local index l is g = c*BS + l - 3, so a block is the global slice
[c*BS, c*BS + BS + 6) in array coordinates.
"""
import numpy as np

import hamr_gamma_wind_amd as H
from tests import oracle_ctypes as O

NG = 3


def single_block_desc(desc):
    """The same grid as one block (the oracle's global metric comes from it)."""
    n = [desc.bs[k] * desc.nb[k] for k in range(3)]
    return H.make_desc(n, (1, 1, 1), gam=desc.gam, spin=desc.spin, cour=desc.cour, rin=desc.rin,
                       rout=desc.rout, rho_floor=desc.rho_floor, u_floor=desc.u_floor, minkowski=desc.minkowski)


def oracle_metric(desc):
    """Global metric in the oracle's layout (hamr_oracle.h)."""
    sd = single_block_desc(desc)
    lay = H.desc_layout(sd)
    ncol = lay["ncol"]
    m = H.desc_geometry(sd, 0)
    gcov = np.ascontiguousarray(m["gcov"][:, :, :ncol].transpose(0, 2, 1))
    gcon = np.ascontiguousarray(m["gcon"][:, :, :ncol].transpose(0, 2, 1))
    gdet = np.ascontiguousarray(m["gdet"][:, :ncol])
    conn = np.ascontiguousarray(m["conn"][:, :ncol].T)
    rcen = np.ascontiguousarray(m["rcen"][:ncol])
    return dict(gcov=gcov, gcon=gcon, gdet=gdet, conn=conn, rcen=rcen)


def oracle_grid(oracle, desc, metric=None, riemann=0):
    n = tuple(desc.bs[k] * desc.nb[k] for k in range(3))
    bs = tuple(desc.bs[k] for k in range(3))
    sx, dx = H.desc_coords(desc)
    if desc.minkowski:
        bc = (O.BC_PERIODIC, O.BC_PERIODIC, O.BC_PERIODIC)
    else:
        bc = (O.BC_OUTFLOW, O.BC_POLE, O.BC_PERIODIC)
    if metric is None:
        metric = oracle_metric(desc)
    return O.Grid(oracle, n, bs, bc, desc.gam, desc.cour, dx, metric, desc.rho_floor, desc.u_floor, riemann=riemann)


def block_coords(desc):
    nb = [desc.nb[k] for k in range(3)]
    out = []
    for c1 in range(nb[0]):
        for c2 in range(nb[1]):
            for c3 in range(nb[2]):
                out.append((c1, c2, c3))
    return out


def blocks_to_global(desc, arr):
    """Interior of every block -> a global array with zero ghosts."""
    bs = [desc.bs[k] for k in range(3)]
    n = [bs[k] * desc.nb[k] for k in range(3)]
    ncomp = arr.shape[1]
    G = np.zeros((ncomp, n[0] + 2 * NG, n[1] + 2 * NG, n[2] + 2 * NG), dtype=arr.dtype)
    for b, (c1, c2, c3) in enumerate(block_coords(desc)):
        G[:, NG + c1 * bs[0]: NG + (c1 + 1) * bs[0], NG + c2 * bs[1]: NG + (c2 + 1) * bs[1],
          NG + c3 * bs[2]: NG + (c3 + 1) * bs[2]] = arr[b][:, NG: NG + bs[0], NG: NG + bs[1], NG: NG + bs[2]]
    return np.ascontiguousarray(G)


def global_to_blocks(desc, G):
    """Every block's ghost-inclusive slice of a global array."""
    bs = [desc.bs[k] for k in range(3)]
    D = [b + 2 * NG for b in bs]
    out = np.zeros((len(block_coords(desc)), G.shape[0]) + tuple(D), dtype=G.dtype)
    for b, (c1, c2, c3) in enumerate(block_coords(desc)):
        out[b] = G[:, c1 * bs[0]: c1 * bs[0] + D[0], c2 * bs[1]: c2 * bs[1] + D[1], c3 * bs[2]: c3 * bs[2] + D[2]]
    return out


def block_view(desc, G):
    """Same as global_to_blocks for a global array without a component axis."""
    return global_to_blocks(desc, G[None])[:, 0]


# smallest normal double: the floor of the pointwise relative error, so that
# an exact zero on both sides is no error and a difference at a zero reference
# value is a huge one
DENORMAL_FLOOR = 2.2250738585072014e-308


def pointwise_relerr(a, b, axis_comp=1, floor=DENORMAL_FLOOR):
    """max over cells of |a - b| / max(|b|, floor) per component (component
    axis `axis_comp`): the FP64 parity bar |a - b| <= tol |b| cell by cell."""
    a = np.moveaxis(np.asarray(a), axis_comp, 0).reshape(np.shape(a)[axis_comp], -1)
    b = np.moveaxis(np.asarray(b), axis_comp, 0).reshape(a.shape[0], -1)
    d = np.abs(a - b)
    return np.max(d / np.maximum(np.abs(b), floor), axis=1, initial=0.0)


def pointwise_relerr_blocks(got, want, floor=DENORMAL_FLOOR):
    """pointwise_relerr over lists of per-block [component][...] arrays."""
    err = None
    for a, b in zip(got, want):
        e = pointwise_relerr(a[None], b[None], floor=floor)
        err = e if err is None else np.maximum(err, e)
    return err


def relerr_by_var(a, b, axis_comp=1):
    """max |a - b| / max |b| per component (component axis `axis_comp`).
    Normalised by the grid maximum: a loose check for the mixed arithmetic's
    error; FP64 parity uses pointwise_relerr."""
    a = np.moveaxis(a, axis_comp, 0).reshape(a.shape[axis_comp], -1)
    b = np.moveaxis(b, axis_comp, 0).reshape(b.shape[axis_comp], -1)
    scale = np.max(np.abs(b), axis=1)
    scale[scale == 0] = 1.0
    return np.max(np.abs(a - b), axis=1) / scale


def device_to_oracle_metric(m, D1, D2):
    """One geometry slab of the device layout -> the oracle's layout (D1 x D2 columns)."""
    ncol = D1 * D2
    return dict(gcov=np.ascontiguousarray(m["gcov"][:, :, :ncol].transpose(0, 2, 1)),
                gcon=np.ascontiguousarray(m["gcon"][:, :, :ncol].transpose(0, 2, 1)),
                gdet=np.ascontiguousarray(m["gdet"][:, :ncol]),
                conn=np.ascontiguousarray(m["conn"][:, :ncol].T),
                rcen=np.ascontiguousarray(m["rcen"][:ncol]))


def amr_oracle(oracle, desc, amr, riemann=0):
    """The oracle's multi-level grid for a block set (metric per block from the
    host geometry of the same set)."""
    info = H.desc_amr(desc, amr)
    rows = info["rows"]
    geos = {}
    metrics = []
    for r in rows:
        geo = int(r[7])
        D1, D2 = int(r[4]) + 2 * NG, int(r[5]) + 2 * NG
        if geo not in geos:
            geos[geo] = device_to_oracle_metric(H.desc_amr_geometry(desc, amr, geo, D1, D2), D1, D2)
        metrics.append(geos[geo])
    n0 = [desc.bs[k] * desc.nb[k] for k in range(3)]
    nlev = amr.nlev
    N = [tuple(x << l for x in n0) for l in range(nlev)]
    _, dx0 = H.desc_coords(desc)
    return O.AmrGrid(oracle, [r[:7] for r in rows], N, metrics, desc.gam, desc.cour, dx0, desc.rho_floor,
                     desc.u_floor, riemann=riemann), rows


def amr_divb(metrics, rows, dx0, ps):
    """Largest |div B| of any interior cell of a multi-level block set,
    relative to its block's scale: div B = sum_d (gdet B^d at the upper d-face
    - at the lower d-face) / dx_d (the block's level spacing), scale = sum_d
    max |gdet B^d| / dx_d over the block's faces.  metrics: per block, the
    oracle layout (gdet[pos][col], FACE1 = 0, FACE2 = 1, FACE3 = 4); rows:
    level, origin, size; ps: per block [3][D1][D2][D3] face fields."""
    worst = 0.0
    for b, r in enumerate(rows):
        bs = [int(k) for k in r[4:7]]
        D = [k + 2 * NG for k in bs]
        dx = [d * 2.0 ** -int(r[0]) for d in dx0]
        gd = metrics[b]["gdet"]
        f = [ps[b][a] * gd[pos].reshape(D[0], D[1])[:, :, None] for a, pos in enumerate((0, 1, 4))]
        I = tuple(slice(NG, NG + k) for k in bs)
        up = [tuple(slice(NG + (1 if k == a else 0), NG + bs[k] + (1 if k == a else 0)) for k in range(3))
              for a in range(3)]
        div = sum((f[a][up[a]] - f[a][I]) / dx[a] for a in range(3))
        scale = sum(np.abs(f[a][up[a]]).max() / dx[a] + np.abs(f[a][I]).max() / dx[a] for a in range(3))
        if scale > 0:
            worst = max(worst, float(np.abs(div).max() / scale))
    return worst


def las_levels(rows, seed=7):
    """Time levels for the three-level torus set: level-0 blocks 4 (those on
    the polar axis, which form the rings, always) or 2, level-1 blocks 2 or 1,
    level-2 blocks 1, by a seeded choice.  Finer blocks at level jumps are
    never slower than their coarser neighbours, and same-level neighbours of
    different levels occur on levels 0 and 1."""
    rng = np.random.default_rng(seed)
    tl = np.zeros(len(rows), dtype=np.int32)
    for b, r in enumerate(rows):
        lev, pole = int(r[0]), int(r[8])
        if lev == 0:
            tl[b] = 4 if (pole or rng.random() < 0.5) else 2
        elif lev == 1:
            tl[b] = 2 if rng.random() < 0.5 else 1
        else:
            tl[b] = 1
    return tl


def las_rule_py(bdt, dt, maxtl, rings, pairs):
    """set_timelevel's rule (LAS.c:5-144) restated: the largest power of 2 not
    above maxtl with bdt / dt >= 1.05 times it (else 1), each ring its minimum,
    finer blocks at level jumps lowered to their coarser neighbour's level,
    repeated until nothing changes."""
    tl = []
    for b in bdt:
        v = 1
        k = maxtl
        while k >= 1:
            if b / dt >= 1.05 * k:
                v = k
                break
            k //= 2
        tl.append(v)
    changed = True
    while changed:
        changed = False
        for ring in rings:
            m = min(tl[x] for x in ring)
            for x in ring:
                if tl[x] != m:
                    tl[x] = m
                    changed = True
        for c, f in pairs:
            if tl[f] > tl[c]:
                tl[f] = tl[c]
                changed = True
    return np.array(tl, dtype=np.int32)


def floored_relerr_blocks(got, want, rel_floor=1e-10):
    """Per component, max over every block's cells of |a - b| / max(|b|,
    rel_floor * max |b| of that component over all blocks): pointwise relative
    for every value above rel_floor of its component's largest, and relative
    to that floor below it (values that rounding-level terms build up from
    zero, such as a field component the step creates, are compared at the
    floor)."""
    nc = got[0].shape[0]
    big = np.zeros(nc)
    for b in want:
        big = np.maximum(big, np.abs(b).reshape(nc, -1).max(axis=1))
    fl = np.where(big > 0, rel_floor * big, DENORMAL_FLOOR)
    err = np.zeros(nc)
    for a, b in zip(got, want):
        d = np.abs(a - b).reshape(nc, -1)
        s = np.maximum(np.abs(b).reshape(nc, -1), fl[:, None])
        err = np.maximum(err, (d / s).max(axis=1))
    return err
