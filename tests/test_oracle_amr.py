"""The oracle's multi-level step (oracle/hamr_amr_oracle.c), pinned on the CPU.

- On one level with blocks of mixed sizes (64x32x32, 32x16x16, 16^3), the
  block-wise ghost fill equals the single-grid ghost fill (oracle_bound, itself
  pinned by test_oracle_invariants) cell for cell, and one step of the block set
  equals one step of the single grid bit for bit: the exchange, the per-block
  polar rules and ring average and the final outflow pass reproduce the
  single-grid semantics.
- On three levels, the exchange (restriction, prolongation, copies) keeps a
  constant state exactly and a state linear in x2 (or in x1, away from the
  radial edge) to rounding, which fixes the child signs, the quarter slopes, the
  MC limiter on linear data and the face rule (odd fine faces from the coarse
  cell-centred field).
- The block-set builder (hamr_desc_amr) gives the multi-level torus workload
  its stated structure, and its exchange tables let each stage's entries run
  concurrently (no entry reads a cell its stage writes).
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H
from tests import gridtools as GT

NG = 3

# single level, mixed sizes: 128x64x32 cells as 2x2x1 superblocks
DESC1 = dict(bs=(64, 32, 32), nb=(2, 2, 1))
# three levels: base 64x64x32; level 1 covers x1 [0,64) x2 [32,96) of 128x128x64;
# level 2 covers x1 [0,64) x2 [96,160) of 256x256x128
BOXES3 = [((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 128))]
DESC3 = dict(bs=(64, 32, 32), nb=(1, 2, 1))


def _desc(d):
    return H.make_desc(d["bs"], d["nb"])


def _blocks_from_global(rows, G):
    """Each block's ghost-inclusive slice of a global level-0 array."""
    out = []
    for r in rows:
        o, bs = r[1:4], r[4:7]
        out.append(np.ascontiguousarray(G[:, o[0]:o[0] + bs[0] + 2 * NG, o[1]:o[1] + bs[1] + 2 * NG,
                                        o[2]:o[2] + bs[2] + 2 * NG]))
    return out


def _interior_only(rows, arrs):
    out = []
    for r, a in zip(rows, arrs):
        b = np.zeros_like(a)
        bs = r[4:7]
        I = (slice(None),) + tuple(slice(NG, NG + int(k)) for k in bs)
        b[I] = a[I]
        out.append(b)
    return out


def _faces_only(rows, arrs):
    """Each block's own face fields: component d on the faces [NG, NG+bs] along
    d, interior across (the upper boundary face is the block's own, updated by
    its CT step; the exchange never writes it)."""
    out = []
    for r, a in zip(rows, arrs):
        b = np.zeros_like(a)
        bs = [int(k) for k in r[4:7]]
        for d in range(3):
            I = [slice(NG, NG + bs[k]) for k in range(3)]
            I[d] = slice(NG, NG + bs[d] + 1)
            b[(d,) + tuple(I)] = a[(d,) + tuple(I)]
        out.append(b)
    return out


def _random_global(n, seed):
    rng = np.random.default_rng(seed)
    D = tuple(k + 2 * NG for k in n)
    P = np.zeros((9,) + D)
    S = np.zeros((3,) + D)
    I = (slice(None),) + tuple(slice(NG, NG + k) for k in n)
    shape = (len(n) and n)
    P[0][I[1:]] = rng.lognormal(0., 0.3, size=n)
    P[1][I[1:]] = rng.lognormal(-1., 0.3, size=n)
    for k in (2, 3, 4):
        P[k][I[1:]] = rng.normal(0., 0.3, size=n)
    for k in (5, 6, 7):
        P[k][I[1:]] = rng.normal(0., 0.1, size=n)
    P[8][I[1:]] = rng.lognormal(0., 0.2, size=n)
    for k in range(3):
        S[k][I[1:]] = rng.normal(0., 0.1, size=n)
    return P, S


def test_block_set_structure_of_the_multilevel_torus():
    # the multi-level torus workload: base 128x96x32, levels 1 and 2 refining
    # r <= 80 M, 30-150 deg and r <= 34.6 M, 60-120 deg
    desc = H.make_desc((64, 32, 32), (2, 3, 1))
    amr = H.make_amr_desc([((0, 160), (32, 160), (0, 64)), ((0, 256), (128, 256), (0, 128))], seed=2)
    r = H.desc_amr(desc, amr, tables=True)
    info, rows = r["info"], r["rows"]
    assert 650 <= info["blocks"] <= 800
    sizes = {tuple(int(x) for x in row[4:7]) for row in rows}
    assert sizes == {(64, 32, 32), (32, 16, 16), (16, 16, 16)}
    assert set(rows[:, 0].tolist()) == {0, 1, 2}
    # the leaves tile each level's region exactly: zones per level
    N0 = (128, 96, 32)
    vol = [N0[0] * N0[1] * N0[2] - 80 * 64 * 32, 160 * 128 * 64 - 128 * 64 * 64, 256 * 128 * 128]
    for l in range(3):
        sel = rows[rows[:, 0] == l]
        assert int(np.prod(sel[:, 4:7], axis=1).sum()) == vol[l]
    # polar blocks are level 0 and their rings close
    polar = rows[(rows[:, 8] & 3) != 0]
    assert (polar[:, 0] == 0).all() and info["rings"] > 0
    # table entries name cells inside their blocks (also checked on the host before upload)
    ncell = np.array([int(np.prod(D)) for D in H.amr_block_dims(rows)])
    for ex in r["ex"]:
        assert ((ex[:, 1] >= 0) & (ex[:, 1] < ncell[ex[:, 0]])).all()
        assert set(np.unique(ex[:, 3] >> 28).tolist()) <= {1, 2}
    assert sum(len(p) for p in r["pro"]) > 0 and sum(len(e) for e in r["ex"]) > 0
    # coarse faces at level jumps along x1 and x2 (flux correction entries);
    # the refined boxes span x3, so none along x3
    assert info["reflux1"] > 0 and info["reflux2"] > 0 and info["reflux3"] == 0
    # coarse edges on those jumps (EMF correction entries)
    assert info["emf"] > 0


def test_single_level_fill_equals_single_grid(oracle):
    desc = _desc(DESC1)
    amr = H.make_amr_desc([], seed=3, p_big=0.3, p_mid=0.4)
    g, rows = GT.amr_oracle(oracle, desc, amr)
    assert len({tuple(int(x) for x in r[4:7]) for r in rows}) >= 2  # mixed sizes
    n = tuple(desc.bs[k] * desc.nb[k] for k in range(3))
    P, S = _random_global(n, seed=7)
    grid = GT.oracle_grid(oracle, desc)
    bp = _blocks_from_global(rows, P)
    bs_ = _blocks_from_global(rows, S)
    want_p = [a.copy() for a in bp]
    want_s = [a.copy() for a in bs_]
    grid.bound(P, S)
    want_p = _blocks_from_global(rows, P)
    want_s = _blocks_from_global(rows, S)
    p = _interior_only(rows, bp)
    ps = _faces_only(rows, want_s)  # the upper x3 face of the last phi block is the grid's wrapped ghost face
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
    g.bound(p, ps, flags)
    for b in range(len(rows)):
        assert np.array_equal(p[b], want_p[b]), b
        assert np.array_equal(ps[b], want_s[b]), b


def test_single_level_step_equals_single_grid(oracle):
    desc = _desc(DESC1)
    amr = H.make_amr_desc([], seed=3, p_big=0.3, p_mid=0.4)
    g, rows = GT.amr_oracle(oracle, desc, amr)
    # single grid with the polar blocks' phi size (the ring average runs over 16-cell blocks)
    sdesc = H.make_desc((64, 32, 16), (2, 2, 2))
    grid = GT.oracle_grid(oracle, sdesc)
    p0, ps0 = H.desc_torus(desc, H.torus_desc())
    P = GT.blocks_to_global(desc, p0)
    S = GT.blocks_to_global(desc, ps0)
    grid.bound(P, S)
    bp = _blocks_from_global(rows, P)
    bs_ = _blocks_from_global(rows, S)
    dt = 0.5 * min(H.desc_coords(desc)[1])
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in bp]
    g.step(bp, bs_, flags, dt)
    grid.step(P, S, dt)
    want_p = _blocks_from_global(rows, P)
    want_s = _blocks_from_global(rows, S)
    for b in range(len(rows)):
        assert np.array_equal(bp[b], want_p[b]), b
        assert np.array_equal(bs_[b], want_s[b]), b


def _level_positions(r, dx0, sx):
    """Code coordinates X1, X2 of the cell centres of block row r (ghost-inclusive)."""
    l = int(r[0])
    dx = [d * 2.0 ** -l for d in dx0]
    i = np.arange(int(r[4]) + 2 * NG) + int(r[1]) - NG
    j = np.arange(int(r[5]) + 2 * NG) + int(r[2]) - NG
    return sx[0] + (i + 0.5) * dx[0], sx[1] + (j + 0.5) * dx[1], dx


@pytest.mark.parametrize("axis", [1, 0])
def test_multilevel_exchange_linear_and_constant_states(oracle, axis):
    desc = _desc(DESC3)
    amr = H.make_amr_desc(BOXES3, seed=5)
    g, rows = GT.amr_oracle(oracle, desc, amr)
    assert set(rows[:, 0].tolist()) == {0, 1, 2}
    sx, dx0 = H.desc_coords(desc)
    # dyadic coefficients: sums of eight equal values are then exact
    coef = np.array([1.0, 0.5, 0.25, -0.25, 0.125, 0.75, -0.375, 0.25, 2.0])

    def fill(linear):
        p, ps = [], []
        for r in rows:
            X1, X2, dx = _level_positions(r, dx0, sx)
            D = tuple(int(b) + 2 * NG for b in r[4:7])
            Xc = (X1 if axis == 0 else X2)
            Xf = (X1 - 0.5 * dx[0]) if axis == 0 else (X2 - 0.5 * dx[1])  # lower face along the varying axis
            shape = [1, 1, 1]
            shape[axis] = D[axis]
            base = np.broadcast_to(Xc.reshape(shape), D) if linear else np.ones(D)
            face = np.broadcast_to(Xf.reshape(shape), D) if linear else np.ones(D)
            P = np.stack([1.0 + c * base for c in coef])
            F = np.stack([0.25 + 0.5 * (face if a == axis else base) for a in range(3)])
            # the cell-centred field B_a matches face a's function
            for a in range(3):
                P[5 + a] = 0.25 + 0.5 * base
            p.append(np.ascontiguousarray(P))
            ps.append(np.ascontiguousarray(F))
        return p, ps

    ex = H.desc_amr(desc, amr, tables=True)["ex"]
    for linear in (False, True):
        p, ps = fill(linear)
        want_p = [a.copy() for a in p]
        want_s = [a.copy() for a in ps]
        # the exchange-normal face field of a restriction ghost is the finer
        # faces' gdet-weighted mean over the coarse face's gdet
        flux_means = _restriction_normal_faces(g, rows, ex, ps)
        known = {}
        for (b, d0, cells), v in flux_means:
            for c, x in zip(cells.tolist(), v.tolist()):
                known[(b, d0, c)] = x
        # a corner ghost of stage d0 copies its face neighbour's ghost of an
        # earlier stage, a restriction's mean included
        for d0 in (1, 2):
            e = ex[d0]
            for row in e[(e[:, 3] >> 28) == 1].tolist():
                for ee in range(d0):
                    key = (row[2], ee, row[3] & 0x0FFFFFFF)
                    if key in known:
                        known[(row[0], ee, row[1])] = known[key]
        for (b, comp, c), x in known.items():
            want_s[b][comp].reshape(-1)[c] = x
        flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
        for stage in (1, 2, 3):
            g.exchange(stage, p, ps, flags)
        for b, r in enumerate(rows):
            if not linear:
                assert np.array_equal(p[b], want_p[b]) and np.array_equal(ps[b], want_s[b]), b
                continue
            # prolongations at the radial edge have no slope across it: skip the
            # cells within 8 cells of r_in when x1 varies
            m = np.ones(p[b].shape[1:], bool)
            if axis == 0:
                i0 = int(r[1]) - NG
                m[: max(0, 8 - i0), :, :] = False
            ep = np.abs(p[b] - want_p[b])[:, m].max()
            es = np.abs(ps[b] - want_s[b])[:, m].max()
            assert ep < 1e-12 and es < 1e-12, (b, ep, es)
        assert sum(len(c) for (b, d0, c), v in flux_means) > 0


FACEPOS = (0, 1, 4)  # metric positions FACE1, FACE2, FACE3


def _own_upper_face(r, d0, loc):
    """Cells (loc: arrays of local i, j, z) that hold block r's own upper
    boundary face along d0: index NG+bs along d0, interior across."""
    bs = [int(k) for k in r[4:7]]
    own = loc[d0] == NG + bs[d0]
    for k in range(3):
        if k != d0:
            own &= (loc[k] >= NG) & (loc[k] < NG + bs[k])
    return own


def _restriction_normal_faces(g, rows, ex, ps):
    """Expected d0-face field of every restriction ghost of the exchange along
    d0 (except a block's own upper face, which the exchange leaves alone):
    (0.25 (sum over the four finer faces of B gdet)) / gdet of the coarse face,
    the finer faces' term order (j,z), (j,z+1), (j+1,z), (j+1,z+1) for x1 and
    its analogues, from the arrays ps before the exchange.  Returns a list of
    ((block, d0, flat cells), values)."""
    out = []
    for d0 in range(3):
        e = ex[d0]
        e = e[(e[:, 3] >> 28) == 2]
        for n in np.unique(e[:, 0]):
            en = e[e[:, 0] == n]
            rn = rows[n]
            Dn = [int(k) + 2 * NG for k in rn[4:7]]
            dc = en[:, 1].astype(np.int64)
            loc = np.unravel_index(dc, Dn)
            keep = ~_own_upper_face(rn, d0, loc)
            en, dc = en[keep], dc[keep]
            loc = tuple(x[keep] for x in loc)
            gc = g._m[n]["gdet"][FACEPOS[d0]][loc[0] * Dn[1] + loc[1]]
            vals = np.empty(len(en))
            for c in np.unique(en[:, 2]):
                sel = en[:, 2] == c
                rc = rows[c]
                Dc = [int(k) + 2 * NG for k in rc[4:7]]
                sc = (en[sel, 3] & 0x0FFFFFFF).astype(np.int64)
                i, j, z = np.unravel_index(sc, Dc)
                f = ps[c][d0]
                gd = g._m[c]["gdet"][FACEPOS[d0]]

                def col(a, b):
                    return gd[a * Dc[1] + b]
                if d0 == 0:
                    m = 0.25 * (f[i, j, z] * col(i, j) + f[i, j, z + 1] * col(i, j) + f[i, j + 1, z] * col(i, j + 1)
                                + f[i, j + 1, z + 1] * col(i, j + 1))
                elif d0 == 1:
                    m = 0.25 * (f[i, j, z] * col(i, j) + f[i, j, z + 1] * col(i, j) + f[i + 1, j, z] * col(i + 1, j)
                                + f[i + 1, j, z + 1] * col(i + 1, j))
                else:
                    m = 0.25 * (f[i, j, z] * col(i, j) + f[i, j + 1, z] * col(i, j + 1) + f[i + 1, j, z] * col(i + 1, j)
                                + f[i + 1, j + 1, z] * col(i + 1, j + 1))
                vals[sel] = m / gc[sel]
            out.append(((int(n), d0, dc), vals))
    return out


# level jumps along x3 too, one across the periodic seam
BOXES3Z = [((0, 64), (32, 96), (0, 64)), ((0, 64), (96, 160), (0, 64))]


@pytest.mark.parametrize("boxes", ["x1x2", "x1x2x3"])
@pytest.mark.parametrize("reflux", [True, False])
def test_level_jump_flux_correction_conserves(oracle, reflux, boxes):
    """Conservation across level jumps: summed over every block, the outward
    flux through the block's faces times the face area (the flux divergence
    summed over its cells) cancels between blocks except on the domain's radial
    edges and polar faces.  Same-level neighbours compute the shared face from
    identical ghost data; at a level jump the coarse face must carry the mean of
    the four finer faces (face area 4x), which the correction provides.  Without
    it the internal faces leave an imbalance, so the check sees a missing,
    misplaced or mis-indexed correction."""
    desc = _desc(DESC3)
    amr = H.make_amr_desc(BOXES3 if boxes == "x1x2" else BOXES3Z, seed=5)
    g, rows = GT.amr_oracle(oracle, desc, amr)
    g.g.reflux = 1 if reflux else 0
    p, ps = H.desc_amr_torus(desc, amr, H.torus_desc())
    p = [np.ascontiguousarray(a) for a in p]
    ps = [np.ascontiguousarray(a) for a in ps]
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
    g.bound(p, ps, flags)
    _, dx0 = H.desc_coords(desc)
    dt = 0.25 * min(dx0)
    F = g.fluxes(p, ps, dt)
    N = [[int(x) << l for x in (desc.bs[k] * desc.nb[k] for k in range(3))] for l in range(3)]
    net = np.zeros(9)
    mag = np.zeros(9)
    for b, r in enumerate(rows):
        l, o, bs = int(r[0]), [int(x) for x in r[1:4]], [int(x) for x in r[4:7]]
        dx = [d * 2.0 ** -l for d in dx0]
        for d in range(3):
            area = dx[(d + 1) % 3] * dx[(d + 2) % 3]
            for side in (0, 1):
                # a face on the radial edge or the polar axis is a domain boundary
                if d < 2 and (o[d] + (bs[d] if side else 0)) in (0, N[l][d]):
                    continue
                idx = [slice(NG, NG + bs[k]) for k in range(3)]
                idx[d] = NG + bs[d] if side else NG
                f = F[d][b][(slice(None),) + tuple(idx)].reshape(9, -1)
                s = (1.0 if side else -1.0) * area * f.sum(axis=1)
                net += s
                mag += area * np.abs(f).sum(axis=1)
    rel = np.abs(net) / mag
    print("reflux" if reflux else "no reflux", "net / sum|terms| per variable:", np.array2string(rel, precision=2))
    if reflux:
        assert (rel < 1e-13).all(), rel
    else:
        assert rel[:5].max() > 1e-8, rel  # the uncorrected coarse faces break the balance


def _leaf_finder(rows, N):
    """leaf(l, Y) -> (block, level) holding cell Y of level l (16-cell units)."""
    U = 16
    own = [np.full(tuple(n // U for n in N[l]), -1, dtype=np.int64) for l in range(len(N))]
    for b, r in enumerate(rows):
        l, o, bs = int(r[0]), [int(x) for x in r[1:4]], [int(x) for x in r[4:7]]
        own[l][o[0] // U:(o[0] + bs[0]) // U, o[1] // U:(o[1] + bs[1]) // U, o[2] // U:(o[2] + bs[2]) // U] = b

    def leaf(l, Y):
        b = own[l][Y[0] // U, Y[1] // U, Y[2] // U]
        if b >= 0:
            return int(b), l
        if l + 1 < len(N):
            b = own[l + 1][2 * Y[0] // U, 2 * Y[1] // U, 2 * Y[2] // U]
            if b >= 0:
                return int(b), l + 1
        return int(own[l - 1][(Y[0] // 2) // U, (Y[1] // 2) // U, (Y[2] // 2) // U]), l - 1
    return leaf


@pytest.mark.parametrize("boxes", ["x1x2", "x1x2x3"])
@pytest.mark.parametrize("correct", [True, False])
def test_level_jump_emf_correction_matches_face_flux_updates(oracle, correct, boxes):
    """div B across level jumps: in the corrector every block updates its own
    copy of a face it shares with a neighbour.  The change of magnetic flux
    (gdet B times the face area) through a coarse face is the line integral of
    its edge EMFs, that of the four finer faces tiling it the integral of the
    finer edges around them.  With the coarse edges on the jump replaced by the
    mean of the two finer edges along them the two changes agree to rounding on
    every coarse face of a level jump; without it they differ at order one.  A missing, misplaced, mis-indexed or mis-weighted
    edge correction breaks the agreement."""
    desc = _desc(DESC3)
    amr = H.make_amr_desc(BOXES3 if boxes == "x1x2" else BOXES3Z, seed=5)
    g, rows = GT.amr_oracle(oracle, desc, amr)
    g.g.reflux = 1 if correct else 0
    p, ps = H.desc_amr_torus(desc, amr, H.torus_desc())
    # a uniform field on top of the torus's (which vanishes outside the torus)
    # so that every level jump sees nonzero EMFs
    bg = np.array([0.02, 0.01, 0.005])
    p = [np.ascontiguousarray(a) for a in p]
    ps = [np.ascontiguousarray(a) for a in ps]
    for a, c in zip(p, ps):
        a[5:8] += bg[:, None, None, None]
        c += bg[:, None, None, None]
    flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
    g.bound(p, ps, flags)
    _, dx0 = H.desc_coords(desc)
    dt = 0.25 * min(dx0)
    psf = g.ct(p, ps, dt)
    n0 = [desc.bs[k] * desc.nb[k] for k in range(3)]
    N = [[x << l for x in n0] for l in range(3)]
    leaf = _leaf_finder(rows, N)
    gdet = [np.asarray(m["gdet"]).reshape(-1) for m in g._m]
    pos = (0, 1, 4)
    worst, scale, count = 0.0, 0.0, 0
    for b, r in enumerate(rows):
        l, o, bs = int(r[0]), [int(x) for x in r[1:4]], [int(x) for x in r[4:7]]
        Dc = g.D[b]
        dxc = [d * 2.0 ** -l for d in dx0]
        dxf = [d * 0.5 for d in dxc]
        for d in range(3):
            ta, tb = [k for k in range(3) if k != d]
            for side in (0, 1):
                Y = [0, 0, 0]
                Y[d] = o[d] + bs[d] if side else o[d] - 1
                if d < 2 and not (0 <= Y[d] < N[l][d]):
                    continue
                Y[2] %= N[l][2]
                for ya in range(o[ta], o[ta] + bs[ta]):
                    for yb in range(o[tb], o[tb] + bs[tb]):
                        Y[ta], Y[tb] = ya, yb
                        f, lf = leaf(l, Y)
                        if lf != l + 1:
                            continue
                        fr = rows[f]
                        fo, fbs = [int(x) for x in fr[1:4]], [int(x) for x in fr[4:7]]
                        Df = g.D[f]
                        Kc = [ya - o[ta] + NG, yb - o[tb] + NG]
                        kc = NG + bs[d] if side else NG
                        ci = [0, 0, 0]
                        ci[d], ci[ta], ci[tb] = kc, Kc[0], Kc[1]
                        cell_c = (ci[0] * Dc[1] + ci[1]) * Dc[2] + ci[2]
                        col_c = ci[0] * Dc[1] + ci[1]
                        area_c = dxc[ta] * dxc[tb]
                        dphi_c = (psf[b][d].reshape(-1)[cell_c] - ps[b][d].reshape(-1)[cell_c]) * \
                            gdet[b][pos[d] * Dc[0] * Dc[1] + col_c] * area_c
                        Zd = (2 * Y[d] if side else 2 * Y[d] + 2)
                        if d == 2 and not (fo[2] <= Zd <= fo[2] + fbs[2]):
                            Zd += N[l + 1][2] if Zd < fo[2] else -N[l + 1][2]
                        dphi_f = 0.0
                        for a in (0, 1):
                            for c in (0, 1):
                                fi = [0, 0, 0]
                                fi[d] = Zd - fo[d] + NG
                                fi[ta] = 2 * ya + a - fo[ta] + NG
                                fi[tb] = 2 * yb + c - fo[tb] + NG
                                cf = (fi[0] * Df[1] + fi[1]) * Df[2] + fi[2]
                                colf = fi[0] * Df[1] + fi[1]
                                dphi_f += (psf[f][d].reshape(-1)[cf] - ps[f][d].reshape(-1)[cf]) * \
                                    gdet[f][pos[d] * Df[0] * Df[1] + colf] * dxf[ta] * dxf[tb]
                        worst = max(worst, abs(dphi_c - dphi_f))
                        scale = max(scale, abs(dphi_c), abs(dphi_f))
                        count += 1
    rel = worst / scale
    print("EMF correction" if correct else "no EMF correction",
          "%d coarse faces of level jumps, max |dPhi_c - sum dPhi_f| / max |dPhi| = %.2e" % (count, rel))
    assert count > 0
    # the changes are small differences of edge terms (a nearly uniform field):
    # rounding shows at ~1e-12 of the largest change (2.3e-12 observed), the
    # uncorrected edges at order one (0.41 observed)
    if correct:
        assert rel < 1e-10, rel
    else:
        assert rel > 1e-3, rel


def test_level_jumps_keep_div_b_at_roundoff(oracle):
    """Each block's constrained-transport update keeps the divergence of every
    interior cell's field at roundoff as long as no ghost fill overwrites one
    of its own faces.  On three levels (level jumps along x1 and x2) the
    torus's div B stays at roundoff over four steps; with the exchange made to
    write the normal field of each block's own upper face (a planted mistake:
    the restricted or prolonged value replaces the face the block's CT
    evolved) it grows to O(1) in the first step."""
    import ctypes

    desc = _desc(DESC3)
    amr = H.make_amr_desc(BOXES3, seed=5)
    g, rows = GT.amr_oracle(oracle, desc, amr)
    _, dx0 = H.desc_coords(desc)
    plant = ctypes.c_int.in_dll(oracle.lib, "oracle_amr_plant_own_face")
    worst = {}
    try:
        for planted in (0, 1):
            plant.value = planted
            p, ps = H.desc_amr_torus(desc, amr, H.torus_desc())
            p = [np.ascontiguousarray(a) for a in p]
            ps = [np.ascontiguousarray(a) for a in ps]
            flags = [np.zeros(a.shape[1:], dtype=np.int32) for a in p]
            assert GT.amr_divb(g._m, rows, dx0, ps) < 1e-14
            g.bound(p, ps, flags)
            dt = 0.125 * min(dx0)
            w = []
            for _ in range(4 if not planted else 1):
                dt = g.step(p, ps, flags, dt)
                w.append(GT.amr_divb(g._m, rows, dx0, ps))
            worst[planted] = max(w)
    finally:
        plant.value = 0
    print("div B relative to the block scale: %.2e (planted overwrite: %.2e)" % (worst[0], worst[1]))
    assert worst[0] < 1e-12
    assert worst[1] > 1e-3


@pytest.mark.parametrize("boxes", ["x1x2", "x1x2x3"])
def test_exchange_entries_never_read_a_cell_their_stage_writes(boxes):
    """The exchange kernels of one stage run their entries concurrently: no
    entry may read (as a copied cell, one of the 2x2x2 children of a
    restriction, a coarse cell or one of its six slope neighbours) a cell that
    an entry of the same stage writes, and no ghost is written twice."""
    desc = _desc(DESC3)
    amr = H.make_amr_desc(BOXES3 if boxes == "x1x2" else BOXES3Z, seed=5)
    info = H.desc_amr(desc, amr, tables=True)
    rows = info["rows"]
    D = np.array([[int(k) + 2 * NG for k in r[4:7]] for r in rows], dtype=np.int64)
    for s in range(3):
        ex, pro = info["ex"][s].astype(np.int64), info["pro"][s].astype(np.int64)
        dst = np.concatenate([(ex[:, 0] << 32) | ex[:, 1], (pro[:, 0] << 32) | pro[:, 1]])
        assert len(np.unique(dst)) == len(dst), s
        kind, cell = ex[:, 3] >> 28, ex[:, 3] & 0x0FFFFFFF
        srcs = [(ex[:, 2] << 32) | cell]
        avg = kind == 2
        b = ex[avg, 2]
        sj, si = D[b, 2], D[b, 1] * D[b, 2]
        c = cell[avg]
        for off in (1, sj, sj + 1, si, si + 1, si + sj, si + sj + 1):
            srcs.append((b << 32) | (c + off))
        srcs.append((pro[:, 2] << 32) | pro[:, 3])
        for k in range(6):
            code = pro[:, 5 + 2 * k]
            has = (code >> 28) != 0
            srcs.append((pro[has, 4 + 2 * k] << 32) | (code[has] & 0x0FFFFFFF))
        src = np.concatenate(srcs)
        bad = np.isin(src, dst)
        print("stage %d: %d entries, %d source reads, %d of written cells" % (s + 1, len(dst), len(src), bad.sum()))
        assert not bad.any(), s


@pytest.mark.parametrize("which", ["x1x2", "x1x2x3", "torus722"])
def test_stage3_tail_is_the_rows_the_update_writes(which):
    """Stage 3's last info['push3'] entries are exactly the x3 ghosts of the
    rows inside a block's x1 and x2 range whose position (the cell across the
    block's x3 face) lies in a block of the same level, each copying that
    block's cell at the position (found here from the block rows: blocks tile
    each level in 16-cell units); the step's fill leaves them out because that
    block's update writes them.  No other entry has such a destination."""
    from hamr_gamma_wind_amd import workloads as W
    if which == "torus722":
        A = W.AMR_TORUS
        desc, amr = H.make_desc(A["bs"], A["nb"]), H.make_amr_desc(A["boxes"], seed=A["seed"])
    else:
        desc = _desc(DESC3)
        amr = H.make_amr_desc(BOXES3 if which == "x1x2" else BOXES3Z, seed=5)
    info = H.desc_amr(desc, amr, tables=True)
    rows = info["rows"].astype(np.int64)
    n0 = [desc.bs[k] * desc.nb[k] for k in range(3)]
    U = 16
    owner = {}
    for b, r in enumerate(rows):
        lev, o, bs = int(r[0]), r[1:4], r[4:7]
        for u1 in range(o[0] // U, (o[0] + bs[0]) // U):
            for u2 in range(o[1] // U, (o[1] + bs[1]) // U):
                for u3 in range(o[2] // U, (o[2] + bs[2]) // U):
                    owner[(lev, u1, u2, u3)] = b
    want_dst, want_src = [], []
    for b, r in enumerate(rows):
        lev, o, bs = int(r[0]), r[1:4], r[4:7]
        D = bs + 2 * NG
        N3 = n0[2] << lev
        for side in (0, 1):
            zpos = (o[2] - 1) % N3 if side == 0 else (o[2] + bs[2]) % N3
            for u1 in range(o[0] // U, (o[0] + bs[0]) // U):
                for u2 in range(o[1] // U, (o[1] + bs[1]) // U):
                    s_ = owner.get((lev, u1, u2, zpos // U))
                    if s_ is None:
                        continue
                    S = rows[s_]
                    DS = S[4:7] + 2 * NG
                    i, j, g = np.meshgrid(np.arange(U), np.arange(U), np.arange(NG), indexing="ij")
                    Y1, Y2 = u1 * U + i, u2 * U + j
                    z = g if side == 0 else NG + bs[2] + g
                    Y3 = (o[2] + z - NG) % N3
                    dc = ((Y1 - o[0] + NG) * D[1] + (Y2 - o[1] + NG)) * D[2] + z
                    sc = ((Y1 - S[1] + NG) * DS[1] + (Y2 - S[2] + NG)) * DS[2] + ((Y3 - S[3]) % N3 + NG)
                    want_dst.append((b << 32) | dc.ravel())
                    want_src.append((s_ << 32) | sc.ravel())
    want_dst, want_src = np.concatenate(want_dst), np.concatenate(want_src)
    ex3 = info["ex"][2].astype(np.int64)
    npush = info["info"]["push3"]
    tail, head = ex3[len(ex3) - npush:], ex3[:len(ex3) - npush]
    print("%s: %d stage-3 entries, %d written by the update" % (which, len(ex3), npush))
    assert npush == len(want_dst) and npush > 0
    assert np.all((tail[:, 3] >> 28) == 1)
    got_dst = (tail[:, 0] << 32) | tail[:, 1]
    got_src = (tail[:, 2] << 32) | (tail[:, 3] & 0x0FFFFFFF)
    a, b_ = np.argsort(got_dst), np.argsort(want_dst)
    assert np.array_equal(got_dst[a], want_dst[b_])
    assert np.array_equal(got_src[a], want_src[b_])
    assert not np.isin((head[:, 0] << 32) | head[:, 1], want_dst).any()
