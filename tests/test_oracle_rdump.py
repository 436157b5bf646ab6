"""The oracle's restart record of one block (rdump_block_write /
rdump_block_read, restart.c:52-67 and :178-269) and its dump record
(dump_block, dump.c:276-356), pinned without the reference.

Pins: the record layout against numpy's own reshape of the ghost-inclusive
arrays (distinct extents per direction, distinct gdet per position, so a
swapped index or position fails); the write/read round trip; the reduction
(a record of a finer block) against closed-form means of linear and constant
fields, including the upper-row quirk of the reference's range test
(restart.c:244: fine cells past the last upper ghost row are skipped, with the
full 1/(red1 red2 red3) weight kept); and against an independent numpy slicing
of the fine record.  The dump record: layout by numpy reshape, the 4-velocity
a unit timelike vector equal to the closed form from the Lorentz factor.
"""
import numpy as np
import pytest

from . import oracle_ctypes as oc

NG, NPR, RREC = 3, 9, 13
FACE1, FACE2, CORN, CENT, FACE3 = range(5)


@pytest.fixture(scope="module")
def orc():
    return oc.load()


def _block(bs, rng, gdet=None):
    D = [b + 2 * NG for b in bs]
    nc, ncol = D[0] * D[1] * D[2], D[0] * D[1]
    p = rng.standard_normal((NPR, nc))
    ps = rng.standard_normal((3, nc))
    if gdet is None:
        gdet = rng.uniform(0.5, 2.0, (5, ncol))
    return D, p, ps, np.ascontiguousarray(gdet)


def test_record_layout(orc):
    bs = (4, 5, 6)
    rng = np.random.default_rng(1)
    D, p, ps, g = _block(bs, rng)
    rec = orc.rdump_write(bs, p, ps, g).reshape(D + [RREC])
    for k in range(NPR):
        assert np.array_equal(rec[..., k], p[k].reshape(D))
    assert np.all(rec[..., NPR] == 0.0)
    gcol = g.reshape(5, D[0], D[1], 1)
    for d, pos in enumerate((FACE1, FACE2, FACE3)):
        assert np.array_equal(rec[..., NPR + 1 + d], ps[d].reshape(D) * gcol[pos])


def _region(D, bs):
    """Mask of the cells a read assigns: [NG, NG+BS] along each direction."""
    m = np.zeros(D, dtype=bool)
    m[NG:NG + bs[0] + 1, NG:NG + bs[1] + 1, NG:NG + bs[2] + 1] = True
    return m.reshape(-1)


@pytest.mark.parametrize("spherical", [0, 1])
def test_round_trip(orc, spherical):
    bs = (4, 5, 6)
    rng = np.random.default_rng(2)
    D, p, ps, g = _block(bs, rng)
    rec = orc.rdump_write(bs, p, ps, g)
    sentinel_p = np.full_like(p, 7.25)
    sentinel_ps = np.full_like(ps, -3.5)
    p2, ps2 = orc.rdump_read(bs, (1, 1, 1), spherical, 40, rec, g, sentinel_p, sentinel_ps)
    m = _region(D, bs)
    assert np.array_equal(p2[:, m], p[:, m])
    assert np.all(p2[:, ~m] == 7.25) and np.all(ps2[:, ~m] == -3.5)
    # (x g) / g and, on the sphere, (((x g) / g) f) / f with f = 1 - 2/40:
    # equal to x up to a few roundings
    np.testing.assert_allclose(ps2[:, m], ps[:, m], rtol=4e-16 * 4, atol=0)
    # powers of two as gdet: exact
    g2 = np.ascontiguousarray(2.0 ** rng.integers(-3, 4, g.shape).astype(float))
    rec = orc.rdump_write(bs, p, ps, g2)
    p3, ps3 = orc.rdump_read(bs, (1, 1, 1), 0, 40, rec, g2, sentinel_p, sentinel_ps)
    assert np.array_equal(ps3[:, m], ps[:, m])


def _fine_record(bs, red, fields):
    """Record of the finer block with p_k(i, j, z) and the three face fields given
    as functions of the fine ghost-inclusive indices (gdet folded in)."""
    R = [b * r + 2 * NG for b, r in zip(bs, red)]
    i, j, z = np.meshgrid(*[np.arange(n, dtype=float) for n in R], indexing="ij")
    rec = np.zeros(R + [RREC])
    for k in range(RREC):
        rec[..., k] = fields(k, i, j, z)
    return rec.reshape(-1)


def test_reduction_linear_fields(orc):
    bs, red = (3, 4, 2), (2, 2, 2)
    D = [b + 2 * NG for b in bs]
    nc = D[0] * D[1] * D[2]
    coef = np.array([(k + 1) * 0.25 for k in range(RREC)])

    def lin(k, i, j, z):
        return coef[k] + 0.5 * i - 0.125 * j + 2.0 * z

    rec = _fine_record(bs, red, lin)
    g = np.ones((5, D[0] * D[1]))
    p, ps = orc.rdump_read(bs, red, 0, 16, rec, g, np.zeros((NPR, nc)), np.zeros((3, nc)))
    p = p.reshape((NPR,) + tuple(D))
    ps = ps.reshape((3,) + tuple(D))
    for a in range(bs[0] + 1):
        for b in range(bs[1] + 1):
            for c in range(bs[2] + 1):
                f = [NG + a * red[0], NG + b * red[1], NG + c * red[2]]
                n = [red[0] if a < bs[0] else 1, red[1] if b < bs[1] else 1, red[2] if c < bs[2] else 1]
                # closed form of the sum of a linear field over an n1 x n2 x n3 box
                ctr = [f[d] + (n[d] - 1) / 2.0 for d in range(3)]
                cnt = n[0] * n[1] * n[2]
                for k in range(NPR):
                    want = cnt / 8.0 * lin(k, ctr[0], ctr[1], ctr[2])
                    assert p[k, NG + a, NG + b, NG + c] == want
                # faces: the box's face at its lower edge along the direction
                w1 = n[1] * n[2] / 4.0 * lin(NPR + 1, f[0], ctr[1], ctr[2])
                w2 = n[0] * n[2] / 4.0 * lin(NPR + 2, ctr[0], f[1], ctr[2])
                w3 = n[0] * n[1] / 4.0 * lin(NPR + 3, ctr[0], ctr[1], f[2])
                assert ps[0, NG + a, NG + b, NG + c] == w1
                assert ps[1, NG + a, NG + b, NG + c] == w2
                assert ps[2, NG + a, NG + b, NG + c] == w3


def test_reduction_upper_row_weights(orc):
    """A constant record: interior cells get the constant, cells of the first
    upper ghost row get it times the fraction of the fine box inside the range
    (1/red per direction on that row)."""
    bs, red = (2, 2, 2), (2, 4, 1)
    D = [b + 2 * NG for b in bs]
    nc = D[0] * D[1] * D[2]
    rec = _fine_record(bs, red, lambda k, i, j, z: np.full(i.shape, 3.0))
    g = np.ones((5, D[0] * D[1]))
    p, ps = orc.rdump_read(bs, red, 0, 16, rec, g, np.zeros((NPR, nc)), np.zeros((3, nc)))
    p = p.reshape((NPR,) + tuple(D))
    ps = ps.reshape((3,) + tuple(D))
    for a in range(bs[0] + 1):
        for b in range(bs[1] + 1):
            for c in range(bs[2] + 1):
                frac = (1.0 if a < bs[0] else 1.0 / red[0]) * (1.0 if b < bs[1] else 1.0 / red[1]) * \
                    (1.0 if c < bs[2] else 1.0 / red[2])
                assert np.all(p[:, NG + a, NG + b, NG + c] == 3.0 * frac)
                f1 = (1.0 if b < bs[1] else 1.0 / red[1]) * (1.0 if c < bs[2] else 1.0 / red[2])
                assert ps[0, NG + a, NG + b, NG + c] == 3.0 * f1


def test_spherical_theta_ratio(orc):
    """B^2 carries fractheta / fractheta_old: 1 - 2/n2 over 1 - 2/(n2 red2)."""
    bs, red, n2 = (2, 2, 2), (1, 2, 1), 8
    D = [b + 2 * NG for b in bs]
    nc = D[0] * D[1] * D[2]
    rec = _fine_record(bs, red, lambda k, i, j, z: np.full(i.shape, 1.5))
    g = np.full((5, D[0] * D[1]), 1.0)
    _, ps_s = orc.rdump_read(bs, red, 1, n2, rec, g, np.zeros((NPR, nc)), np.zeros((3, nc)))
    _, ps_c = orc.rdump_read(bs, red, 0, n2, rec, g, np.zeros((NPR, nc)), np.zeros((3, nc)))
    m = _region(D, bs)
    ratio = (1 - 2.0 / n2) / (1 - 2.0 / (n2 * red[1]))
    np.testing.assert_allclose(ps_s[1, m], ps_c[1, m] * ratio, rtol=1e-15)
    assert np.array_equal(ps_s[0], ps_c[0]) and np.array_equal(ps_s[2], ps_c[2])


def test_reduction_matches_slicing(orc):
    """Random record of a block (2, 1, 3) times finer against numpy slice means."""
    bs, red = (3, 4, 2), (2, 1, 3)
    rng = np.random.default_rng(5)
    D = [b + 2 * NG for b in bs]
    nc = D[0] * D[1] * D[2]
    R = [b * r + 2 * NG for b, r in zip(bs, red)]
    rec = rng.standard_normal(R[0] * R[1] * R[2] * RREC)
    g = rng.uniform(0.5, 2.0, (5, D[0] * D[1]))
    p, ps = orc.rdump_read(bs, red, 0, 16, rec, g, np.zeros((NPR, nc)), np.zeros((3, nc)))
    p = p.reshape((NPR,) + tuple(D))
    ps = ps.reshape((3,) + tuple(D))
    fine = rec.reshape(R + [RREC])
    gg = g.reshape(5, D[0], D[1])
    w = 1.0 / (red[0] * red[1] * red[2])
    for a in range(bs[0] + 1):
        for b in range(bs[1] + 1):
            for c in range(bs[2] + 1):
                lo = [NG + a * red[0], NG + b * red[1], NG + c * red[2]]
                hi = [min(lo[d] + red[d], NG + bs[d] * red[d] + 1) for d in range(3)]
                box = fine[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
                i, j = NG + a, NG + b
                np.testing.assert_allclose(p[:, i, j, NG + c], box[..., :NPR].sum(axis=(0, 1, 2)) * w,
                                           rtol=1e-13, atol=1e-15)
                want = [box[0, :, :, NPR + 1].sum() / (red[1] * red[2]) / gg[FACE1, i, j],
                        box[:, 0, :, NPR + 2].sum() / (red[0] * red[2]) / gg[FACE2, i, j],
                        box[:, :, 0, NPR + 3].sum() / (red[0] * red[1]) / gg[FACE3, i, j]]
                np.testing.assert_allclose(ps[:, i, j, NG + c], want, rtol=1e-13, atol=1e-15)


PAIRS = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def _ks_columns(D1, D2):
    """Kerr-Schild centre metric of a D1 x D2 patch of columns, packed (ncol, 10)."""
    import hamr_gamma_wind_amd as H
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    sx, dx = H.desc_coords(desc)
    gcov = np.zeros((D1 * D2, 10))
    gcon = np.zeros((D1 * D2, 10))
    for i in range(D1):
        for j in range(D2):
            X = np.array([0.0, sx[0] + (i + 10.5) * dx[0], sx[1] + (j + 4.5) * dx[1], 0.0])
            gc, gi, _, _, _ = H.desc_metric_at(desc, X)
            gcov[i * D2 + j], gcon[i * D2 + j] = gc, gi
    return gcov, gcon


def test_dump_record_layout_and_four_velocity(orc):
    """dump_block: rho, u and B of the interior cells in index_3D order (numpy
    reshape), u^mu a unit timelike vector (g u u = -1 to float rounding) whose
    spatial part is U^i - gamma beta^i / alpha (gamma the Lorentz factor of U)."""
    bs = (3, 4, 5)
    D = [b + 2 * NG for b in bs]
    rng = np.random.default_rng(7)
    p = np.zeros((NPR,) + tuple(D))
    p[0] = rng.uniform(0.5, 2.0, D)
    p[1] = rng.uniform(0.1, 1.0, D)
    gcov, gcon = _ks_columns(D[0], D[1])
    # velocities of order 0.3 in a local unit basis
    for k, s in ((2, 4), (3, 7), (4, 9)):
        p[k] = rng.uniform(-0.3, 0.3, D) / np.sqrt(gcov[:, s].reshape(D[0], D[1], 1))
    p[5:8] = rng.standard_normal((3,) + tuple(D))
    rec = orc.dump_block(bs, p.reshape(NPR, -1), gcov, gcon).reshape(bs + (9,))
    I = tuple(slice(NG, NG + b) for b in bs)
    for slot, k in ((0, 0), (1, 1), (6, 5), (7, 6), (8, 7)):
        assert np.array_equal(rec[..., slot], p[k][I].astype(np.float32))
    for a in range(bs[0]):
        for b in range(bs[1]):
            col = (a + NG) * D[1] + (b + NG)
            g = np.zeros((4, 4))
            gi = np.zeros((4, 4))
            for s_, (x, y) in enumerate(PAIRS):
                g[x, y] = g[y, x] = gcov[col, s_]
                gi[x, y] = gi[y, x] = gcon[col, s_]
            alpha = 1.0 / np.sqrt(-gi[0, 0])
            for c in range(bs[2]):
                u = rec[a, b, c, 2:6].astype(np.float64)
                assert abs(u @ g @ u + 1.0) < 1e-5
                U = p[2:5, a + NG, b + NG, c + NG]
                gam = np.sqrt(1.0 + U @ g[1:, 1:] @ U)
                want = U - gam * gi[0, 1:] * alpha * alpha / alpha
                np.testing.assert_allclose(u[1:], want, rtol=2e-6, atol=1e-6 * abs(u[0]))
                np.testing.assert_allclose(u[0], gam / alpha, rtol=2e-6)
