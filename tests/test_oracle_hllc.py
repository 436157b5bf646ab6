"""The oracle's HLLC solver (fluxcalc2D_FT with set_Mud, calc_HLLC_hydro and
calc_HLLC; GPU_hllc.h:3-400, hllc.c:20-213), pinned without the reference (its
CPU HLLC path never computes wave speeds, and the GPU one cannot run here).

Pins:
- the face tetrads on the Kerr-Schild metric: orthonormal (e_a . e_b = eta_ab),
  the inverse is the inverse, e_(0) is the normal observer (g e_(0) =
  -alpha dt), the zero pattern of the construction (e_(c) along d_c, e_(b) in
  the (b, c) plane, row 0 only in e_(0)) and positive diagonal; identity on
  flat space; the host build equal to the oracle's;
- a face with equal states on both sides: the HLLC flux equals the
  coordinate flux of that state (transform, frame fluxes and back-transform
  against the plain coordinate primtoflux), with and without field, on the
  Kerr-Schild metric;
- flat space, a moving contact (rho jumps; pressure, velocity and field
  continuous): the flux is the upwind state's physical flux, without field
  (hydro HLLC) and with a normal and a tangential field (MHD HLLC), where HLL
  smears the density;
- a stationary contact (contact speed exactly 0, rejected by the reference's
  admissibility test) and a supersonic face fall back to HLL and to the
  upwind flux.
"""
import numpy as np
import pytest

import hamr_gamma_wind_amd as H

from . import oracle_ctypes as oc

GAM = 13.0 / 9.0
ETA = np.diag([-1.0, 1.0, 1.0, 1.0])
PAIRS = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


@pytest.fixture(scope="module")
def orc():
    return oc.load()


def full(g10):
    m = np.zeros((4, 4))
    for s, (a, b) in enumerate(PAIRS):
        m[a, b] = m[b, a] = g10[s]
    return m


def ks_points():
    desc = H.make_desc((16, 8, 8), (4, 4, 4))
    sx, dx = H.desc_coords(desc)
    rng = np.random.default_rng(3)
    pts = []
    for _ in range(12):
        X = np.array([0.0, sx[0] + rng.uniform(0, 64) * dx[0], sx[1] + rng.uniform(0.5, 31.5) * dx[1],
                      rng.uniform(0, 6.2)])
        gcov, gcon, gdet, _, _ = H.desc_metric_at(desc, X)
        pts.append((gcov, gcon, gdet))
    return pts


@pytest.mark.parametrize("d", [1, 2, 3])
def test_tetrad_is_orthonormal_frame_of_the_normal_observer(orc, d):
    for gcov, gcon, _ in ks_points():
        mud, mudi = orc.tetrad(gcov, gcon, d)
        g = full(gcov)
        np.testing.assert_allclose(mud.T @ g @ mud, ETA, atol=1e-12)
        np.testing.assert_allclose(mudi @ mud, np.eye(4), atol=1e-12)
        alpha = 1.0 / np.sqrt(-full(gcon)[0, 0])
        np.testing.assert_allclose(g @ mud[:, 0], [-alpha, 0, 0, 0], atol=1e-12)
        a, b = d, d % 3 + 1
        c = b % 3 + 1
        nz = np.zeros((4, 4), dtype=bool)
        nz[:, 0] = True
        nz[[a, b, c], a] = True
        nz[[b, c], b] = True
        nz[c, c] = True
        assert np.all(mud[~nz] == 0.0)
        assert np.all(mud[0, 1:] == 0.0)
        assert mud[0, 0] > 0 and mud[a, a] > 0 and mud[b, b] > 0 and mud[c, c] > 0
        hm, hmi = H.face_tetrad(gcov, gcon, d)
        assert np.array_equal(hm, mud) and np.array_equal(hmi, mudi)


def test_tetrad_flat_is_identity(orc):
    g = np.array([-1.0, 0, 0, 0, 1.0, 0, 0, 1.0, 0, 1.0])
    for d in (1, 2, 3):
        mud, mudi = orc.tetrad(g, g, d)
        assert np.array_equal(mud, np.eye(4)) and np.array_equal(mudi, np.eye(4))


def _state(rng, bmag, gcov=None):
    """rho, u O(1); velocity and field components O(0.3) and O(bmag) in a
    local unit basis (code components divided by sqrt(g_ii))."""
    scale = np.ones(3) if gcov is None else 1.0 / np.sqrt(np.array([gcov[4], gcov[7], gcov[9]]))
    p = np.zeros(9)
    p[0] = rng.uniform(0.5, 2.0)
    p[1] = rng.uniform(0.1, 1.0)
    p[2:5] = rng.uniform(-0.3, 0.3, 3) * scale
    p[5:8] = rng.uniform(-1, 1, 3) * bmag * scale
    p[8] = (GAM - 1) * p[1] * p[0] ** -GAM
    return p


@pytest.mark.parametrize("bmag", [0.0, 0.5])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_equal_states_give_the_coordinate_flux(orc, d, bmag):
    rng = np.random.default_rng(10 * d + int(bmag * 10))
    for gcov, gcon, gdet in ks_points():
        p = _state(rng, bmag, gcov)
        F, _ = orc.hllc_face(p, p, gcov, gcon, gdet, GAM, d)
        want, _, _ = orc.primtoflux(p, gcov, gcon, gdet, GAM, d)
        scale = np.abs(want).max()
        np.testing.assert_allclose(F, want, rtol=0, atol=1e-11 * scale)


def _flat():
    g = np.array([-1.0, 0, 0, 0, 1.0, 0, 0, 1.0, 0, 1.0])
    return g, g.copy(), 1.0


@pytest.mark.parametrize("field", ["none", "normal+tangential"])
@pytest.mark.parametrize("v", [0.3, -0.25])
def test_moving_contact_is_resolved(orc, field, v):
    """rho jumps across the face; p, v and B are the same on both sides.  The
    HLLC flux is the upwind side's physical flux (v > 0: left)."""
    gcov, gcon, g = _flat()
    pl = np.array([1.0, 0.6, v, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    pr = pl.copy()
    pr[0] = 0.125
    if field != "none":
        pl[5] = pr[5] = 0.7
        pl[6] = pr[6] = 0.4
    for q in (pl, pr):
        q[8] = (GAM - 1) * q[1] * q[0] ** -GAM
    F, _ = orc.hllc_face(pl, pr, gcov, gcon, g, GAM, 1)
    up = pl if v > 0 else pr
    want, _, _ = orc.primtoflux(up, gcov, gcon, g, GAM, 1)
    np.testing.assert_allclose(F, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # HLL smears the contact: its mass flux lies strictly between the two sides'
    fl, _, _ = orc.primtoflux(pl, gcov, gcon, g, GAM, 1)
    fr, _, _ = orc.primtoflux(pr, gcov, gcon, g, GAM, 1)
    ul, _, _ = orc.primtoflux(pl, gcov, gcon, g, GAM, 0)
    ur, _, _ = orc.primtoflux(pr, gcov, gcon, g, GAM, 0)
    _, cmaxl, cminl = orc.primtoflux(pl, gcov, gcon, g, GAM, 1)
    _, cmaxr, cminr = orc.primtoflux(pr, gcov, gcon, g, GAM, 1)
    cmax = abs(max(0.0, cmaxl, cmaxr))
    cmin = abs(max(0.0, -cminl, -cminr))
    hll = (cmax * fl + cmin * fr - cmax * cmin * (ur - ul)) / (cmax + cmin + 1e-20)
    assert abs(hll[0] - want[0]) > 1e-3 * abs(want[0])


def test_stationary_contact_falls_back_to_hll(orc):
    gcov, gcon, g = _flat()
    pl = np.array([1.0, 0.6, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    pr = pl.copy()
    pr[0] = 0.125
    for q in (pl, pr):
        q[8] = (GAM - 1) * q[1] * q[0] ** -GAM
    F, _ = orc.hllc_face(pl, pr, gcov, gcon, g, GAM, 1)
    fl, cmaxl, cminl = orc.primtoflux(pl, gcov, gcon, g, GAM, 1)
    fr, cmaxr, cminr = orc.primtoflux(pr, gcov, gcon, g, GAM, 1)
    ul, _, _ = orc.primtoflux(pl, gcov, gcon, g, GAM, 0)
    ur, _, _ = orc.primtoflux(pr, gcov, gcon, g, GAM, 0)
    cmax = abs(max(0.0, cmaxl, cmaxr))
    cmin = abs(max(0.0, -cminl, -cminr))
    hll = (cmax * fl + cmin * fr - cmax * cmin * (ur - ul)) / (cmax + cmin + 1e-20)
    np.testing.assert_allclose(F, hll, rtol=0, atol=1e-13 * np.abs(hll).max())
    assert F[0] != 0.0  # HLL's diffusive mass flux, not the exact 0 of the contact


def test_supersonic_face_is_upwind(orc):
    gcov, gcon, g = _flat()
    pl = np.array([1.0, 0.01, 3.0, 0.2, 0.0, 0.1, 0.05, 0.0, 0.0])
    pr = np.array([0.5, 0.02, 2.5, 0.0, 0.1, 0.1, 0.0, 0.05, 0.0])
    for q in (pl, pr):
        q[8] = (GAM - 1) * q[1] * q[0] ** -GAM
    F, _ = orc.hllc_face(pl, pr, gcov, gcon, g, GAM, 1)
    want, _, _ = orc.primtoflux(pl, gcov, gcon, g, GAM, 1)
    np.testing.assert_allclose(F, want, rtol=0, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("field", ["none", "normal"])
def test_star_state_carries_the_upwind_specific_quantities(orc, field):
    """A shock tube with transverse velocities (flat space; the contact moves
    right).  Across the left wave the Rankine-Hugoniot conditions give a star
    state whose rho, entropy tracer and (without field) transverse momenta are
    the left state's scaled by one factor, all moving with the contact: the
    flux ratios F_K / F_rho and (hydro) F_t / F_rho equal the left state's
    U_K / U_rho and U_t / U_rho."""
    gcov, gcon, g = _flat()
    pl = np.array([1.0, 2.0, 0.1, 0.3, -0.2, 0.0, 0.0, 0.0, 0.0])
    pr = np.array([0.2, 0.1, 0.0, -0.1, 0.4, 0.0, 0.0, 0.0, 0.0])
    if field != "none":
        pl[5] = pr[5] = 0.5
    for q in (pl, pr):
        q[8] = (GAM - 1) * q[1] * q[0] ** -GAM
    F, _ = orc.hllc_face(pl, pr, gcov, gcon, g, GAM, 1)
    ul, _, _ = orc.primtoflux(pl, gcov, gcon, g, GAM, 0)
    fl, _, _ = orc.primtoflux(pl, gcov, gcon, g, GAM, 1)
    fr, _, _ = orc.primtoflux(pr, gcov, gcon, g, GAM, 1)
    assert not np.allclose(F, fl) and not np.allclose(F, fr)  # a star state, not an outer one
    np.testing.assert_allclose(F[8] / F[0], ul[8] / ul[0], rtol=1e-13)
    if field == "none":
        np.testing.assert_allclose(F[3] / F[0], ul[3] / ul[0], rtol=1e-13)
        np.testing.assert_allclose(F[4] / F[0], ul[4] / ul[0], rtol=1e-13)
