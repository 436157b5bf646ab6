"""Pins the FP64 oracle's pointwise physics.

Against golden vectors written by the reference's own CPU C (compiled in this
container by oracle/ref/Makefile, fixture tests/golden/ref_physics.json):
get_state, primtoflux (dirs 0-3), the magnetosonic speeds of vchar, and the
three Noble inversions including their return codes.  The geometric source term
has no reference golden vector (phys_mhd.c:158 reads global grid arrays); it is
pinned against an independent full-tensor contraction built from the golden
connection.  PPM reconstruction has no CPU reference (main.c:595 forbids PPM on
the CPU path); it is pinned by closed forms.
"""
import json
import os

import numpy as np
import pytest

from tests.oracle_ctypes import NPR

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_physics.json")
PAIRS = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


def unpack10(v):
    g = np.zeros((4, 4))
    for s, (a, b) in enumerate(PAIRS):
        g[a, b] = g[b, a] = v[s]
    return g


def conn40_from_full(conn64):
    """Golden conn is Gamma^i_{jk} as conn[i][j][k]; pack to [lam*10 + pair]."""
    c = np.array(conn64).reshape(4, 4, 4)
    out = np.zeros(40)
    for lam in range(4):
        for s, (a, b) in enumerate(PAIRS):
            out[lam * 10 + s] = c[lam, a, b]
    return out


def relerr(a, b, scale=None):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    s = np.max(np.abs(b)) if scale is None else scale
    return np.max(np.abs(a - b)) / max(s, 1e-300)


def test_golden_metadata(gold):
    assert abs(gold["gamma"] - 13.0 / 9.0) < 1e-15
    assert gold["spin"] == 0.9375
    assert len(gold["cases"]) >= 50


def test_get_state_matches_reference(oracle, gold):
    for c in gold["cases"]:
        ucon, ucov, bcon, bcov = oracle.get_state(c["prim"], c["gcov"], c["gcon"])
        assert relerr(ucon, c["ucon"]) < 1e-13
        assert relerr(ucov, c["ucov"]) < 1e-13
        assert relerr(bcon, c["bcon"], np.max(np.abs(c["bcon"])) + 1e-300) < 1e-12
        assert relerr(bcov, c["bcov"], np.max(np.abs(c["bcov"])) + 1e-300) < 1e-12


def test_primtoflux_matches_reference(oracle, gold):
    gam = gold["gamma"]
    for c in gold["cases"]:
        g = c["gdet"][0]
        for d in range(4):
            f, vmax, vmin = oracle.primtoflux(c["prim"], c["gcov"], c["gcon"], g, gam, d)
            ref = np.array(c["flux%d" % d])
            # per-component tolerance relative to the flux scale of this direction
            assert relerr(f, ref) < 1e-12, (d, f, ref)
            if d > 0:
                assert abs(vmax - c["vmax"][d - 1]) <= 1e-12 * max(1.0, abs(c["vmax"][d - 1]))
                assert abs(vmin - c["vmin"][d - 1]) <= 1e-12 * max(1.0, abs(c["vmin"][d - 1]))


def test_source_against_full_contraction(oracle, gold):
    """dU_nu = gdet T^kap_lam Gamma^lam_{nu kap} with T from the golden state."""
    gam = gold["gamma"]
    for c in gold["cases"]:
        pr = np.array(c["prim"])
        ucon, ucov, bcon, bcov = (np.array(c[k]) for k in ("ucon", "ucov", "bcon", "bcov"))
        P = (gam - 1.0) * pr[1]
        bsq = bcon @ bcov
        T = (pr[0] + pr[1] + P + bsq) * np.outer(ucon, ucov) + (P + 0.5 * bsq) * np.eye(4) - np.outer(bcon, bcov)
        G = np.array(c["conn"]).reshape(4, 4, 4)  # G[lam, nu, kap]
        want = c["gdet"][0] * np.einsum("kl,lnk->n", T, G)
        dU = oracle.source(pr, c["gcov"], c["gcon"], c["gdet"][0], conn40_from_full(c["conn"]), gam)
        assert dU[0] == 0.0 and np.all(dU[5:] == 0.0)
        # roundoff scale: magnitudes of every product entering the contraction
        # (T itself cancels strongly when b^2/rho is large)
        Tabs = (abs(pr[0] + pr[1] + P + bsq) * np.outer(np.abs(ucon), np.abs(ucov))
                + abs(P + 0.5 * bsq) * np.eye(4) + np.outer(np.abs(bcon), np.abs(bcov)))
        scale = c["gdet"][0] * np.einsum("kl,lnk->n", Tabs, np.abs(G))
        assert np.all(np.abs(dU[1:5] - want) <= 1e-13 * scale + 1e-300), (dU[1:5], want, scale)


@pytest.mark.parametrize("which,key", [(0, "2d"), (1, "1dvsq"), (2, "1dfix")])
def test_inversions_match_reference(oracle, gold, which, key):
    gam = gold["gamma"]
    ncodes = set()
    for c in gold["cases"]:
        ret, prim = oracle.utoprim(which, c["U"], c["gcov"], c["gcon"], c["gdet"][0], gam, c["guess"])
        assert ret == c["ret" + key], (which, c["fail_kind"], ret, c["ret" + key])
        ncodes.add(ret)
        ref = np.array(c["prim" + key])
        if ret == 0:
            assert relerr(prim[:5], ref[:5]) < 1e-10, (prim, ref)
        else:
            # on failure only the face fields are rewritten; the guess survives
            assert np.array_equal(prim[:5], np.array(c["guess"])[:5])
        if ret != -100:
            assert relerr(prim[5:8], ref[5:8], np.max(np.abs(ref[5:8])) + 1e-300) < 1e-14
    assert len(ncodes) >= 3


def test_inversion_round_trip(oracle, gold):
    """P -> U -> P recovers the primitive state (converged cases)."""
    gam = gold["gamma"]
    for c in gold["cases"]:
        if c["fail_kind"] != 0:
            continue
        U = np.array(c["flux0"])
        ret, prim = oracle.utoprim(0, U, c["gcov"], c["gcon"], c["gdet"][0], gam, c["guess"])
        assert ret == 0
        assert relerr(prim[:8], np.array(c["prim"])[:8]) < 1e-9


def test_ppm_closed_forms(oracle):
    # constants are preserved
    l, r = oracle.para([2.5] * 5)
    assert l == 2.5 and r == 2.5
    # linear data: the edges are the linear interpolants
    for a, b in [(1.0, 0.3), (-2.0, 1e-3), (5.0, -7.0)]:
        y = [a + b * x for x in (-2, -1, 0, 1, 2)]
        l, r = oracle.para(y)
        assert abs(l - (a - 0.5 * b)) <= 4e-16 * max(1, abs(a) + abs(b))
        assert abs(r - (a + 0.5 * b)) <= 4e-16 * max(1, abs(a) + abs(b))
    # smooth monotone data where the limiter is inactive: 4th-order edge formula
    y = [np.exp(0.1 * x) for x in (-2, -1, 0, 1, 2)]
    l, r = oracle.para(y)
    assert abs(l - (7 * (y[1] + y[2]) - (y[0] + y[3])) / 12) < 1e-15
    assert abs(r - (7 * (y[2] + y[3]) - (y[1] + y[4])) / 12) < 1e-15
    # local extremum: flattened to the cell value
    l, r = oracle.para([0.0, 1.0, 2.0, 1.0, 0.0])
    assert l == 2.0 and r == 2.0
    # monotonicity: edges stay within the neighbouring values
    rng = np.random.default_rng(7)
    for _ in range(2000):
        y = np.cumsum(rng.exponential(size=5)) * rng.choice([-1, 1])
        if rng.random() < 0.3:
            y[rng.integers(5)] += rng.normal() * 5
        l, r = oracle.para(y)
        lo = min(y[1], y[2])
        hi = max(y[1], y[2])
        assert lo - 1e-12 * abs(hi) <= l <= hi + 1e-12 * abs(hi)
        lo = min(y[2], y[3])
        hi = max(y[2], y[3])
        assert lo - 1e-12 * abs(hi) <= r <= hi + 1e-12 * abs(hi)


def test_ppm_mutation_is_caught(oracle):
    """A PPM edge that drops the slope correction would fail the 4th-order check."""
    y = [np.exp(0.1 * x) for x in (-2, -1, 0, 1, 2)]
    l, _ = oracle.para(y)
    assert abs(l - 0.5 * (y[1] + y[2])) > 1e-5


@pytest.fixture(scope="module")
def failures():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "ref_failures.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("which,key", [(0, "2d"), (1, "1dvsq"), (2, "1dfix")])
def test_inversions_in_cpu_arithmetic_match_reference_bitwise(failures, gold, which, key):
    """With the reference CPU file's arithmetic (u~^2 as its nine-term loop, no
    fma in func_vsq; ORACLE_CPU_ARITH) the restatement reproduces every return
    code of the compiled reference on the regular and the crafted failure
    cases, and its primitives to 1e-13 (the remaining differences: libm's pow
    and sqrt against the same calls)."""
    from tests import oracle_ctypes as O

    cpu = O.load_cpu_arith()
    worst = 0.0
    for data in (gold, failures):
        gam = data["gamma"]
        for c in data["cases"]:
            ret, prim = cpu.utoprim(which, c["U"], c["gcov"], c["gcon"], c["gdet"][0], gam, c["guess"])
            assert ret == c["ret" + key], (which, ret, c["ret" + key])
            ref = np.array(c["prim" + key])
            if ret == 0:
                worst = max(worst, relerr(prim[:5], ref[:5]))
            elif ret != -100:
                assert np.array_equal(prim[:5], ref[:5])
    print("worst relative primitive error %.2e" % worst)
    assert worst < 1e-13


@pytest.mark.parametrize("which,key", [(0, "2d"), (1, "1dvsq"), (2, "1dfix")])
def test_inversion_failure_cases_match_reference(oracle, failures, which, key):
    """Crafted conserved states and guesses (oracle/ref/gen_golden.c failures:
    energy, momentum, field or density distorted, guesses with huge Lorentz
    factors) on which the reference's own Utoprim_2d, Utoprim_1dvsq2fix1 and
    Utoprim_1dfix1 return their failure codes: the oracle returns the same
    code and leaves the same primitives.  Codes covered for the 2D inversion:
    0, 2 (guess's u^2 too big), 6 (negative u), 101 (Newton not converged),
    201 (non-finite residual), -100 (D <= 0); cascades 2D fail -> 1D v^2 ok,
    2D and 1D v^2 fail -> 1D ok, all three fail.  Codes 1, 3, 4 and 5 of the
    2D inversion (utoprim_2d.c:308-356) are not reachable from finite inputs
    in 400,000 draws: validate_x (utoprim_2d.c:417-427) clamps v^2 below 1
    (code 4) and keeps W below W_TOO_BIG = 1e20 unless the start exceeds it,
    which the Newton steps then leave (code 3); rho0 = D sqrt(1 - v^2) > 0
    when D > 0 and v^2 < 1 (code 5); code 1 needs W == FAIL_VAL = 1e30.
    One crafted input (a guess with Lorentz factor ~1e8) sits where the
    reference's CPU and GPU paths part ways (the CPU file's residual
    overflows, code 201; with the GPU path's u~^2 order and fma the 2D
    Newton converges): the oracle, which follows the GPU path, is checked
    there against the CPU-arithmetic build instead."""
    from tests import oracle_ctypes as O

    cpu = O.load_cpu_arith()
    gam = failures["gamma"]
    codes = set()
    differ = 0
    for c in failures["cases"]:
        ret, prim = oracle.utoprim(which, c["U"], c["gcov"], c["gcon"], c["gdet"][0], gam, c["guess"])
        codes.add(ret)
        if ret != c["ret" + key]:
            # the reference's two paths part ways on this input: the CPU file's
            # arithmetic reproduces the golden code (checked bitwise in
            # test_inversions_in_cpu_arithmetic_match_reference_bitwise), the
            # GPU path's arithmetic, which the library follows, does not
            assert cpu.utoprim(which, c["U"], c["gcov"], c["gcon"], c["gdet"][0], gam, c["guess"])[0] == \
                c["ret" + key]
            differ += 1
            continue
        ref = np.array(c["prim" + key])
        if ret == 0:
            assert relerr(prim[:5], ref[:5]) < 1e-10, (prim, ref)
        else:
            assert np.array_equal(prim[:5], ref[:5]), (prim, ref)
        if ret != -100:
            assert relerr(prim[5:8], ref[5:8], np.max(np.abs(ref[5:8])) + 1e-300) < 1e-14
    assert differ <= 1, differ
    if which == 0:
        assert {0, 2, 6, 101, 201, -100} <= codes, codes
    paths = {(c["ret2d"] != 0, c["ret1dvsq"] != 0, c["ret1dfix"] != 0) for c in failures["cases"]}
    assert {(True, False, False), (True, True, False), (True, True, True)} <= paths
