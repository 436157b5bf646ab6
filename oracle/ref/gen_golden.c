/*
 * gen_golden.c -- writes golden vectors from the reference's OWN CPU C
 * (metric.c, coord.c, lu.c, phys_mhd.c, u2p_util.c, utoprim_2d.c,
 * utoprim_1dvsq2fix1.c, utoprim_1dfix1.c of tp2872/HAMR_gamma_wind), compiled
 * where the sources lie by oracle/ref/Makefile against a config.h patched to the
 * torus setup (Kerr a = 0.9375, Gamma = 13/9, 3D, CPU build).
 *
 * Test infrastructure: it runs only in a container that holds the reference, and
 * its output (tests/golden/ref_physics.json) is the committed fixture.
 *
 * Inputs are deterministic (splitmix64 from fixed seeds).  For each case it
 * records the code-coordinate metric (gcov_func, gdet_func, gcon_func,
 * conn_func), the fluid state (get_state), fluxes and conserved variables
 * (primtoflux, dirs 0-3), magnetosonic speeds (vchar, dirs 1-3) and the three
 * inversions (Utoprim_2d, Utoprim_1dvsq2fix1, Utoprim_1dfix1) started from a
 * perturbed guess, plus crafted failure cases for the inversion return codes.
 */
#include "include.h"
#include "decs_MPI.h"
#include "defs.h"

static unsigned long long rng_state;
static unsigned long long splitmix(void)
{
    unsigned long long z = (rng_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double urand(void) { return (double)(splitmix() >> 11) * (1.0 / 9007199254740992.0); }
static double nrand(void)
{
    double u1 = urand(), u2 = urand();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2. * log(u1)) * cos(2. * M_PI * u2);
}

static int first_item;
static void jarr(const char *name, const double *v, int n)
{
    int i;
    printf("%s\"%s\": [", first_item ? "" : ", ", name);
    first_item = 0;
    for (i = 0; i < n; i++) printf("%s%.17g", i ? ", " : "", v[i]);
    printf("]");
}
static void jint(const char *name, int v)
{
    printf("%s\"%s\": %d", first_item ? "" : ", ", name, v);
    first_item = 0;
}

static void pack10(double g[NDIM][NDIM], double *o)
{
    o[0] = g[0][0]; o[1] = g[0][1]; o[2] = g[0][2]; o[3] = g[0][3];
    o[4] = g[1][1]; o[5] = g[1][2]; o[6] = g[1][3];
    o[7] = g[2][2]; o[8] = g[2][3]; o[9] = g[3][3];
}

static void emit_case(double *X, double *pr, int fail_kind)
{
    double gcov[NDIM][NDIM], gcon[NDIM][NDIM], conn[NDIM][NDIM][NDIM], g10[10], c10[10], conn_flat[64];
    double r, th, ph, flux[4][NPR], vmax[4], vmin[4], U[NPR], prim[NPR], guess[NPR];
    struct of_geom geom;
    struct of_state q;
    int i, j, k, d, ret;
    bl_coord(X, &r, &th, &ph);
    gcov_func(X, gcov);
    geom.g = gdet_func(gcov);
    gcon_func(gcov, gcon);
    for (i = 0; i < NDIM; i++)
        for (j = 0; j < NDIM; j++) { geom.gcov[i][j] = gcov[i][j]; geom.gcon[i][j] = gcon[i][j]; }
    conn_func(X, &geom, conn);
    for (i = 0; i < 4; i++)
        for (j = 0; j < 4; j++)
            for (k = 0; k < 4; k++) conn_flat[(i * 4 + j) * 4 + k] = conn[i][j][k];
    get_state(pr, &geom, &q);
    for (d = 0; d < 4; d++) primtoflux(pr, &q, NULL, d, &geom, flux[d]);
    vmax[0] = vmin[0] = 0.;
    for (d = 1; d < 4; d++) vchar(pr, &q, &geom, d, &vmax[d], &vmin[d]);

    printf("{");
    first_item = 1;
    jarr("X", X, 4);
    {
        double v[3] = {r, th, ph};
        jarr("rthph", v, 3);
    }
    pack10(gcov, g10);
    jarr("gcov", g10, 10);
    pack10(gcon, c10);
    jarr("gcon", c10, 10);
    jarr("gdet", &geom.g, 1);
    jarr("conn", conn_flat, 64);
    jarr("prim", pr, NPR);
    jarr("ucon", q.ucon, 4);
    jarr("ucov", q.ucov, 4);
    jarr("bcon", q.bcon, 4);
    jarr("bcov", q.bcov, 4);
    jarr("flux0", flux[0], NPR);
    jarr("flux1", flux[1], NPR);
    jarr("flux2", flux[2], NPR);
    jarr("flux3", flux[3], NPR);
    jarr("vmax", vmax + 1, 3);
    jarr("vmin", vmin + 1, 3);

    /* conserved state handed to the inversions */
    for (k = 0; k < NPR; k++) U[k] = flux[0][k];
    if (fail_kind == 1) U[0] = -fabs(U[0]);                 /* D <= 0 */
    if (fail_kind == 2) U[UU] = U[RHO] + 0.5 * (U[UU] - U[RHO]);     /* total energy halved */
    if (fail_kind == 5) U[UU] = U[RHO] + 0.97 * (U[UU] - U[RHO]);    /* small energy deficit */
    if (fail_kind == 6) { U[B1] *= 30.; U[B2] *= 30.; U[B3] *= 30.; } /* magnetically dominated */
    if (fail_kind == 3) { U[U1] *= 50.; U[U2] *= 50.; U[U3] *= 50.; }                          /* momentum excess */
    jint("fail_kind", fail_kind);
    jarr("U", U, NPR);
    for (k = 0; k < NPR; k++) guess[k] = pr[k];
    guess[RHO] *= 1.003; guess[UU] *= 0.995; guess[U1] += 1e-3; guess[U2] -= 2e-3; guess[U3] += 1.5e-3;
    if (fail_kind == 4) { guess[U1] = 1e11; }                                                   /* utsq too big */
    jarr("guess", guess, NPR);
    for (k = 0; k < NPR; k++) prim[k] = guess[k];
    ret = Utoprim_2d(U, gcov, gcon, geom.g, prim, NEWT_TOL, BASIC);
    jint("ret2d", ret);
    jarr("prim2d", prim, NPR);
    for (k = 0; k < NPR; k++) prim[k] = guess[k];
    ret = Utoprim_1dvsq2fix1(U, gcov, gcon, geom.g, prim, NEWT_TOL, BASIC);
    jint("ret1dvsq", ret);
    jarr("prim1dvsq", prim, NPR);
    for (k = 0; k < NPR; k++) prim[k] = guess[k];
    ret = Utoprim_1dfix1(U, gcov, gcon, geom.g, prim, NEWT_TOL, BASIC);
    jint("ret1dfix", ret);
    jarr("prim1dfix", prim, NPR);
    printf("}");
}

/* ---- crafted inversion failures (gen_golden failures) ----
 * Random states as in the main cases, then conserved variables and guesses
 * distorted at random (energy, momentum and field scaled, the guess's velocity
 * inflated, the density scaled or made non-positive, a guess with a huge
 * Lorentz factor and internal energy), kept when Utoprim_2d returns a code not yet
 * collected often enough, or when the cascade (2D, then 1D v^2, then 1D)
 * takes a path not yet collected.  Only finite inputs. */
#define FKINDS 24
static int fseen[FKINDS];
static int fslot(int code)
{
    switch (code) {
    case 0: return 0;
    case 1: return 1;
    case 101: return 2;
    case 201: return 3;
    case 2: return 4;
    case 3: return 5;
    case 4: return 6;
    case 5: return 7;
    case 6: return 8;
    case -100: return 9;
    default: return 10;
    }
}

static int emit_failure(double *X, double *pr, double *U, double *guess, int first, int *slots, int nslots)
{
    double gcov[NDIM][NDIM], gcon[NDIM][NDIM], g10[10], c10[10], prim[NPR], gdet;
    int k, r2, r1v, r1;
    gcov_func(X, gcov);
    gdet = gdet_func(gcov);
    gcon_func(gcov, gcon);
    for (k = 0; k < NPR; k++) prim[k] = guess[k];
    r2 = Utoprim_2d(U, gcov, gcon, gdet, prim, NEWT_TOL, BASIC);
    {
        double p1[NPR], p2[NPR];
        int want = 0, q;
        for (k = 0; k < NPR; k++) p1[k] = guess[k];
        r1v = Utoprim_1dvsq2fix1(U, gcov, gcon, gdet, p1, NEWT_TOL, BASIC);
        for (k = 0; k < NPR; k++) p2[k] = guess[k];
        r1 = Utoprim_1dfix1(U, gcov, gcon, gdet, p2, NEWT_TOL, BASIC);
        /* slot of the 2D code, and of the cascade path of a failing 2D solve */
        slots[0] = fslot(r2);
        slots[1] = r2 ? 11 + (r1v != 0) + (r1v != 0 && r1 != 0) : -1;
        slots[2] = r2 ? 14 + fslot(r1v) % 10 : -1;
        for (q = 0; q < nslots; q++)
            if (slots[q] >= 0 && fseen[slots[q]] < 3) want = 1;
        if (!want) return 0;
        for (q = 0; q < nslots; q++)
            if (slots[q] >= 0) fseen[slots[q]]++;
        if (!first) printf(",\n");
        printf("{");
        first_item = 1;
        jarr("X", X, 4);
        pack10(gcov, g10);
        jarr("gcov", g10, 10);
        pack10(gcon, c10);
        jarr("gcon", c10, 10);
        jarr("gdet", &gdet, 1);
        jarr("base_prim", pr, NPR);
        jarr("U", U, NPR);
        jarr("guess", guess, NPR);
        jint("ret2d", r2);
        jarr("prim2d", prim, NPR);
        jint("ret1dvsq", r1v);
        jarr("prim1dvsq", p1, NPR);
        jint("ret1dfix", r1);
        jarr("prim1dfix", p2, NPR);
        printf("}");
        return 1;
    }
}

static int failures(void)
{
    int c, n = 0, slots[3];
    a = BH_SPIN;
    rng_state = 20261203ULL;
    printf("{\"gamma\": %.17g, \"spin\": %.17g, \"cases\": [\n", (double)GAMMA, (double)a);
    for (c = 0; c < 400000 && n < 60; c++) {
        double X[4], pr[NPR], U[NPR], guess[NPR], gcov[NDIM][NDIM], gcon[NDIM][NDIM], flux[NPR], rho, sig;
        struct of_geom geom;
        struct of_state q;
        int i, j, k, mode = (int)(urand() * 8.0);
        X[0] = 0.;
        X[1] = log(1.25) + urand() * (log(60.) - log(1.25));
        X[2] = -0.97 + 1.94 * urand();
        X[3] = 2. * M_PI * urand();
        rho = exp(-2. + 1.5 * nrand());
        sig = exp(log(0.01) + urand() * (log(100.) - log(0.01)));
        pr[RHO] = rho;
        pr[UU] = rho * exp(-2.5 + 0.5 * nrand());
        pr[U1] = 0.4 * nrand();
        pr[U2] = 0.05 * nrand();
        pr[U3] = 0.3 * nrand();
        pr[B1] = sqrt(sig * rho) * 0.3 * nrand();
        pr[B2] = sqrt(sig * rho) * 0.05 * nrand();
        pr[B3] = sqrt(sig * rho) * 0.3 * nrand();
        pr[KTOT] = 0.;
        gcov_func(X, gcov);
        geom.g = gdet_func(gcov);
        gcon_func(gcov, gcon);
        for (i = 0; i < NDIM; i++)
            for (j = 0; j < NDIM; j++) { geom.gcov[i][j] = gcov[i][j]; geom.gcon[i][j] = gcon[i][j]; }
        get_state(pr, &geom, &q);
        primtoflux(pr, &q, NULL, 0, &geom, flux);
        for (k = 0; k < NPR; k++) U[k] = flux[k];
        for (k = 0; k < NPR; k++) guess[k] = pr[k];
        guess[RHO] *= 1.003; guess[UU] *= 0.995; guess[U1] += 1e-3; guess[U2] -= 2e-3; guess[U3] += 1.5e-3;
        if (mode == 0) U[UU] = U[RHO] + (-3. + 6. * urand()) * (U[UU] - U[RHO]);
        if (mode == 1) { double m = exp(urand() * log(1e4)); U[U1] *= m; U[U2] *= m; U[U3] *= m; }
        if (mode == 2) { double b = exp(urand() * log(1e4)); U[B1] *= b; U[B2] *= b; U[B3] *= b; }
        if (mode == 3) { double v = exp(urand() * log(1e11)); guess[U1] *= v; guess[U2] *= v; guess[U3] *= v; }
        if (mode == 4) { double d = exp(-20. + 40. * urand()); U[RHO] *= d; }
        if (mode == 5) {
            double m = exp(urand() * log(1e3)), e = -2. + 4. * urand();
            U[UU] = U[RHO] + e * (U[UU] - U[RHO]);
            U[U1] *= m; U[U2] *= m; U[U3] *= m;
        }
        if (mode == 6) {
            /* the guess's 4-velocity near UTSQ_TOO_BIG with a large internal
             * energy: the starting W exceeds W_TOO_BIG */
            double want = exp(log(10.) * (18.5 + 1.4 * urand())), usq = 0., sc;
            for (i = 1; i < 4; i++)
                for (j = 1; j < 4; j++) usq += gcov[i][j] * guess[U1 + i - 1] * guess[U1 + j - 1];
            sc = sqrt(want / (usq + 1e-300));
            guess[U1] *= sc; guess[U2] *= sc; guess[U3] *= sc;
            guess[UU] = exp(log(10.) * (-1. + 4. * urand()));
        }
        if (mode == 7) U[RHO] = -fabs(U[RHO]) * urand();
        n += emit_failure(X, pr, U, guess, n == 0, slots, 3);
    }
    printf("\n], \"searched\": %d}\n", c);
    return 0;
}

int main(int argc, char **argv)
{
    int c, ncase = 0;
    if (argc > 1 && !strcmp(argv[1], "failures")) return failures();
    a = BH_SPIN;
    rng_state = 20260412ULL;
    printf("{\"gamma\": %.17g, \"spin\": %.17g, \"cases\": [\n", (double)GAMMA, (double)a);
    for (c = 0; c < 54; c++) {
        double X[4], pr[NPR], rho, sig;
        int fail_kind = (c >= 40) ? 1 + (c - 40) % 7 : 0;
        X[0] = 0.;
        X[1] = log(1.25) + urand() * (log(60.) - log(1.25));
        X[2] = -0.97 + 1.94 * urand();
        X[3] = 2. * M_PI * urand();
        rho = exp(-2. + 1.5 * nrand());
        sig = (c % 3 == 0) ? 0.02 : (c % 3 == 1) ? 1.0 : 20.0;   /* b^2/rho scale */
        pr[RHO] = rho;
        pr[UU] = rho * exp(-2.5 + 0.5 * nrand());
        pr[U1] = 0.4 * nrand();
        pr[U2] = 0.05 * nrand();
        pr[U3] = 0.3 * nrand();
        pr[B1] = sqrt(sig * rho) * 0.3 * nrand();
        pr[B2] = sqrt(sig * rho) * 0.05 * nrand();
        pr[B3] = sqrt(sig * rho) * 0.3 * nrand();
        pr[KTOT] = 0.;
        if (ncase++) printf(",\n");
        emit_case(X, pr, fail_kind);
    }
    printf("\n]}\n");
    return 0;
}
