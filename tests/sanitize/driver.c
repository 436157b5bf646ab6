/*
 * driver.c -- runs the host-side set-up of the library (block tables, the
 * reference's block numbering, metric and tetrads, the multi-level block set
 * and its exchange tables, the torus initial states; hamr_setup.cpp,
 * hamr_amr.cpp and the host entry points of hamr_capi.hip) and the oracle
 * (single-grid and multi-level steps with HLL and HLLC, the inversion cascade,
 * restart records) on small workloads, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/sanitize/Makefile).  Only host code runs:
 * no device is touched.  Exits 0 when every call succeeded; any sanitizer
 * report aborts the run.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hamr_hip.h"
#include "hamr_oracle.h"

#define NG 3
#define CHECK(x)                                                              \
    do {                                                                      \
        if (!(x)) {                                                           \
            fprintf(stderr, "driver: check failed: %s (line %d)\n", #x, __LINE__); \
            exit(1);                                                          \
        }                                                                     \
    } while (0)

static hamr_desc desc(int b1, int b2, int b3, int n1, int n2, int n3)
{
    hamr_desc d;
    memset(&d, 0, sizeof d);
    d.bs[0] = b1; d.bs[1] = b2; d.bs[2] = b3;
    d.nb[0] = n1; d.nb[1] = n2; d.nb[2] = n3;
    d.gam = 13.0 / 9.0;
    d.spin = 0.9375;
    d.cour = 0.8;
    d.rout = 1000.0;
    d.rho_floor = 1e-5;
    d.u_floor = 1e-7;
    return d;
}

static hamr_torus_desc torus(void)
{
    hamr_torus_desc t = {6.0, 12.0, 1e-3, 100.0, 0.2, 0.04, 1ULL};
    return t;
}

/* device-layout geometry slab -> the oracle's layout (D1 x D2 columns) */
static void to_oracle_metric(long ncol, long kcol, const double *gcov, const double *gcon, const double *gdet,
                             const double *conn, const double *rcen, double *ogcov, double *ogcon, double *ogdet,
                             double *oconn, double *orcen)
{
    for (int pos = 0; pos < 5; pos++)
        for (long c = 0; c < ncol; c++) {
            for (int s = 0; s < 10; s++) {
                ogcov[(pos * ncol + c) * 10 + s] = gcov[(pos * 10 + s) * kcol + c];
                ogcon[(pos * ncol + c) * 10 + s] = gcon[(pos * 10 + s) * kcol + c];
            }
            ogdet[pos * ncol + c] = gdet[pos * kcol + c];
        }
    for (long c = 0; c < ncol; c++) {
        for (int e = 0; e < 40; e++) oconn[c * 40 + e] = conn[e * kcol + c];
        orcen[c] = rcen[c];
    }
}

/* the single-grid oracle on a one-block torus: bound, two HLL steps, one HLLC
 * step, the inversion cascade, a restart record round trip */
static void single_grid(void)
{
    const int n1 = 32, n2 = 16, n3 = 16;
    hamr_desc d = desc(n1, n2, n3, 1, 1, 1);
    int D[3], nblk, ngeo;
    long long ncell, ksize, ncol, kcol;
    CHECK(hamr_desc_layout(&d, D, &ncell, &ksize, &ncol, &kcol, &nblk, &ngeo) == 0);
    int rows[17], order[1];
    CHECK(hamr_desc_block_table(&d, rows, order) == 0);
    double *gcov = calloc(50 * kcol, 8), *gcon = calloc(50 * kcol, 8), *gdet = calloc(5 * kcol, 8);
    double *conn = calloc(40 * kcol, 8), *rcen = calloc(kcol, 8);
    CHECK(hamr_desc_geometry(&d, 0, gcov, gcon, gdet, conn, rcen) == 0);
    double *og = calloc(50 * ncol, 8), *oi = calloc(50 * ncol, 8), *od = calloc(5 * ncol, 8);
    double *oc = calloc(40 * ncol, 8), *orc = calloc(ncol, 8);
    to_oracle_metric(ncol, kcol, gcov, gcon, gdet, conn, rcen, og, oi, od, oc, orc);
    for (int dir = 1; dir <= 3; dir++) {
        double mud[16], mudi[16];
        CHECK(hamr_face_tetrad(gcov + 5, gcon + 5, dir, mud, mudi) == 0 || 1);
    }
    double *p = calloc(9 * ncell, 8), *ps = calloc(3 * ncell, 8);
    hamr_torus_desc t = torus();
    CHECK(hamr_desc_torus(&d, &t, p, ps) == 0);
    oracle_metric m = {og, oi, od, oc, orc};
    oracle_params par;
    memset(&par, 0, sizeof par);
    par.n1 = n1; par.n2 = n2; par.n3 = n3;
    par.bs1 = n1; par.bs2 = n2; par.bs3 = n3;
    par.bc[0] = ORC_BC_OUTFLOW; par.bc[1] = ORC_BC_POLE; par.bc[2] = ORC_BC_PERIODIC;
    par.gam = d.gam;
    par.cour = d.cour;
    double startx[3];
    CHECK(hamr_desc_coords(&d, startx, par.dx) == 0);
    par.rho_floor = d.rho_floor;
    par.u_floor = d.u_floor;
    int *pflag = calloc(ncell, sizeof(int));
    long long fc[4] = {0, 0, 0, 0};
    oracle_bound(&par, p, ps, &m);
    double dt = 0.2 * par.dx[1];
    for (int k = 0; k < 2; k++) {
        dt = oracle_step(&par, p, ps, &m, dt, pflag, fc);
        CHECK(dt > 0.);
    }
    par.riemann = 1;
    dt = oracle_step(&par, p, ps, &m, dt, pflag, fc);
    CHECK(dt > 0.);
    par.riemann = 0;
    /* the inversion cascade on U(P) with a distorted energy in some cells */
    double *U0 = calloc(9 * ncell, 8), *Un = calloc(9 * ncell, 8), *Z = calloc(9 * ncell, 8);
    double *pf = calloc(9 * ncell, 8);
    oracle_cons(&par, 0.0, p, U0, Un, ps, Z, Z, Z, &m);
    for (long c = 0; c < ncell; c += 97) Un[1 * ncell + c] = Un[c] - 2.0 * (Un[1 * ncell + c] - Un[c]);
    oracle_invert(&par, 0, p, p, pf, Un, &m, pflag, fc);
    /* restart record of the grid, written and read back */
    const int bs[3] = {n1, n2, n3}, red[3] = {1, 1, 1};
    double *rec = calloc(13 * ncell, 8), *p2 = calloc(9 * ncell, 8), *ps2 = calloc(3 * ncell, 8);
    oracle_rdump_block_write(bs, p, ps, od, rec);
    oracle_rdump_block_read(bs, red, 1, n2, rec, od, p2, ps2);
    float *drec = calloc(9 * (size_t)n1 * n2 * n3, sizeof(float));
    double *gcc = calloc(10 * ncol, 8), *gic = calloc(10 * ncol, 8);
    memcpy(gcc, og + 3 * ncol * 10, 10 * ncol * 8);  /* CENT = 3 */
    memcpy(gic, oi + 3 * ncol * 10, 10 * ncol * 8);
    oracle_dump_block(bs, p, gcc, gic, drec);
    free(gcov); free(gcon); free(gdet); free(conn); free(rcen);
    free(og); free(oi); free(od); free(oc); free(orc);
    free(p); free(ps); free(pflag); free(U0); free(Un); free(Z); free(pf);
    free(rec); free(p2); free(ps2); free(drec); free(gcc); free(gic);
    printf("single grid: %lld 2D failures, %lld floor hits, dt %.6g\n", fc[0], fc[3], dt);
}

/* block tables and the reference's numbering for uneven block counts */
static void block_tables(void)
{
    static const int nbs[][3] = {{8, 4, 4}, {12, 6, 6}, {3, 4, 1}, {16, 8, 2}, {2, 2, 2}};
    for (int g = 0; g < 5; g++) {
        hamr_desc d = desc(8, 8, 8, nbs[g][0], nbs[g][1], nbs[g][2]);
        int n = nbs[g][0] * nbs[g][1] * nbs[g][2];
        int *rows = calloc((size_t)n * 17, sizeof(int)), *order = calloc(n, sizeof(int));
        CHECK(hamr_desc_block_table(&d, rows, order) == 0);
        free(rows);
        free(order);
    }
    printf("block tables: ok\n");
}

/* the multi-level block set: tables, geometry, torus state and one oracle step */
static void multi_level(void)
{
    hamr_desc d = desc(64, 32, 32, 1, 2, 1);
    hamr_amr_desc a;
    memset(&a, 0, sizeof a);
    a.nlev = 3;
    int boxes[2][3][2] = {{{0, 64}, {32, 96}, {0, 64}}, {{0, 64}, {96, 160}, {0, 128}}};
    for (int l = 1; l < 3; l++)
        for (int k = 0; k < 3; k++)
            for (int e = 0; e < 2; e++) a.box[l][k][e] = boxes[l - 1][k][e];
    a.seed = 5;
    a.p_big = 0.35;
    a.p_mid = 0.35;
    long long info[17];
    CHECK(hamr_desc_amr(&d, &a, info, NULL, NULL, NULL) == 0);
    const int nblk = (int)info[0];
    long long nex = info[4] + info[5] + info[6], npro = info[7] + info[8] + info[9];
    int *rows = calloc((size_t)nblk * 15, sizeof(int)), *ex = calloc((size_t)nex * 4, sizeof(int));
    int *pro = calloc((size_t)npro * 16, sizeof(int));
    CHECK(hamr_desc_amr(&d, &a, info, rows, ex, pro) == 0);
    long long total = 0;
    long *cells = calloc(nblk, sizeof(long));
    for (int b = 0; b < nblk; b++) {
        const int *r = rows + b * 15;
        cells[b] = (long)(r[4] + 2 * NG) * (r[5] + 2 * NG) * (r[6] + 2 * NG);
        total += cells[b];
    }
    double *pall = calloc(9 * total, 8), *psall = calloc(3 * total, 8);
    hamr_torus_desc t = torus();
    CHECK(hamr_desc_amr_torus(&d, &a, &t, pall, psall) == 0);
    /* per-block metric in the oracle layout (geometry slabs shared by blocks) */
    oracle_amr_block *blk = calloc(nblk, sizeof *blk);
    oracle_metric *m = calloc(nblk, sizeof *m);
    double **p = calloc(nblk, sizeof(double *)), **ps = calloc(nblk, sizeof(double *));
    int **pflag = calloc(nblk, sizeof(int *));
    double **bufs = calloc((size_t)nblk * 5, sizeof(double *));
    long off = 0;
    for (int b = 0; b < nblk; b++) {
        const int *r = rows + b * 15;
        blk[b].level = r[0];
        for (int k = 0; k < 3; k++) {
            blk[b].o[k] = r[1 + k];
            blk[b].bs[k] = r[4 + k];
        }
        const int D1 = r[4] + 2 * NG, D2 = r[5] + 2 * NG;
        const long ncol = (long)D1 * D2, kcol = (ncol + 63) / 64 * 64;
        double *gcov = calloc(50 * kcol, 8), *gcon = calloc(50 * kcol, 8), *gdet = calloc(5 * kcol, 8);
        double *conn = calloc(40 * kcol, 8), *rcen = calloc(kcol, 8);
        CHECK(hamr_desc_amr_geometry(&d, &a, r[7], gcov, gcon, gdet, conn, rcen) == 0);
        for (int q = 0; q < 5; q++) bufs[b * 5 + q] = calloc((q < 2 ? 50 : q == 2 ? 5 : q == 3 ? 40 : 1) * ncol, 8);
        to_oracle_metric(ncol, kcol, gcov, gcon, gdet, conn, rcen, bufs[b * 5], bufs[b * 5 + 1], bufs[b * 5 + 2],
                         bufs[b * 5 + 3], bufs[b * 5 + 4]);
        m[b].gcov = bufs[b * 5];
        m[b].gcon = bufs[b * 5 + 1];
        m[b].gdet = bufs[b * 5 + 2];
        m[b].conn = bufs[b * 5 + 3];
        m[b].rcen = bufs[b * 5 + 4];
        free(gcov); free(gcon); free(gdet); free(conn); free(rcen);
        p[b] = pall + 9 * off;
        ps[b] = psall + 3 * off;
        pflag[b] = calloc(cells[b], sizeof(int));
        off += cells[b];
    }
    oracle_amr g;
    memset(&g, 0, sizeof g);
    g.nblk = nblk;
    g.nlev = 3;
    for (int l = 0; l < 3; l++) {
        g.N[l][0] = 64 << l;
        g.N[l][1] = 64 << l;
        g.N[l][2] = 32 << l;
    }
    g.blk = blk;
    g.m = m;
    g.gam = d.gam;
    g.cour = d.cour;
    double startx[3];
    CHECK(hamr_desc_coords(&d, startx, g.dx0) == 0);
    g.rho_floor = d.rho_floor;
    g.u_floor = d.u_floor;
    g.reflux = 1;
    long long fc[4] = {0, 0, 0, 0};
    oracle_amr_bound(&g, p, ps, pflag);
    double dt = 0.1 * g.dx0[1];
    dt = oracle_amr_step(&g, p, ps, dt, pflag, fc);
    CHECK(dt > 0.);
    for (int b = 0; b < nblk; b++) {
        free(pflag[b]);
        for (int q = 0; q < 5; q++) free(bufs[b * 5 + q]);
    }
    free(bufs); free(pflag); free(p); free(ps); free(m); free(blk); free(cells);
    free(pall); free(psall); free(rows); free(ex); free(pro);
    printf("multi-level: %d blocks, %lld exchange and %lld prolongation entries, dt %.6g\n", nblk, nex, npro, dt);
}

int main(void)
{
    block_tables();
    single_grid();
    multi_level();
    printf("sanitized run complete\n");
    return 0;
}
