/*
 * hamr_amr_oracle.c -- FP64 CPU restatement of the step on a multi-level set
 * of blocks of uneven sizes (level ratio 2).
 *
 * TEST INFRASTRUCTURE ONLY (what the HIP step is tested against; see hamr_oracle.h).
 *
 * Each block is advanced by the single-grid stage functions of hamr_oracle.c on
 * its own ghost-inclusive arrays (its level's spacing, its metric); the polar
 * E1 ring average runs over the blocks of each ring (E_average,
 * const_trans.c:268-543) and the ghost fill is the multi-level exchange below.
 * Everything is found by searching the block list (which leaf holds a
 * position, which block is a polar block's partner, which blocks form a ring),
 * independently of the device tables.
 *
 * Ghost fill, in the order of bound_prim (bounds.c:11-105), stage d filling the
 * ghosts along d over the full extent of the earlier directions and the
 * interior of the later ones:
 *   1. radial outflow faces (inflow check at the ghost's geometry), then the
 *      exchange of the internal x1 faces;
 *   2. polar faces (mirror from the pi-shifted partner), then the exchange of
 *      the internal x2 faces;
 *   3. the exchange of the x3 faces (periodic);
 *   (in stages 2 and 3 a corner ghost, beyond the block along an earlier
 *   direction, is a copy of the storage cell at Y of the same-level leaf
 *   across the face that holds the position clamped into the block's range,
 *   when there is one: that leaf's ghost of the earlier stage, its own
 *   boundary faces included);
 *   4. radial outflow again over every row except the polar ghost rows (the
 *      radial ghosts of stages 2 and 3 are the outflow image of the block's own
 *      filled edge cells).
 * Exchange of a ghost cell at position Y (cells of its block's level l, x3
 * wrapped): in a polar ghost row (stage 3), a copy of the same-level leaf
 * holding the row's clamped position, at Y; beyond the radial edge along an
 * earlier direction, nothing (step 4); otherwise, for the leaf L holding Y:
 *   level l    : copy of the cell;
 *   level l+1  : restriction, the mean of the 2x2x2 children (packsendaverage1,
 *                GPU_program2.cu:154-197); face fields: the mean of the four
 *                children's faces that tile the coarse face;
 *   level l-1  : prolongation (unpackreceivecoarse1, GPU_program2.cu:615-803):
 *                coarse value + 0.25 s_d dq_d summed over d, s_d = -1 on the
 *                lower and +1 on the upper child, dq_d the MC slope
 *                (slope_lim_2, GPU_program2.cu:10-45) of the coarse cell and
 *                its two neighbours along d (the neighbour's coarse-level value:
 *                a cell, the mean of finer children, or the cell of a coarser
 *                leaf; beyond the radial edge or the axis: no slope); a face
 *                field along a has no slope term along a and starts from the
 *                coarse face (even child) or the coarse cell-centred field (odd
 *                child, the face that halves the coarse cell).
 * pflag of a ghost: the source cell's (first child / coarse cell).
 */
#include "hamr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NG ORC_NG
#define NPR ORC_NPR
enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7, KTOT = 8 };
enum { FACE1 = 0, FACE2 = 1, CORN = 2, CENT = 3, FACE3 = 4 };
#define NV (NPR + 3)

typedef struct {
    int D[3];
    long ncell;
} bgeo;

static bgeo geo_of(const oracle_amr_block *B)
{
    bgeo r;
    for (int k = 0; k < 3; k++) r.D[k] = B->bs[k] + 2 * NG;
    r.ncell = (long)r.D[0] * r.D[1] * r.D[2];
    return r;
}

static long cell_local(const oracle_amr_block *B, int i, int j, int z)
{
    return ((long)i * (B->bs[1] + 2 * NG) + j) * (B->bs[2] + 2 * NG) + z;
}

/* cell of block B at position Y of B's level (may be a ghost) */
static long cell_at(const oracle_amr_block *B, const int Y[3])
{
    return cell_local(B, Y[0] - B->o[0] + NG, Y[1] - B->o[1] + NG, Y[2] - B->o[2] + NG);
}

static int inside(const oracle_amr_block *B, const int Y[3])
{
    for (int k = 0; k < 3; k++)
        if (Y[k] < B->o[k] || Y[k] >= B->o[k] + B->bs[k]) return 0;
    return 1;
}

/* the leaf holding cell Y of level l (Y inside the grid); *lc = its level */
static int find_leaf(const oracle_amr *g, int l, const int Y[3], int *lc)
{
    for (int b = 0; b < g->nblk; b++) {
        const oracle_amr_block *B = &g->blk[b];
        int Z[3], s = B->level - l;
        for (int k = 0; k < 3; k++) Z[k] = s >= 0 ? (Y[k] << s) : (Y[k] >> (-s));
        if (inside(B, Z)) {
            *lc = B->level;
            return b;
        }
    }
    return -1;
}

static int wrap(int y, int n) { return ((y % n) + n) % n; }

/* all 12 values of one cell */
static void cell_values(const oracle_amr *g, int b, long c, double **p, double **ps, double *v)
{
    const long N = geo_of(&g->blk[b]).ncell;
    for (int k = 0; k < NPR; k++) v[k] = p[b][k * N + c];
    for (int a = 0; a < 3; a++) v[NPR + a] = ps[b][a * N + c];
}

/* restriction: mean of the 2x2x2 cells of block b from cell c (lowest corner) */
static void avg8_values(const oracle_amr *g, int b, long c, double **p, double **ps, double *v)
{
    const oracle_amr_block *B = &g->blk[b];
    const long N = geo_of(B).ncell, sj = B->bs[2] + 2 * NG, si = (long)(B->bs[1] + 2 * NG) * sj;
    for (int k = 0; k < NPR; k++) {
        const double *q = p[b] + k * N + c;
        v[k] = 0.125 * (q[0] + q[1] + q[sj] + q[sj + 1] + q[si] + q[si + 1] + q[si + sj] + q[si + sj + 1]);
    }
    const double *f0 = ps[b] + c, *f1 = ps[b] + N + c, *f2 = ps[b] + 2 * N + c;
    v[NPR + 0] = 0.25 * (f0[0] + f0[1] + f0[sj] + f0[sj + 1]);
    v[NPR + 1] = 0.25 * (f1[0] + f1[1] + f1[si] + f1[si + 1]);
    v[NPR + 2] = 0.25 * (f2[0] + f2[sj] + f2[si] + f2[si + sj]);
}

/* slope_lim_2 (GPU_program2.cu:10-45, the MC limiter with factor 2) */
static double slope_lim(double y1, double y2, double y3)
{
    double Dqm = 2.0 * (y2 - y1), Dqp = 2.0 * (y3 - y2), Dqc = 0.5 * (y3 - y1), s = Dqm * Dqp;
    if (s <= 0.) return 0.;
    if (fabs(Dqm) < fabs(Dqp) && fabs(Dqm) < fabs(Dqc)) return Dqm;
    else if (fabs(Dqp) < fabs(Dqc)) return Dqp;
    else return Dqc;
}

/* Test switch (planted-mistake tests only): 1 makes the exchange also write
 * the normal face field of a block's own upper boundary face, which the
 * exchange rule forbids. */
int oracle_amr_plant_own_face = 0;
/* Test switch (planted-mistake tests only): 1 drops the corner-ghost rule of
 * stages 2 and 3 (corner ghosts then come from the leaf holding the corner). */
int oracle_amr_plant_corner = 0;

/* write the ghost values; the component `skip` (0..2, or -1) of the face field is left as it is */
/* Local-adaptive-time-stepping exchange (NULL members: the plain exchange):
 * only blocks with act[b] != 0 (stepping in the sub-iteration) exchange, with
 * sources read from the arrays passed (each block's arrays of its phase) and
 * ghosts written into those and into p2/ps2 (each block's other array pair). */
typedef struct {
    const int *act;
    double **p2, **ps2;
} las_x;

static void write_ghost(const oracle_amr *g, int b, long c, double **p, double **ps, int **pflag, const double *v,
                        int flag, int skip, const las_x *lx)
{
    const long N = geo_of(&g->blk[b]).ncell;
    for (int pass = 0; pass < (lx && lx->p2 ? 2 : 1); pass++) {
        double *pp = pass ? lx->p2[b] : p[b], *qq = pass ? lx->ps2[b] : ps[b];
        for (int k = 0; k < NPR; k++) pp[k * N + c] = v[k];
        for (int a = 0; a < 3; a++)
            if (a != skip) qq[a * N + c] = v[NPR + a];
    }
    if (pflag) pflag[b][c] = flag;
}

/* gdet of face position pos of column (i, j) of block b */
static double face_gdet(const oracle_amr *g, int b, int pos, int i, int j)
{
    const oracle_amr_block *B = &g->blk[b];
    const long ncol = (long)(B->bs[0] + 2 * NG) * (B->bs[1] + 2 * NG);
    return g->m[b].gdet[pos * ncol + (long)i * (B->bs[1] + 2 * NG) + j];
}

/* gdet-weighted mean of the four finer faces normal to direction dd of the
 * 2x2x2 cells of block c from local cell (i, j, z) (the flux mean of
 * packsendaverage1-3, GPU_program2.cu:176-185, :231-236, :281-285, same term
 * order), divided by the coarse face's gdet gc */
static double flux_mean(const oracle_amr *g, int c, int dd, int i, int j, int z, double **ps, double gc)
{
    const oracle_amr_block *B = &g->blk[c];
    const long N = geo_of(B).ncell;
    const double *f = ps[c] + dd * N;
    double m;
#define F(a, b2, e) f[cell_local(B, a, b2, e)]
    if (dd == 0) {
        const double g0 = face_gdet(g, c, FACE1, i, j), g1 = face_gdet(g, c, FACE1, i, j + 1);
        m = 0.25 * (F(i, j, z) * g0 + F(i, j, z + 1) * g0 + F(i, j + 1, z) * g1 + F(i, j + 1, z + 1) * g1);
    } else if (dd == 1) {
        const double g0 = face_gdet(g, c, FACE2, i, j), g1 = face_gdet(g, c, FACE2, i + 1, j);
        m = 0.25 * (F(i, j, z) * g0 + F(i, j, z + 1) * g0 + F(i + 1, j, z) * g1 + F(i + 1, j, z + 1) * g1);
    } else {
        m = 0.25 * (F(i, j, z) * face_gdet(g, c, FACE3, i, j) + F(i, j + 1, z) * face_gdet(g, c, FACE3, i, j + 1) +
                    F(i + 1, j, z) * face_gdet(g, c, FACE3, i + 1, j) +
                    F(i + 1, j + 1, z) * face_gdet(g, c, FACE3, i + 1, j + 1));
    }
#undef F
    return m / gc;
}

/* value at position Q of level lq's coarse neighbour of a prolongation; 0 if none */
static int neighbour_values(const oracle_amr *g, int lq, int cblk, int Q[3], int d, double **p, double **ps,
                            double *v)
{
    if (d == 2) Q[2] = wrap(Q[2], g->N[lq][2]);
    if (Q[d] < 0 || Q[d] >= g->N[lq][d]) return 0;
    if (inside(&g->blk[cblk], Q)) {
        cell_values(g, cblk, cell_at(&g->blk[cblk], Q), p, ps, v);
        return 1;
    }
    int ll, q = find_leaf(g, lq, Q, &ll);
    if (q < 0) abort();
    const oracle_amr_block *S = &g->blk[q];
    if (ll == lq) {
        cell_values(g, q, cell_at(S, Q), p, ps, v);
    } else if (ll == lq + 1) {
        int F[3] = {2 * Q[0], 2 * Q[1], 2 * Q[2]};
        avg8_values(g, q, cell_at(S, F), p, ps, v);
    } else {
        int sh = lq - ll, C[3] = {Q[0] >> sh, Q[1] >> sh, Q[2] >> sh};
        cell_values(g, q, cell_at(S, C), p, ps, v);
    }
    return 1;
}

/* Exchange of ghost cell (i, j, z) of block b in the stage along dd (0..2).
 * The dd-component of the face field at the block's own upper boundary face
 * (ghost index NG+bs along dd, interior across) is never written: the block's
 * CT update owns it (the reference's receive stores no normal component,
 * unpackreceive1, GPU_program2.cu:376-380). */
static void exchange_cell(const oracle_amr *g, int dd, int b, int i, int j, int z, double **p, double **ps,
                          int **pflag, const las_x *lx)
{
    const oracle_amr_block *B = &g->blk[b];
    const int l = B->level;
    const long dc = cell_local(B, i, j, z);
    const int loc[3] = {i, j, z};
    int own = loc[dd] == NG + B->bs[dd];
    for (int k = 0; k < 3; k++)
        if (k != dd && (loc[k] < NG || loc[k] >= NG + B->bs[k])) own = 0;
    const int skip = (own && !oracle_amr_plant_own_face) ? dd : -1;
    int Y[3] = {B->o[0] + i - NG, B->o[1] + j - NG, B->o[2] + z - NG};
    Y[2] = wrap(Y[2], g->N[l][2]);
    const int out1 = Y[0] < 0 || Y[0] >= g->N[l][0], out2 = Y[1] < 0 || Y[1] >= g->N[l][1];
    double v[NV];
    if (out2) {
        int Yc[3] = {Y[0] < 0 ? 0 : (Y[0] >= g->N[l][0] ? g->N[l][0] - 1 : Y[0]), Y[1] < 0 ? 0 : g->N[l][1] - 1, Y[2]};
        int lc, c = find_leaf(g, l, Yc, &lc);
        if (c < 0 || lc != l) abort();
        if (lx && lx->act && !lx->act[c]) return;
        const long sc = cell_at(&g->blk[c], Y);
        cell_values(g, c, sc, p, ps, v);
        write_ghost(g, b, dc, p, ps, pflag, v, pflag ? pflag[c][sc] : 0, skip, lx);
        return;
    }
    /* a corner ghost (beyond the block along an earlier direction): the
     * same-level leaf across the face that holds the position clamped into the
     * block's range has it in its storage (its ghost of the earlier stage) */
    {
        int corner = 0, Yk[3] = {Y[0], Y[1], Y[2]};
        for (int e = 0; e < dd; e++)
            if (loc[e] < NG || loc[e] >= NG + B->bs[e]) {
                corner = 1;
                Yk[e] = Y[e] < B->o[e] ? B->o[e] : (Y[e] >= B->o[e] + B->bs[e] ? B->o[e] + B->bs[e] - 1 : Y[e]);
            }
        if (corner && !oracle_amr_plant_corner) {
            int lk, k = find_leaf(g, l, Yk, &lk);
            if (k >= 0 && lk == l) {
                if (lx && lx->act && !lx->act[k]) return;
                const long sc = cell_at(&g->blk[k], Y);
                cell_values(g, k, sc, p, ps, v);
                write_ghost(g, b, dc, p, ps, pflag, v, pflag ? pflag[k][sc] : 0, skip, lx);
                return;
            }
        }
    }
    if (out1) return;
    int lc, c = find_leaf(g, l, Y, &lc);
    if (c < 0) abort();
    if (lx && lx->act && !lx->act[c]) return;
    const oracle_amr_block *C = &g->blk[c];
    if (lc == l) {
        const long sc = cell_at(C, Y);
        cell_values(g, c, sc, p, ps, v);
        write_ghost(g, b, dc, p, ps, pflag, v, pflag ? pflag[c][sc] : 0, skip, lx);
    } else if (lc == l + 1) {
        int F[3] = {2 * Y[0], 2 * Y[1], 2 * Y[2]};
        const long sc = cell_at(C, F);
        avg8_values(g, c, sc, p, ps, v);
        const int pos = dd == 0 ? FACE1 : dd == 1 ? FACE2 : FACE3;
        v[NPR + dd] = flux_mean(g, c, dd, F[0] - C->o[0] + NG, F[1] - C->o[1] + NG, F[2] - C->o[2] + NG, ps,
                                face_gdet(g, b, pos, i, j));
        write_ghost(g, b, dc, p, ps, pflag, v, pflag ? pflag[c][sc] : 0, skip, lx);
    } else if (lc == l - 1) {
        int P[3] = {Y[0] >> 1, Y[1] >> 1, Y[2] >> 1};
        const long pc = cell_at(C, P);
        double c0[NV], dq[3][NV], vm[NV], vp[NV], sgn[3];
        cell_values(g, c, pc, p, ps, c0);
        for (int d = 0; d < 3; d++) {
            int Qm[3] = {P[0], P[1], P[2]}, Qp[3] = {P[0], P[1], P[2]};
            Qm[d] -= 1;
            Qp[d] += 1;
            const int hm = neighbour_values(g, l - 1, c, Qm, d, p, ps, vm);
            const int hp = neighbour_values(g, l - 1, c, Qp, d, p, ps, vp);
            for (int k = 0; k < NV; k++) dq[d][k] = (hm && hp) ? slope_lim(vm[k], c0[k], vp[k]) : 0.;
            sgn[d] = (Y[d] & 1) ? 1.0 : -1.0;
        }
        for (int k = 0; k < NPR; k++)
            v[k] = c0[k] + 0.25 * sgn[0] * dq[0][k] + 0.25 * sgn[1] * dq[1][k] + 0.25 * sgn[2] * dq[2][k];
        for (int a = 0; a < 3; a++) {
            const int kk = (sgn[a] > 0.) ? B1 + a : NPR + a;
            double w = c0[kk];
            for (int d = 0; d < 3; d++)
                if (d != a) w = w + 0.25 * sgn[d] * dq[d][kk];
            v[NPR + a] = w;
        }
        write_ghost(g, b, dc, p, ps, pflag, v, pflag ? pflag[c][pc] : 0, skip, lx);
    } else {
        abort();
    }
}

static int radial_face(const oracle_amr *g, const oracle_amr_block *B, int side)
{
    return side ? (B->o[0] + B->bs[0] == g->N[B->level][0]) : (B->o[0] == 0);
}

static int polar_face(const oracle_amr *g, const oracle_amr_block *B, int side)
{
    return side ? (B->o[1] + B->bs[1] == g->N[B->level][1]) : (B->o[1] == 0);
}

static void exchange_stage(const oracle_amr *g, int stage, double **p, double **ps, int **pflag, const las_x *lx)
{
    const int dd = stage - 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) {
        if (lx && lx->act && !lx->act[b]) continue;
        const oracle_amr_block *B = &g->blk[b];
        const bgeo G = geo_of(B);
        for (int side = 0; side < 2; side++) {
            if ((dd == 0 && radial_face(g, B, side)) || (dd == 1 && polar_face(g, B, side))) continue;
            int lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                lo[k] = k < dd ? 0 : NG;
                hi[k] = k < dd ? G.D[k] : NG + B->bs[k];
            }
            lo[dd] = side ? NG + B->bs[dd] : 0;
            hi[dd] = side ? G.D[dd] : NG;
            for (int i = lo[0]; i < hi[0]; i++)
                for (int j = lo[1]; j < hi[1]; j++)
                    for (int z = lo[2]; z < hi[2]; z++) exchange_cell(g, dd, b, i, j, z, p, ps, pflag, lx);
        }
    }
}

void oracle_amr_exchange(const oracle_amr *g, int stage, double **p, double **ps, int **pflag)
{
    exchange_stage(g, stage, p, ps, pflag, NULL);
}

/* metric of block b at position pos of column (i, j) */
static void geom(const oracle_amr *g, int b, int pos, int i, int j, const double **gcov, const double **gcon)
{
    const oracle_amr_block *B = &g->blk[b];
    const long ncol = (long)(B->bs[0] + 2 * NG) * (B->bs[1] + 2 * NG), col = (long)i * (B->bs[1] + 2 * NG) + j;
    *gcov = g->m[b].gcov + ((long)pos * ncol + col) * 10;
    *gcon = g->m[b].gcon + ((long)pos * ncol + col) * 10;
}

/* radial outflow of block b (GPU_bounds.h:3-175): every row, or (final pass)
 * every row but the polar ghost rows */
static void outflow(const oracle_amr *g, int b, double **p, double **ps, int **pflag, int skip_polar_rows)
{
    const oracle_amr_block *B = &g->blk[b];
    const bgeo G = geo_of(B);
    const long N = G.ncell;
    for (int side = 0; side < 2; side++) {
        if (!radial_face(g, B, side)) continue;
        const int isrc = side ? NG + B->bs[0] - 1 : NG;
        for (int j = 0; j < G.D[1]; j++) {
            if (skip_polar_rows && ((polar_face(g, B, 0) && j < NG) || (polar_face(g, B, 1) && j >= NG + B->bs[1])))
                continue;
            for (int z = 0; z < G.D[2]; z++)
                for (int gg = 0; gg < NG; gg++) {
                    const int i = side ? NG + B->bs[0] + gg : gg;
                    const long c = cell_local(B, i, j, z), cs = cell_local(B, isrc, j, z);
                    double pr[NPR];
                    const double *gcov, *gcon;
                    for (int k = 0; k < NPR; k++) pr[k] = p[b][k * N + cs];
                    geom(g, b, CENT, i, j, &gcov, &gcon);
                    oracle_inflow_check(pr, gcov, gcon, side ? 1 : 0);
                    for (int k = 0; k < NPR; k++) p[b][k * N + c] = pr[k];
                    ps[b][1 * N + c] = ps[b][1 * N + cs];
                    ps[b][2 * N + c] = ps[b][2 * N + cs];
                    if (pflag) pflag[b][c] = pflag[b][cs];
                }
        }
    }
}

/* polar partner: the block of the same level and shape half a turn away */
static int partner(const oracle_amr *g, int b)
{
    const oracle_amr_block *B = &g->blk[b];
    const int o3 = (B->o[2] + g->N[B->level][2] / 2) % g->N[B->level][2];
    for (int q = 0; q < g->nblk; q++) {
        const oracle_amr_block *Q = &g->blk[q];
        if (Q->level == B->level && Q->o[0] == B->o[0] && Q->o[1] == B->o[1] && Q->o[2] == o3 && Q->bs[0] == B->bs[0] &&
            Q->bs[1] == B->bs[1] && Q->bs[2] == B->bs[2])
            return q;
    }
    return -1;
}

/* polar mirror (GPU_bounds.h:510-606) over every i and z of block b */
static void polar(const oracle_amr *g, int b, double **p, double **ps, int **pflag)
{
    const oracle_amr_block *B = &g->blk[b];
    const bgeo G = geo_of(B);
    const long N = G.ncell;
    for (int side = 0; side < 2; side++) {
        if (!polar_face(g, B, side)) continue;
        const int q = partner(g, b);
        if (q < 0) abort();
        for (int i = 0; i < G.D[0]; i++)
            for (int z = 0; z < G.D[2]; z++) {
                int zs = z;
                if (q == b) zs = (((z - NG + B->bs[2] / 2) % B->bs[2]) + B->bs[2]) % B->bs[2] + NG;
                for (int gg = 0; gg < NG; gg++) {
                    const int j = side ? NG + B->bs[1] + gg : NG - 1 - gg;
                    const int jsrc = side ? NG + B->bs[1] - 1 - gg : NG + gg;
                    const long c = cell_local(B, i, j, z), cs = cell_local(B, i, jsrc, zs);
                    for (int k = 0; k < NPR; k++) p[b][k * N + c] = p[q][k * N + cs];
                    p[b][(long)U2 * N + c] *= -1.0;
                    p[b][(long)U3 * N + c] *= -1.0;
                    p[b][(long)B2 * N + c] *= -1.0;
                    p[b][(long)B3 * N + c] *= -1.0;
                    ps[b][0 * N + c] = ps[q][0 * N + cs];
                    ps[b][2 * N + c] = -ps[q][2 * N + cs];
                    if (pflag) pflag[b][c] = pflag[q][cs];
                }
            }
    }
}

void oracle_amr_bound(const oracle_amr *g, double **p, double **ps, int **pflag)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) outflow(g, b, p, ps, pflag, 0);
    oracle_amr_exchange(g, 1, p, ps, pflag);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) polar(g, b, p, ps, pflag);
    oracle_amr_exchange(g, 2, p, ps, pflag);
    oracle_amr_exchange(g, 3, p, ps, pflag);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) outflow(g, b, p, ps, pflag, 1);
}

static oracle_params block_params(const oracle_amr *g, int b)
{
    const oracle_amr_block *B = &g->blk[b];
    oracle_params par;
    memset(&par, 0, sizeof par);
    par.n1 = par.bs1 = B->bs[0];
    par.n2 = par.bs2 = B->bs[1];
    par.n3 = par.bs3 = B->bs[2];
    par.bc[0] = (radial_face(g, B, 0) || radial_face(g, B, 1)) ? ORC_BC_OUTFLOW : ORC_BC_NONE;
    const int lo = polar_face(g, B, 0), hi = polar_face(g, B, 1);
    par.bc[1] = (lo && hi) ? ORC_BC_POLE : lo ? ORC_BC_POLE_LO : hi ? ORC_BC_POLE_HI : ORC_BC_NONE;
    par.bc[2] = ORC_BC_NONE;
    par.gam = g->gam;
    par.cour = g->cour;
    for (int k = 0; k < 3; k++) par.dx[k] = g->dx0[k] * ldexp(1.0, -B->level);
    par.rho_floor = g->rho_floor;
    par.u_floor = g->u_floor;
    par.riemann = g->riemann;
    return par;
}

/* polar E1 average over each ring of blocks (same side, level, radial range and
 * shape, ordered in phi), as the single-grid oracle averages over its phi-blocks */
static void ring_average_las(const oracle_amr *g, double **E, double dt, const int *tl, int las_s);
static void ring_average(const oracle_amr *g, double **E, double dt) { ring_average_las(g, E, dt, NULL, -1); }

/* tl != NULL: the rings stepping in sub-iteration las_s, each weighted by tl dt */
static void ring_average_las(const oracle_amr *g, double **E, double dt0, const int *tl, int las_s)
{
    int *done = calloc((size_t)g->nblk * 2, sizeof(int));
    int *ring = malloc(sizeof(int) * (size_t)g->nblk);
    for (int b0 = 0; b0 < g->nblk; b0++)
        for (int side = 0; side < 2; side++) {
            const oracle_amr_block *A = &g->blk[b0];
            if (!polar_face(g, A, side) || done[b0 * 2 + side]) continue;
            int len = 0;
            for (int q = 0; q < g->nblk; q++) {
                const oracle_amr_block *Q = &g->blk[q];
                if (polar_face(g, Q, side) && Q->level == A->level && Q->o[0] == A->o[0] && Q->bs[0] == A->bs[0] &&
                    Q->bs[1] == A->bs[1] && Q->bs[2] == A->bs[2])
                    ring[len++] = q;
            }
            /* phi order */
            for (int x = 1; x < len; x++)
                for (int y = x; y > 0 && g->blk[ring[y]].o[2] < g->blk[ring[y - 1]].o[2]; y--) {
                    int t = ring[y];
                    ring[y] = ring[y - 1];
                    ring[y - 1] = t;
                }
            for (int x = 0; x < len; x++) done[ring[x] * 2 + side] = 1;
            if (tl && oracle_las_phase(tl[ring[0]], las_s) == 0) continue;
            const double dt = tl ? dt0 * (double)tl[ring[0]] : dt0;
            const int j = side ? NG + A->bs[1] : NG;
            for (int i = NG; i <= NG + A->bs[0]; i++) {
                double avg = 0.;
                for (int b = 0; b < len; b++) {
                    const oracle_amr_block *B = &g->blk[ring[b]];
                    double s = 0.;
                    for (int zz = 0; zz < B->bs[2]; zz++) s += dt * E[ring[b]][cell_local(B, i, j, NG + zz)];
                    s /= (double)(B->bs[2]);
                    if (b == 0) avg = s / ((double)len);
                    else avg += s / ((double)len);
                }
                for (int b = 0; b < len; b++) {
                    const oracle_amr_block *B = &g->blk[ring[b]];
                    const long N = geo_of(B).ncell;
                    for (int z = NG; z <= NG + B->bs[2]; z++) {
                        E[ring[b]][cell_local(B, i, j, z)] = avg / dt;
                        E[ring[b]][2 * N + cell_local(B, i, j, z)] = 0.;
                    }
                }
            }
        }
    free(done);
    free(ring);
}

/* Flux correction at level jumps (flux_send / flux_rec with calc_corr = 1,
 * boundcomF.c:5-1315; packsendfluxaverage1-3, GPU_program2.cu:1618-1771;
 * unpackreceive1flux-3flux, :1258-1617): a face of a block whose neighbour
 * across it is one level finer takes the flux ((0.25 dt) (f00 + f01 + f10 +
 * f11)) / dt of the four finer faces that tile it, f_ab at transverse offsets
 * a (the slower direction) and b; the finer faces keep their own fluxes
 * (AVG_FLUXES 0).  The leaf across each face cell is found by searching the
 * block list. */
void oracle_amr_reflux(const oracle_amr *g, double **F1, double **F2, double **F3, double dt)
{
    double **F[3] = {F1, F2, F3};
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) {
        const oracle_amr_block *C = &g->blk[b];
        const long NC = geo_of(C).ncell;
        for (int d = 0; d < 3; d++) {
            const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2; /* transverse directions, slower first */
            for (int side = 0; side < 2; side++) {
                int Y[3];
                Y[d] = side ? C->o[d] + C->bs[d] : C->o[d] - 1;
                if (d < 2 && (Y[d] < 0 || Y[d] >= g->N[C->level][d])) continue; /* radial edge or axis */
                if (d == 2) Y[2] = wrap(Y[2], g->N[C->level][2]);
                for (int ya = C->o[ta]; ya < C->o[ta] + C->bs[ta]; ya++)
                    for (int yb = C->o[tb]; yb < C->o[tb] + C->bs[tb]; yb++) {
                        Y[ta] = ya;
                        Y[tb] = yb;
                        int lc;
                        const int f = find_leaf(g, C->level, Y, &lc);
                        if (f < 0 || lc != C->level + 1) continue;
                        const oracle_amr_block *Fb = &g->blk[f];
                        const long NF = geo_of(Fb).ncell;
                        /* the finer cell beside the face (lowest transverse corner) and its face */
                        int Z[3], L[3], K[3];
                        Z[d] = side ? 2 * Y[d] : 2 * Y[d] + 1;
                        Z[ta] = 2 * ya;
                        Z[tb] = 2 * yb;
                        for (int k = 0; k < 3; k++) L[k] = Z[k] - Fb->o[k] + NG;
                        L[d] += side ? 0 : 1;
                        for (int k = 0; k < 3; k++) K[k] = (k == d) ? (side ? NG + C->bs[d] : NG) : Y[k] - C->o[k] + NG;
                        int Lb[3] = {L[0], L[1], L[2]}, La[3] = {L[0], L[1], L[2]}, Lab[3] = {L[0], L[1], L[2]};
                        Lb[tb] += 1;
                        La[ta] += 1;
                        Lab[ta] += 1;
                        Lab[tb] += 1;
                        const long f00 = cell_local(Fb, L[0], L[1], L[2]), f01 = cell_local(Fb, Lb[0], Lb[1], Lb[2]);
                        const long f10 = cell_local(Fb, La[0], La[1], La[2]), f11 = cell_local(Fb, Lab[0], Lab[1], Lab[2]);
                        const long cc = cell_local(C, K[0], K[1], K[2]);
                        for (int k = 0; k < NPR; k++) {
                            const double *q = F[d][f] + k * NF;
                            const double sum = q[f00] + q[f01] + q[f10] + q[f11];
                            F[d][b][k * NC + cc] = ((0.25 * dt) * sum) / dt;
                        }
                    }
            }
        }
    }
}

/* EMF correction at level jumps (E_send1-3 / E_rec1-3 with calc_corr = 1 and
 * the edge exchanges of boundcomE.c; packsendEaverage1-3 and unpackreceive1-3E,
 * GPU_program2.cu:1875-2484): every edge EMF on a block's boundary that a leaf
 * one level finer touches (one of the four cells around the edge, searched in
 * the fixed order (-,-), (-,+), (+,-), (+,+) of the two directions across it)
 * takes ((0.5 dt) (e0 + e1)) / dt of that leaf's two edges along it. */
void oracle_amr_emf_correct(const oracle_amr *g, double **E, double dt)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) {
        const oracle_amr_block *C = &g->blk[b];
        const int l = C->level;
        if (l + 1 >= g->nlev) continue;
        const long NC = geo_of(C).ncell;
        for (int t = 0; t < 3; t++) {
            const int p = (t + 1) % 3, q = (t + 2) % 3;
            int K[3];
            for (K[t] = C->o[t]; K[t] < C->o[t] + C->bs[t]; K[t]++)
                for (K[p] = C->o[p]; K[p] <= C->o[p] + C->bs[p]; K[p]++)
                    for (K[q] = C->o[q]; K[q] <= C->o[q] + C->bs[q]; K[q]++) {
                        if (K[p] != C->o[p] && K[p] != C->o[p] + C->bs[p] && K[q] != C->o[q] &&
                            K[q] != C->o[q] + C->bs[q])
                            continue; /* not on the block's boundary */
                        int f = -1, Z[3];
                        for (int dp = -1; dp <= 0 && f < 0; dp++)
                            for (int dq = -1; dq <= 0 && f < 0; dq++) {
                                int Y[3] = {K[0], K[1], K[2]}, ok = 1, lc;
                                Y[p] += dp;
                                Y[q] += dq;
                                for (int x = 0; x < 2; x++)
                                    if (Y[x] < 0 || Y[x] >= g->N[l][x]) ok = 0; /* beyond the radial edge or the axis */
                                Y[2] = wrap(Y[2], g->N[l][2]);
                                if (!ok) continue;
                                const int c = find_leaf(g, l, Y, &lc);
                                if (c >= 0 && lc == l + 1) f = c;
                            }
                        if (f < 0) continue;
                        const oracle_amr_block *Fb = &g->blk[f];
                        for (int x = 0; x < 3; x++) Z[x] = 2 * K[x];
                        /* the finer leaf across the periodic x3 seam */
                        if (Z[2] < Fb->o[2]) Z[2] += g->N[l + 1][2];
                        if (Z[2] > Fb->o[2] + Fb->bs[2]) Z[2] -= g->N[l + 1][2];
                        int Z1[3] = {Z[0], Z[1], Z[2]};
                        Z1[t] += 1;
                        const double *e = E[f] + (long)t * geo_of(Fb).ncell;
                        const double sum = e[cell_at(Fb, Z)] + e[cell_at(Fb, Z1)];
                        E[b][(long)t * NC + cell_at(C, K)] = ((0.5 * dt) * sum) / dt;
                    }
        }
    }
}

void oracle_amr_fluxes(const oracle_amr *g, double **p, double **ps, double dt, double **F1, double **F2, double **F3)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < g->nblk; b++) {
        oracle_params par = block_params(g, b);
        oracle_flux(&par, 1, p[b], ps[b], &g->m[b], F1[b], NULL);
        oracle_flux(&par, 2, p[b], ps[b], &g->m[b], F2[b], NULL);
        oracle_flux(&par, 3, p[b], ps[b], &g->m[b], F3[b], NULL);
    }
    if (g->reflux) oracle_amr_reflux(g, F1, F2, F3, dt);
}

void oracle_amr_ct(const oracle_amr *g, double **p, double **ps, double dt, double **psf)
{
    const int nb = g->nblk;
    double **F1 = calloc(nb, sizeof(double *)), **F2 = calloc(nb, sizeof(double *)), **F3 = calloc(nb, sizeof(double *));
    double **E = calloc(nb, sizeof(double *)), **Ec = calloc(nb, sizeof(double *));
    int ok = F1 && F2 && F3 && E && Ec;
    for (int b = 0; ok && b < nb; b++) {
        const long N = geo_of(&g->blk[b]).ncell;
        F1[b] = calloc(N * NPR, sizeof(double));
        F2[b] = calloc(N * NPR, sizeof(double));
        F3[b] = calloc(N * NPR, sizeof(double));
        E[b] = calloc(N * 3, sizeof(double));
        Ec[b] = calloc(N * 3, sizeof(double));
        ok = F1[b] && F2[b] && F3[b] && E[b] && Ec[b];
    }
    if (ok) {
        oracle_amr_fluxes(g, p, ps, dt, F1, F2, F3);
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            oracle_emf(&par, p[b], F1[b], F2[b], F3[b], &g->m[b], dt, E[b], Ec[b]);
        }
        ring_average(g, E, dt);
        if (g->reflux) oracle_amr_emf_correct(g, E, dt);
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            memcpy(psf[b], ps[b], sizeof(double) * 3 * geo_of(&g->blk[b]).ncell);
            oracle_ct_update(&par, dt, ps[b], psf[b], E[b], &g->m[b]);
        }
    }
    for (int b = 0; b < nb; b++) {
        if (F1) free(F1[b]);
        if (F2) free(F2[b]);
        if (F3) free(F3[b]);
        if (E) free(E[b]);
        if (Ec) free(Ec[b]);
    }
    free(F1); free(F2); free(F3); free(E); free(Ec);
}

double oracle_amr_step(const oracle_amr *g, double **p, double **ps, double dt, int **pflag, long long *failcount)
{
    const int nb = g->nblk;
    double **ph = calloc(nb, sizeof(double *)), **psh = calloc(nb, sizeof(double *));
    double **F1 = calloc(nb, sizeof(double *)), **F2 = calloc(nb, sizeof(double *)), **F3 = calloc(nb, sizeof(double *));
    double **U0 = calloc(nb, sizeof(double *)), **E = calloc(nb, sizeof(double *)), **Ec = calloc(nb, sizeof(double *));
    double **d1 = calloc(nb, sizeof(double *)), **d2 = calloc(nb, sizeof(double *)), **d3 = calloc(nb, sizeof(double *));
    double ndt = -1.;
    int ok = ph && psh && F1 && F2 && F3 && U0 && E && Ec && d1 && d2 && d3;
    for (int b = 0; ok && b < nb; b++) {
        const long N = geo_of(&g->blk[b]).ncell;
        ph[b] = malloc(sizeof(double) * N * NPR);
        psh[b] = malloc(sizeof(double) * N * 3);
        F1[b] = calloc(N * NPR, sizeof(double));
        F2[b] = calloc(N * NPR, sizeof(double));
        F3[b] = calloc(N * NPR, sizeof(double));
        U0[b] = calloc(N * NPR, sizeof(double));
        E[b] = calloc(N * 3, sizeof(double));
        Ec[b] = calloc(N * 3, sizeof(double));
        d1[b] = malloc(sizeof(double) * N);
        d2[b] = malloc(sizeof(double) * N);
        d3[b] = malloc(sizeof(double) * N);
        ok = ph[b] && psh[b] && F1[b] && F2[b] && F3[b] && U0[b] && E[b] && Ec[b] && d1[b] && d2[b] && d3[b];
        if (ok) {
            memcpy(ph[b], p[b], sizeof(double) * N * NPR);
            memcpy(psh[b], ps[b], sizeof(double) * N * 3);
        }
    }
    if (ok) {
        /* predictor: p -> ph with Dt = dt/2 */
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            oracle_flux(&par, 1, p[b], ps[b], &g->m[b], F1[b], NULL);
            oracle_flux(&par, 2, p[b], ps[b], &g->m[b], F2[b], NULL);
            oracle_flux(&par, 3, p[b], ps[b], &g->m[b], F3[b], NULL);
            oracle_emf(&par, p[b], F1[b], F2[b], F3[b], &g->m[b], dt, E[b], Ec[b]);
        }
        ring_average(g, E, dt);
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            oracle_ct_update(&par, 0.5 * dt, ps[b], psh[b], E[b], &g->m[b]);
            oracle_update(&par, 0, 0.5 * dt, p[b], p[b], ph[b], U0[b], psh[b], F1[b], F2[b], F3[b], &g->m[b],
                          pflag[b], failcount);
        }
        oracle_amr_bound(g, ph, psh, pflag);
        /* corrector: U0 -> p with Dt = dt, fluxes from ph, corrected at level
         * jumps before the EMFs use them (step_ch.c:1142-1196) */
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            oracle_flux(&par, 1, ph[b], psh[b], &g->m[b], F1[b], d1[b]);
            oracle_flux(&par, 2, ph[b], psh[b], &g->m[b], F2[b], d2[b]);
            oracle_flux(&par, 3, ph[b], psh[b], &g->m[b], F3[b], d3[b]);
        }
        if (g->reflux) oracle_amr_reflux(g, F1, F2, F3, dt);
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            oracle_emf(&par, ph[b], F1[b], F2[b], F3[b], &g->m[b], dt, E[b], Ec[b]);
        }
        ring_average(g, E, dt);
        if (g->reflux) oracle_amr_emf_correct(g, E, dt);
        ndt = 1e9;
        for (int b = 0; b < nb; b++) {
            oracle_params par = block_params(g, b);
            oracle_ct_update(&par, dt, ps[b], ps[b], E[b], &g->m[b]);
            oracle_update(&par, 1, dt, p[b], ph[b], p[b], U0[b], ps[b], F1[b], F2[b], F3[b], &g->m[b],
                          pflag[b], failcount);
            const double bd = oracle_block_dt(&par, d1[b], d2[b], d3[b]);
            if (bd < ndt) ndt = bd;
        }
        oracle_amr_bound(g, p, ps, pflag);
        if (ndt > 1.3 * dt) ndt = 1.3 * dt;
    }
    for (int b = 0; b < nb; b++) {
        if (ph) free(ph[b]);
        if (psh) free(psh[b]);
        if (F1) free(F1[b]);
        if (F2) free(F2[b]);
        if (F3) free(F3[b]);
        if (U0) free(U0[b]);
        if (E) free(E[b]);
        if (Ec) free(Ec[b]);
        if (d1) free(d1[b]);
        if (d2) free(d2[b]);
        if (d3) free(d3[b]);
    }
    free(ph); free(psh); free(F1); free(F2); free(F3); free(U0); free(E); free(Ec); free(d1); free(d2); free(d3);
    return ndt;
}

/* ---- local adaptive time stepping ----
 * One cycle of 2 maxtl sub-iterations (GPU_step_ch, step_ch.c:916-979, and
 * advance_GPU, :997-1336, with PRESTEP = PRESTEP2 = 0 and AMR_SWITCHTIMELEVEL
 * = AMR_MAXTIMELEVEL = maxtl).  In sub-iteration s a block of time level tl
 * (oracle_las_phase) takes its predictor (p -> ph, Dt = tl dt / 2) or its
 * corrector (ph -> p, Dt = tl dt) or rests.  Within a sub-iteration: fluxes of
 * the stepping blocks (their phase's arrays); flux corrections; EMFs; the polar
 * ring average of the stepping rings (weight tl dt); EMF corrections; CT and
 * conserved updates; the ghost fill of the stepping blocks (physical faces on
 * their updated arrays; exchanges only between stepping blocks, each source
 * read from its phase's arrays, each ghost written into both array pairs of
 * its block, as unpackreceive1 writes p and ph, GPU_program2.cu:345-388).
 *
 * Corrections (boundcomF.c / boundcomE.c with calc_corr = 1; the factor dt tl
 * and first-step rule of GPU_boundcomF.cu:7-8 and :109-110; the receive
 * acc / factor of unpackreceive1flux, GPU_program2.cu:1270-1277): a receiver
 * face or edge takes, at its corrector, acc / (dt tl_r), where acc sums
 * c (values) over the sender's correctors since the receiver's last one, c =
 * w dt tl_s, restarted when s % 2tl_r == 2tl_s - 1.  Senders, found by
 * searching the block list:
 *   faces: the one level finer leaf across (w = 1/4, its four faces, as
 *          oracle_amr_reflux), or a same-level leaf of smaller time level (w = 1);
 *   edges: first the same-level rule: an edge on the block boundary that no
 *          finer leaf touches takes the EMF of the same-level leaf around it
 *          with the smallest time level (first in the quadrant order on ties)
 *          when that is smaller than the block's (w = 1); then the level-jump
 *          rule of oracle_amr_emf_correct (w = 1/2).
 * The next dt comes from the last sub-iteration, where every block takes its
 * corrector: min over blocks of the harmonic Courant dt, clamped to 1.3 dt. */

/* Test switch (planted-mistake tests only): 1 leaves out the flux
 * corrections of a cycle, 2 the same-level EMF corrections. */
int oracle_las_plant = 0;

int oracle_las_phase(int tl, int s)
{
    if (tl < 1) tl = 1;
    if (s % (2 * tl) == 2 * tl - 1) return 2;
    if (s % tl == tl - 1) return 1;
    return 0;
}

/* same-level or finer sender of face cell (K: the face, Y: the cell across) of
 * block b along d; *mean = 1 for a finer leaf; -1: none */
static int face_sender(const oracle_amr *g, const int *tl, int b, int d, const int Y[3], int *mean)
{
    const oracle_amr_block *B = &g->blk[b];
    int lc;
    const int m = find_leaf(g, B->level, Y, &lc);
    *mean = 0;
    if (m < 0) return -1;
    if (lc == B->level + 1) {
        if (tl[m] > tl[b]) abort(); /* finer block slower than its coarser neighbour */
        *mean = 1;
        return m;
    }
    if (lc == B->level && m != b && tl[m] < tl[b]) return m;
    return -1;
}

static void las_flux_corr(const oracle_amr *g, const int *tl, int s, double dt, double **F[3], double **acc[3])
{
    for (int stage = 0; stage < 2; stage++)
#pragma omp parallel for schedule(dynamic, 1)
        for (int b = 0; b < g->nblk; b++) {
            const oracle_amr_block *C = &g->blk[b];
            const long NC = geo_of(C).ncell;
            const int tr = tl[b];
            if (stage == 1 && oracle_las_phase(tr, s) != 2) continue;
            for (int d = 0; d < 3; d++) {
                const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
                for (int side = 0; side < 2; side++) {
                    int Y[3];
                    Y[d] = side ? C->o[d] + C->bs[d] : C->o[d] - 1;
                    if (d < 2 && (Y[d] < 0 || Y[d] >= g->N[C->level][d])) continue;
                    if (d == 2) Y[2] = wrap(Y[2], g->N[C->level][2]);
                    for (int ya = C->o[ta]; ya < C->o[ta] + C->bs[ta]; ya++)
                        for (int yb = C->o[tb]; yb < C->o[tb] + C->bs[tb]; yb++) {
                            Y[ta] = ya;
                            Y[tb] = yb;
                            int mean;
                            const int m = face_sender(g, tl, b, d, Y, &mean);
                            if (m < 0) continue;
                            int K[3];
                            for (int k = 0; k < 3; k++)
                                K[k] = (k == d) ? (side ? NG + C->bs[d] : NG) : Y[k] - C->o[k] + NG;
                            const long cc = cell_local(C, K[0], K[1], K[2]);
                            if (stage == 1) {
                                for (int k = 0; k < NPR; k++)
                                    F[d][b][k * NC + cc] = acc[d][b][k * NC + cc] / (dt * (double)tr);
                                continue;
                            }
                            const int ts = tl[m];
                            if (oracle_las_phase(ts, s) != 2) continue;
                            const int first = s % (2 * tr) == 2 * ts - 1;
                            const oracle_amr_block *S = &g->blk[m];
                            const long NS = geo_of(S).ncell;
                            long f00, f01 = 0, f10 = 0, f11 = 0;
                            if (mean) {
                                int Z[3], L[3];
                                Z[d] = side ? 2 * Y[d] : 2 * Y[d] + 1;
                                Z[ta] = 2 * ya;
                                Z[tb] = 2 * yb;
                                for (int k = 0; k < 3; k++) L[k] = Z[k] - S->o[k] + NG;
                                L[d] += side ? 0 : 1;
                                int Lb[3] = {L[0], L[1], L[2]}, La[3] = {L[0], L[1], L[2]}, Lab[3] = {L[0], L[1], L[2]};
                                Lb[tb] += 1;
                                La[ta] += 1;
                                Lab[ta] += 1;
                                Lab[tb] += 1;
                                f00 = cell_local(S, L[0], L[1], L[2]);
                                f01 = cell_local(S, Lb[0], Lb[1], Lb[2]);
                                f10 = cell_local(S, La[0], La[1], La[2]);
                                f11 = cell_local(S, Lab[0], Lab[1], Lab[2]);
                            } else {
                                /* the same face in the neighbour: its lower face if it lies above */
                                int L[3];
                                for (int k = 0; k < 3; k++) L[k] = Y[k] - S->o[k] + NG;
                                L[d] = side ? NG : NG + S->bs[d];
                                f00 = cell_local(S, L[0], L[1], L[2]);
                            }
                            const double c = (mean ? 0.25 : 1.0) * (dt * (double)ts);
                            for (int k = 0; k < NPR; k++) {
                                const double *q = F[d][m] + k * NS;
                                const double sum = mean ? q[f00] + q[f01] + q[f10] + q[f11] : q[f00];
                                double *a = &acc[d][b][k * NC + cc];
                                *a = first ? c * sum : *a + c * sum;
                            }
                        }
                }
            }
        }
}

/* edge sender of boundary edge K (component t) of block b: rule 0 same-level,
 * rule 1 level jump; -1: none.  Z receives the sender's edge (level of the sender). */
static int edge_sender(const oracle_amr *g, const int *tl, int b, int t, const int K[3], int rule, int Z[3])
{
    const oracle_amr_block *C = &g->blk[b];
    const int l = C->level, p = (t + 1) % 3, q = (t + 2) % 3;
    int best = -1, finer = -1;
    for (int dp = -1; dp <= 0; dp++)
        for (int dq = -1; dq <= 0; dq++) {
            int Y[3] = {K[0], K[1], K[2]}, ok = 1, lc;
            Y[p] += dp;
            Y[q] += dq;
            for (int x = 0; x < 2; x++)
                if (Y[x] < 0 || Y[x] >= g->N[l][x]) ok = 0;
            Y[2] = wrap(Y[2], g->N[l][2]);
            if (!ok) continue;
            const int c = find_leaf(g, l, Y, &lc);
            if (c < 0) continue;
            if (lc == l + 1 && finer < 0) finer = c;
            if (lc == l && (best < 0 || tl[c] < tl[best])) best = c;
        }
    if (rule == 1) {
        if (finer < 0) return -1;
        if (tl[finer] > tl[b]) abort();
        const oracle_amr_block *Fb = &g->blk[finer];
        for (int x = 0; x < 3; x++) Z[x] = 2 * K[x];
        if (Z[2] < Fb->o[2]) Z[2] += g->N[l + 1][2];
        if (Z[2] > Fb->o[2] + Fb->bs[2]) Z[2] -= g->N[l + 1][2];
        return finer;
    }
    if (finer >= 0 || best < 0 || tl[best] >= tl[b]) return -1;
    const oracle_amr_block *S = &g->blk[best];
    for (int x = 0; x < 3; x++) Z[x] = K[x];
    if (Z[2] < S->o[2]) Z[2] += g->N[l][2];
    if (Z[2] > S->o[2] + S->bs[2]) Z[2] -= g->N[l][2];
    return best;
}

static void las_emf_corr(const oracle_amr *g, const int *tl, int s, double dt, double **E, double **acc, int rule)
{
    for (int stage = 0; stage < 2; stage++)
#pragma omp parallel for schedule(dynamic, 1)
        for (int b = 0; b < g->nblk; b++) {
            const oracle_amr_block *C = &g->blk[b];
            const long NC = geo_of(C).ncell;
            const int tr = tl[b];
            if (stage == 1 && oracle_las_phase(tr, s) != 2) continue;
            if (rule == 1 && C->level + 1 >= g->nlev) continue;
            for (int t = 0; t < 3; t++) {
                const int p = (t + 1) % 3, q = (t + 2) % 3;
                int K[3];
                for (K[t] = C->o[t]; K[t] < C->o[t] + C->bs[t]; K[t]++)
                    for (K[p] = C->o[p]; K[p] <= C->o[p] + C->bs[p]; K[p]++)
                        for (K[q] = C->o[q]; K[q] <= C->o[q] + C->bs[q]; K[q]++) {
                            if (K[p] != C->o[p] && K[p] != C->o[p] + C->bs[p] && K[q] != C->o[q] &&
                                K[q] != C->o[q] + C->bs[q])
                                continue;
                            int Z[3];
                            const int m = edge_sender(g, tl, b, t, K, rule, Z);
                            if (m < 0) continue;
                            const long cc = (long)t * NC + cell_at(C, K);
                            if (stage == 1) {
                                E[b][cc] = acc[b][cc] / (dt * (double)tr);
                                continue;
                            }
                            const int ts = tl[m];
                            if (oracle_las_phase(ts, s) != 2) continue;
                            const int first = s % (2 * tr) == 2 * ts - 1;
                            const oracle_amr_block *S = &g->blk[m];
                            const double *e = E[m] + (long)t * geo_of(S).ncell;
                            double sum, c;
                            if (rule == 1) {
                                int Z1[3] = {Z[0], Z[1], Z[2]};
                                Z1[t] += 1;
                                sum = e[cell_at(S, Z)] + e[cell_at(S, Z1)];
                                c = 0.5 * (dt * (double)ts);
                            } else {
                                sum = e[cell_at(S, Z)];
                                c = 1.0 * (dt * (double)ts);
                            }
                            acc[b][cc] = first ? c * sum : acc[b][cc] + c * sum;
                        }
            }
        }
}

double oracle_amr_las_cycle(const oracle_amr *g, const int *tl, int maxtl, double **p, double **ps, double dt,
                            int **pflag, long long *failcount, double *bdt_out, double **Fint, double **Eint)
{
    const int nb = g->nblk;
    const int nsub = 2 * maxtl;
    double **ph = calloc(nb, sizeof(double *)), **psh = calloc(nb, sizeof(double *));
    double **F1 = calloc(nb, sizeof(double *)), **F2 = calloc(nb, sizeof(double *)), **F3 = calloc(nb, sizeof(double *));
    double **U0 = calloc(nb, sizeof(double *)), **E = calloc(nb, sizeof(double *)), **Ec = calloc(nb, sizeof(double *));
    double **d1 = calloc(nb, sizeof(double *)), **d2 = calloc(nb, sizeof(double *)), **d3 = calloc(nb, sizeof(double *));
    double **a1 = calloc(nb, sizeof(double *)), **a2 = calloc(nb, sizeof(double *)), **a3 = calloc(nb, sizeof(double *));
    double **aes = calloc(nb, sizeof(double *)), **aej = calloc(nb, sizeof(double *));
    double **rp = calloc(nb, sizeof(double *)), **rps = calloc(nb, sizeof(double *));
    double **op = calloc(nb, sizeof(double *)), **ops = calloc(nb, sizeof(double *));
    int *act = calloc(nb, sizeof(int)), *phs = calloc(nb, sizeof(int));
    double ndt = -1.;
    int ok = ph && psh && F1 && F2 && F3 && U0 && E && Ec && d1 && d2 && d3 && a1 && a2 && a3 && aes && aej && rp &&
             rps && op && ops && act && phs;
    for (int b = 0; ok && b < nb; b++) {
        const long N = geo_of(&g->blk[b]).ncell;
        if (tl[b] < 1 || tl[b] > maxtl || (tl[b] & (tl[b] - 1))) ok = 0;
        ph[b] = malloc(sizeof(double) * N * NPR);
        psh[b] = malloc(sizeof(double) * N * 3);
        F1[b] = calloc(N * NPR, sizeof(double));
        F2[b] = calloc(N * NPR, sizeof(double));
        F3[b] = calloc(N * NPR, sizeof(double));
        U0[b] = calloc(N * NPR, sizeof(double));
        E[b] = calloc(N * 3, sizeof(double));
        Ec[b] = calloc(N * 3, sizeof(double));
        d1[b] = malloc(sizeof(double) * N);
        d2[b] = malloc(sizeof(double) * N);
        d3[b] = malloc(sizeof(double) * N);
        a1[b] = calloc(N * NPR, sizeof(double));
        a2[b] = calloc(N * NPR, sizeof(double));
        a3[b] = calloc(N * NPR, sizeof(double));
        aes[b] = calloc(N * 3, sizeof(double));
        aej[b] = calloc(N * 3, sizeof(double));
        ok = ok && ph[b] && psh[b] && F1[b] && F2[b] && F3[b] && U0[b] && E[b] && Ec[b] && d1[b] && d2[b] && d3[b] &&
             a1[b] && a2[b] && a3[b] && aes[b] && aej[b];
        if (ok) {
            memcpy(ph[b], p[b], sizeof(double) * N * NPR);
            memcpy(psh[b], ps[b], sizeof(double) * N * 3);
        }
    }
    if (ok) {
        double **F[3] = {F1, F2, F3}, **acc[3] = {a1, a2, a3};
        for (int s = 0; s < nsub; s++) {
            const int last = s == nsub - 1;
            for (int b = 0; b < nb; b++) {
                phs[b] = oracle_las_phase(tl[b], s);
                act[b] = phs[b] != 0;
            }
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < nb; b++) {
                if (!act[b]) continue;
                oracle_params par = block_params(g, b);
                const double *pi = phs[b] == 1 ? p[b] : ph[b], *psi = phs[b] == 1 ? ps[b] : psh[b];
                oracle_flux(&par, 1, pi, psi, &g->m[b], F1[b], last ? d1[b] : NULL);
                oracle_flux(&par, 2, pi, psi, &g->m[b], F2[b], last ? d2[b] : NULL);
                oracle_flux(&par, 3, pi, psi, &g->m[b], F3[b], last ? d3[b] : NULL);
            }
            if (oracle_las_plant != 1) las_flux_corr(g, tl, s, dt, F, acc);
            if (Fint)
                for (int b = 0; b < nb; b++) {
                    if (phs[b] != 2) continue;
                    const long n = geo_of(&g->blk[b]).ncell * NPR;
                    const double Dt = dt * (double)tl[b];
                    for (int d = 0; d < 3; d++)
                        for (long k = 0; k < n; k++) Fint[3 * b + d][k] += Dt * F[d][b][k];
                }
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < nb; b++) {
                if (!act[b]) continue;
                oracle_params par = block_params(g, b);
                oracle_emf(&par, phs[b] == 1 ? p[b] : ph[b], F1[b], F2[b], F3[b], &g->m[b], dt * (double)tl[b], E[b],
                           Ec[b]);
            }
            ring_average_las(g, E, dt, tl, s);
            if (oracle_las_plant != 2) las_emf_corr(g, tl, s, dt, E, aes, 0);
            las_emf_corr(g, tl, s, dt, E, aej, 1);
            if (Eint)
                for (int b = 0; b < nb; b++) {
                    if (phs[b] != 2) continue;
                    const long n = geo_of(&g->blk[b]).ncell * 3;
                    const double Dt = dt * (double)tl[b];
                    for (long k = 0; k < n; k++) Eint[b][k] += Dt * E[b][k];
                }
            for (int b = 0; b < nb; b++) {
                if (!act[b]) continue;
                oracle_params par = block_params(g, b);
                const double Dt = phs[b] == 1 ? 0.5 * dt * (double)tl[b] : dt * (double)tl[b];
                if (phs[b] == 1) {
                    oracle_ct_update(&par, Dt, ps[b], psh[b], E[b], &g->m[b]);
                    oracle_update(&par, 0, Dt, p[b], p[b], ph[b], U0[b], psh[b], F1[b], F2[b], F3[b], &g->m[b],
                                  pflag[b], failcount);
                } else {
                    oracle_ct_update(&par, Dt, ps[b], ps[b], E[b], &g->m[b]);
                    oracle_update(&par, 1, Dt, p[b], ph[b], p[b], U0[b], ps[b], F1[b], F2[b], F3[b], &g->m[b],
                                  pflag[b], failcount);
                }
            }
            /* ghost fill of the stepping blocks */
            for (int b = 0; b < nb; b++) {
                rp[b] = phs[b] == 1 ? ph[b] : p[b];
                rps[b] = phs[b] == 1 ? psh[b] : ps[b];
                op[b] = phs[b] == 1 ? p[b] : ph[b];
                ops[b] = phs[b] == 1 ? ps[b] : psh[b];
            }
            const las_x lx = {act, op, ops};
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < nb; b++)
                if (act[b]) outflow(g, b, rp, rps, pflag, 0);
            exchange_stage(g, 1, rp, rps, pflag, &lx);
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < nb; b++)
                if (act[b]) polar(g, b, rp, rps, pflag);
            exchange_stage(g, 2, rp, rps, pflag, &lx);
            exchange_stage(g, 3, rp, rps, pflag, &lx);
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < nb; b++)
                if (act[b]) outflow(g, b, rp, rps, pflag, 1);
            if (last) {
                ndt = 1e9;
                for (int b = 0; b < nb; b++) {
                    oracle_params par = block_params(g, b);
                    const double bd = oracle_block_dt(&par, d1[b], d2[b], d3[b]);
                    if (bdt_out) bdt_out[b] = bd;
                    if (bd < ndt) ndt = bd;
                }
                if (ndt > 1.3 * dt) ndt = 1.3 * dt;
            }
        }
    }
    for (int b = 0; b < nb; b++) {
        if (ph) free(ph[b]);
        if (psh) free(psh[b]);
        if (F1) free(F1[b]);
        if (F2) free(F2[b]);
        if (F3) free(F3[b]);
        if (U0) free(U0[b]);
        if (E) free(E[b]);
        if (Ec) free(Ec[b]);
        if (d1) free(d1[b]);
        if (d2) free(d2[b]);
        if (d3) free(d3[b]);
        if (a1) free(a1[b]);
        if (a2) free(a2[b]);
        if (a3) free(a3[b]);
        if (aes) free(aes[b]);
        if (aej) free(aej[b]);
    }
    free(ph); free(psh); free(F1); free(F2); free(F3); free(U0); free(E); free(Ec); free(d1); free(d2); free(d3);
    free(a1); free(a2); free(a3); free(aes); free(aej); free(rp); free(rps); free(op); free(ops); free(act); free(phs);
    return ndt;
}
