/*
 * hamr_rdump_oracle.c -- CPU restatement of the restart record of one block
 * (rdump_block_write / rdump_block_read, restart.c:52-67 and :178-269) and of
 * its dump record (dump_block, dump.c:276-356).
 *
 * TEST INFRASTRUCTURE ONLY (what the HIP record kernels are tested against; see
 * hamr_oracle.h).
 *
 * Arrays are one block's ghost-inclusive grid in the layout of hamr_oracle.h
 * (p[k*ncell + cell], ps[d*ncell + cell], gdet[pos*ncol + col] with
 * ncol = D1*D2).  The reader follows the reference's own loop: it walks the
 * record in file order, block-relative fine index r = i - N1_GPU_offset*red_1
 * (restart.c:233, the block offset drops out of every test below since it is a
 * multiple of red), and scatters each fine cell into its coarse cell.
 */
#include "hamr_oracle.h"

#include <math.h>

#define RREC (ORC_NPR + 4)

/* restart.c:55-63: p[0..8], ps[0], ps[1] gdet(FACE1), ps[2] gdet(FACE2),
 * ps[3] gdet(FACE3) per cell, cells in index_3D order.  This layout keeps no
 * face-field component 0: its slot is written as 0 (the reader skips it). */
void oracle_rdump_block_write(const int bs[3], const double *p, const double *ps, const double *gdet, double *rec)
{
    const int D1 = bs[0] + 2 * ORC_NG, D2 = bs[1] + 2 * ORC_NG, D3 = bs[2] + 2 * ORC_NG;
    const long nc = (long)D1 * D2 * D3, ncol = (long)D1 * D2;
    for (int i = 0; i < D1; i++)
        for (int j = 0; j < D2; j++)
            for (int z = 0; z < D3; z++) {
                const long cell = ((long)i * D2 + j) * D3 + z, col = (long)i * D2 + j;
                double *r = rec + cell * RREC;
                for (int k = 0; k < ORC_NPR; k++) r[k] = p[k * nc + cell];
                r[ORC_NPR + 0] = 0.;
                r[ORC_NPR + 1] = ps[0 * nc + cell] * gdet[0 * ncol + col]; /* FACE1 */
                r[ORC_NPR + 2] = ps[1 * nc + cell] * gdet[1 * ncol + col]; /* FACE2 */
                r[ORC_NPR + 3] = ps[2 * nc + cell] * gdet[4 * ncol + col]; /* FACE3 */
            }
}

/* restart.c:233-269 with NPR_U = NPR, STAGGERED, no optional physics.  The
 * record's block is red[d] times finer (its D_d = bs_d red_d + 2 NG).  Cells
 * whose fine index r_d lies in [0, bs_d red_d] along all d (restart.c:244) are
 * read; the coarse cell is r_d / red_d (block-relative, + NG in the arrays).
 * Visiting the first fine cell of a coarse cell (all r_d % red_d == 0) zeroes
 * it; then p += v reduce_factor and each face-field component accumulates the
 * fine faces on the coarse face (r_d % red_d == 0 along its own direction),
 * divided by this grid's gdet.  spherical: B^2 is also scaled by
 * fractheta / fractheta_old (restart.c:249-257), with this grid's
 * fractheta = 1 - 2/n2 and the record's 1 - 2/(n2 red_2) (the x2 boundary of
 * the 3D spherical grid is transmissive, coord.c:397). */
void oracle_rdump_block_read(const int bs[3], const int red[3], int spherical, int n2, const double *rec,
                             const double *gdet, double *p, double *ps)
{
    const int D1 = bs[0] + 2 * ORC_NG, D2 = bs[1] + 2 * ORC_NG, D3 = bs[2] + 2 * ORC_NG;
    const long nc = (long)D1 * D2 * D3, ncol = (long)D1 * D2;
    const int R1 = bs[0] * red[0] + 2 * ORC_NG, R2 = bs[1] * red[1] + 2 * ORC_NG, R3 = bs[2] * red[2] + 2 * ORC_NG;
    const double fractheta = 1.0 - 2.0 / ((double)n2) * 1;
    const double fractheta_old = 1.0 - 2.0 / ((double)n2 * red[1]) * 1;
    const double *v = rec;
    for (int fi = 0; fi < R1; fi++)
        for (int fj = 0; fj < R2; fj++)
            for (int fz = 0; fz < R3; fz++, v += RREC) {
                const int r1 = fi - ORC_NG, r2 = fj - ORC_NG, r3 = fz - ORC_NG;
                if (!(r1 >= 0 && r1 <= bs[0] * red[0] && r2 >= 0 && r2 <= bs[1] * red[1] && r3 >= 0 &&
                      r3 <= bs[2] * red[2]))
                    continue;
                const int i = ORC_NG + r1 / red[0], j = ORC_NG + r2 / red[1], z = ORC_NG + r3 / red[2];
                const long cell = ((long)i * D2 + j) * D3 + z, col = (long)i * D2 + j;
                if (r1 % red[0] == 0 && r2 % red[1] == 0 && r3 % red[2] == 0) {
                    for (int k = 0; k < ORC_NPR; k++) p[k * nc + cell] = 0.0;
                    for (int d = 0; d < 3; d++) ps[d * nc + cell] = 0.0;
                }
                double rf = 1.0 / (double)(red[0] * red[1] * red[2]);
                for (int k = 0; k < ORC_NPR; k++) p[k * nc + cell] += v[k] * rf;
                rf = 1.0 / (double)(red[1] * red[2]);
                if (r1 % red[0] == 0) ps[0 * nc + cell] += v[RREC - 3] * rf / gdet[0 * ncol + col];
                rf = 1.0 / (double)(red[0] * red[2]);
                if (r2 % red[1] == 0) {
                    if (spherical)
                        ps[1 * nc + cell] += v[RREC - 2] * rf / gdet[1 * ncol + col] * fractheta / fractheta_old;
                    else
                        ps[1 * nc + cell] += v[RREC - 2] * rf / gdet[1 * ncol + col];
                }
                rf = 1.0 / (double)(red[0] * red[1]);
                if (r3 % red[2] == 0) ps[2 * nc + cell] += v[RREC - 1] * rf / gdet[4 * ncol + col];
            }
}

/* dump_block (dump.c:276-356; ideal gas, CALC_MDOT and the optional physics
 * off): 9 floats per interior cell in index_3D order: rho, u, ucon[0..3]
 * (ucon_calc, phys.c:309-331, with gamma_calc :359 and the centre metric of the
 * cell's column, gcov/gcon packed [col*10 + s]) and the cell-centred B^1..B^3. */
void oracle_dump_block(const int bs[3], const double *p, const double *gcov_cent, const double *gcon_cent, float *rec)
{
    const int D1 = bs[0] + 2 * ORC_NG, D2 = bs[1] + 2 * ORC_NG, D3 = bs[2] + 2 * ORC_NG;
    const long nc = (long)D1 * D2 * D3;
    long o = 0;
    (void)D1;
    for (int i = ORC_NG; i < ORC_NG + bs[0]; i++)
        for (int j = ORC_NG; j < ORC_NG + bs[1]; j++)
            for (int z = ORC_NG; z < ORC_NG + bs[2]; z++, o += 9) {
                const long cell = ((long)i * D2 + j) * D3 + z, col = (long)i * D2 + j;
                const double *gc = gcov_cent + col * 10, *gi = gcon_cent + col * 10;
                double pr[ORC_NPR], alpha, beta[4], qsq, gamma, ucon[4];
                for (int k = 0; k < ORC_NPR; k++) pr[k] = p[k * nc + cell];
                alpha = 1. / sqrt(-gi[0]);
                for (int a = 1; a < 4; a++) beta[a] = gi[a] * alpha * alpha;
                /* gcov pairs 11 12 13 22 23 33 at 4 5 6 7 8 9 */
                qsq = gc[4] * pr[2] * pr[2] + gc[7] * pr[3] * pr[3] + gc[9] * pr[4] * pr[4] +
                      2. * (gc[5] * pr[2] * pr[3] + gc[6] * pr[2] * pr[4] + gc[8] * pr[3] * pr[4]);
                if (qsq < 0.) qsq = (fabs(qsq) > 1.E-10) ? -1. : 1.E-10; /* gamma = 1 on failure */
                gamma = (qsq == -1.) ? 1. : sqrt(1. + qsq);
                ucon[0] = gamma / alpha;
                for (int a = 1; a < 4; a++) ucon[a] = pr[1 + a] - gamma * beta[a] / alpha;
                rec[o + 0] = (float)pr[0];
                rec[o + 1] = (float)pr[1];
                for (int a = 0; a < 4; a++) rec[o + 2 + a] = (float)ucon[a];
                rec[o + 6] = (float)pr[5];
                rec[o + 7] = (float)pr[6];
                rec[o + 8] = (float)pr[7];
            }
}
