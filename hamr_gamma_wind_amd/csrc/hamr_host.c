/*
 * hamr_host.c -- the reference's host-side step, in C, over the C ABI of
 * libhamr_hip.so (libhamr_host.so).
 *
 * hamr_host_advance is advance_GPU / GPU_step_ch (step_ch.c:916-1336) with
 * every block at timelevel 1: per half step, the fluxes of each block
 * (GPU_fluxcalc2D, dir 3, 2, 1), each block's Courant dt read back to the
 * host (fluxcalc_GPU), the flux correction at level jumps in the full step
 * (flux_send / flux_rec), the EMFs of each block, the polar E1 average, the
 * face-field update of each block, the conserved update and inversion of each
 * block, and the ghost fill; the next dt is the minimum over blocks of the
 * per-block harmonic of the corrector's direction minima (step_ch.c:1313-1322),
 * clamped to 1.3 dt.  It is the per-block call structure the reference's host
 * drives, kept as the drop-in host of the C ABI; the batched entry points
 * (hamr_step, hamr_step_graph) replace it with one stream of whole-set launches.
 */
#include <stdlib.h>

#include "../../include/hamr_hip.h"

int hamr_host_advance(hamr_ctx *ctx, int nblk, double dt, double *ndt_out)
{
    double *bdt = (double *)malloc(sizeof(double) * 3 * (size_t)(nblk > 0 ? nblk : 1));
    double ndt = 1.e9;
    int r = 0, flag, n, d;
    if (!bdt) return -1;
    for (flag = 0; flag <= 1 && !r; flag++) {
        const double Dt = flag ? dt : 0.5 * dt;
        for (d = 3; d >= 1 && !r; d--)
            for (n = 0; n < nblk && !r; n++) r = hamr_fluxcalc(ctx, d, flag, n);
        for (n = 0; n < nblk && !r; n++)
            for (d = 1; d <= 3 && !r; d++) r = hamr_block_dt(ctx, n, d, &bdt[3 * n + d - 1]);
        if (!r) r = hamr_flux_correct(ctx, flag, dt);
        for (n = 0; n < nblk && !r; n++) r = hamr_ct_emf(ctx, flag, n);
        if (!r) r = hamr_ct_bound(ctx, dt);
        if (!r) r = hamr_emf_correct(ctx, flag, dt);
        for (n = 0; n < nblk && !r; n++) r = hamr_ct_update(ctx, flag, Dt, n);
        for (n = 0; n < nblk && !r; n++) r = hamr_update(ctx, flag, n, Dt);
        if (!r) r = hamr_boundprim(ctx, flag ? 1 : 0);
    }
    if (!r) {
        for (n = 0; n < nblk; n++) {
            const double b0 = 1. / (1. / bdt[3 * n] + 1. / bdt[3 * n + 1] + 1. / bdt[3 * n + 2]);
            if (b0 < ndt) ndt = b0;
        }
        if (ndt > 1.3 * dt) ndt = 1.3 * dt;
        r = hamr_sync(ctx);
    }
    free(bdt);
    if (ndt_out) *ndt_out = ndt;
    return r;
}
