// hamr_kernels.h -- kernel parameters and launchers of the gfx950 step.
#pragma once

#include <hip/hip_runtime.h>

#include "hamr_grid.h"
#include "hamr_phys.h"

namespace hamr {

// Passed by value to every kernel.
struct KParams {
    int D[3];              // block extent including ghosts
    int bs[3];             // block interior
    int64_t ksize;         // cells per component slab
    int64_t kcol;          // columns per metric slab
    int nb3;               // blocks around the axis (pole partner shift)
    double gam, cour;
    double dx[3];
    double rho_floor, u_floor;
    const BlockDev *blk;
    const double *gcov, *gcon, *gdet, *conn, *rcen;
};

// `mixed` selects the arithmetic of reconstruction, fluxes and EMFs: 0 = FP64
// (F, Ec, E are double arrays), 1 = FP32 (float arrays with the same slab layout).
void launch_flux(const KParams &P, int dir, const int *blist, int nb, const double *p, const double *ps, void *F,
                 unsigned long long *dtmin, int mixed, hipStream_t s);
void launch_emf(const KParams &P, const int *blist, int nb, const double *p, const void *F1, const void *F2,
                const void *F3, void *Ec, void *E, int mixed, hipStream_t s);
void launch_pole(const KParams &P, const int *ring, int nring, double dt_val, const double *dt_ptr, void *E, int mixed,
                 hipStream_t s);
void launch_ct_update(const KParams &P, const int *blist, int nb, double Dt_val, const double *dt_ptr,
                      double dt_mult, const double *psi, double *psf, const void *E, int mixed, hipStream_t s);
void launch_update(const KParams &P, const int *blist, int nb, int full, double Dt_val, const double *dt_ptr,
                   double dt_mult, const double *pi, const double *pb, double *pf, double *U0, double *Unew,
                   const double *psf, const void *F1, const void *F2, const void *F3, int *pflag, int *failcnt,
                   int mixed, hipStream_t s);
// Conserved update only (k_cons, predictor form with step Dt): U(p) -> U0,
// U(p) - Dt div F + Dt S(p) with the face average of psf -> Unew.
void launch_cons(const KParams &P, const int *blist, int nb, double Dt, const double *p, double *U0, double *Unew,
                 const double *psf, const void *F1, const void *F2, const void *F3, int mixed, hipStream_t s);
// Adjoint of one forward stage of a single periodic flat block (FP64): from the
// weights w of U' (9 slabs) to lP (9 slabs, dL/dP) and lps (3 slabs, dL/d face
// field); F1-F3 are the stage's fluxes; lpsf, lE, lEc, lF1-lF3 are scratch.
void launch_adjoint(const KParams &P, const BlockDev &B, double Dt, const double *p, const double *ps,
                    const double *F1, const double *F2, const double *F3, const double *w, double *lP, double *lps,
                    double *lpsf, double *lE, double *lEc, double *lF1, double *lF2, double *lF3, hipStream_t s);

void launch_bound(const KParams &P, int dir, const int *blist, int nb, double *p, double *ps, int *pflag,
                  hipStream_t s);
void launch_dt_finalize(int nblk, unsigned long long *dtmin, double *dt, hipStream_t s);
void launch_dt_reset(int nblk, unsigned long long *dtmin, hipStream_t s);
void launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t s);

}  // namespace hamr
