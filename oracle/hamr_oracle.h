/*
 * hamr_oracle.h -- FP64 CPU restatement of H-AMR's GRMHD fluid step.
 *
 * TEST INFRASTRUCTURE ONLY: the HIP step in hamr_gamma_wind_amd/ is tested
 * against this library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * It works on ONE global grid with NG=3 ghost zones (whereas the reference and
 * the HIP path work block by block); a correct multi-block step reproduces it,
 * because every block-boundary face, corner and ghost value is a function of the
 * global state only.  The block sizes enter only where the reference's own
 * semantics depend on them: the per-block harmonic Courant dt and the polar E1
 * average (an average of per-block averages).
 *
 * Array layout (all double, structure of arrays):
 *   cell(i,j,z) = (i*D2 + j)*D3 + z  with D_d = n_d + 6 and i,j,z in [0,D_d)
 *   (ghost-inclusive; interior starts at 3; z fastest, as the reference's index_3D)
 *   prims   p [k*ncell + cell]      k = RHO,UU,U1,U2,U3,B1,B2,B3,KTOT
 *   face B  ps[d*ncell + cell]      d = 0,1,2 -> B^1 at FACE1, B^2 at FACE2, B^3 at FACE3
 *   fluxes  F [k*ncell + cell]      flux through the lower face of the cell
 *   EMF     E [d*ncell + cell]      d = 0,1,2 -> corner E_1, E_2, E_3
 *   metric  col(i,j) = i*D2 + j, positions FACE1=0 FACE2=1 CORN=2 CENT=3 FACE3=4
 *           gcov[(pos*ncol + col)*10 + s], gcon likewise, s over the symmetric
 *           pairs 00,01,02,03,11,12,13,22,23,33;  gdet[pos*ncol + col];
 *           conn[col*40 + lam*10 + s] = Gamma^lam_{s} at CENT;  rcen[col] = r at CENT.
 */
#ifndef HAMR_ORACLE_H
#define HAMR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NPR 9
#define ORC_NG 3

/* ORC_BC_POLE: both x2 faces of the grid are on the axis (with the polar E1
 * ring average over the grid's phi range).  ORC_BC_POLE_LO / _HI: only the
 * lower / upper x2 face is on the axis and the ring average is left to the
 * caller (one block of a multi-block ring); ORC_BC_NONE: a block-internal face. */
enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1, ORC_BC_POLE = 2, ORC_BC_POLE_LO = 3, ORC_BC_POLE_HI = 4,
       ORC_BC_NONE = 5 };

typedef struct oracle_params {
    int n1, n2, n3;        /* interior cells of the global grid */
    int bs1, bs2, bs3;     /* block interior size (per-block dt, polar average) */
    int bc[3];             /* boundary type per direction (ORC_BC_*) */
    double gam;            /* adiabatic index */
    double cour;           /* Courant factor */
    double dx[3];          /* code-coordinate spacings dx1, dx2, dx3 */
    double rho_floor;      /* rho >= rho_floor * r^-1.5  (0 disables) */
    double u_floor;        /* u   >= u_floor   * r^-2.5  (0 disables) */
    int riemann;           /* 0: HLL, 1: HLLC in the face's orthonormal frame */
} oracle_params;

typedef struct oracle_metric {
    const double *gcov, *gcon, *gdet, *conn, *rcen;
} oracle_metric;

/* ---- pointwise physics (for golden-vector tests) ---- */
void oracle_get_state(const double *pr, const double *gcov, const double *gcon,
                      double *ucon, double *ucov, double *bcon, double *bcov);
void oracle_primtoflux(const double *pr, const double *gcov, const double *gcon, double g,
                       double gam, int dir, double *flux, double *vmax, double *vmin);
void oracle_source(const double *pr, const double *gcov, const double *gcon, double g,
                   const double *conn40, double gam, double *dU);
void oracle_para(const double *x5, double *l, double *r);
/* which: 0 = Noble 2D, 1 = 1D v^2 entropy fix, 2 = 1D fix.  prim is the guess
 * on input and the solution on success; returns the reference's return code. */
int oracle_utoprim(int which, const double *U, const double *gcov, const double *gcon,
                   double g, double gam, double *prim);

/* set_Mud (hllc.c:20-178): orthonormal tetrad at a face of direction dir from
 * the packed gcov/gcon there; mud[mu*4 + a] = e_(a)^mu, mudi[a*4 + mu] = e^(a)_mu. */
void oracle_tetrad(const double *gcov, const double *gcon, int dir, double *mud, double *mudi);
/* fluxcalc2D_FT (GPU_hllc.h:3-247) at one face from the reconstructed states
 * pl, pr (normal field already set) and the face's metric: the 9 coordinate
 * fluxes F (gdet included); returns the Courant speed ctop. */
double oracle_hllc_face(const double *pl, const double *pr, const double *gcov, const double *gcon, double g,
                        double gam, int dir, double *F);

/* ---- grid stages ---- */
/* HLL flux through the lower dir-faces; dtface (may be NULL) gets cour*dx/ctop
 * per face (0 where no face is computed). */
void oracle_flux(const oracle_params *par, int dir, const double *p, const double *ps,
                 const oracle_metric *m, double *F, double *dtface);
/* E_cent + upwinded corner EMF + polar E1 average (factor = dt of the step). */
void oracle_emf(const oracle_params *par, const double *pb, const double *F1,
                const double *F2, const double *F3, const oracle_metric *m,
                double dt_factor, double *E, double *Ecent_scratch);
void oracle_ct_update(const oracle_params *par, double Dt, const double *psi, double *psf,
                      const double *E, const oracle_metric *m);
/* full=0: U0 <- U(pi), source from pi, guess pi;  full=1: U from U0, source and guess from pb. */
/* The conserved update alone over the interior (predictor form, step Dt):
 * U(pi) into U0, U(pi) - Dt div F + Dt S(pi) with the face average of psf into Unew. */
void oracle_cons(const oracle_params *par, double Dt, const double *pi, double *U0, double *Unew,
                 const double *psf, const double *F1, const double *F2, const double *F3, const oracle_metric *m);
void oracle_update(const oracle_params *par, int full, double Dt, const double *pi,
                   const double *pb, double *pf, double *U0, const double *psf,
                   const double *F1, const double *F2, const double *F3,
                   const oracle_metric *m, int *pflag, long long *failcount);
/* The inversion cascade, floors and flags alone over the interior, on the
 * conserved variables Unew with the guess pi (full = 0) or pb (full = 1). */
void oracle_invert(const oracle_params *par, int full, const double *pi, const double *pb, double *pf_out,
                   const double *Unew, const oracle_metric *m, int *pflag, long long *failcount);
void oracle_bound(const oracle_params *par, double *p, double *ps, const oracle_metric *m);
/* Per-block harmonic dt from the three dtface arrays, minimum over blocks. */
double oracle_block_dt(const oracle_params *par, const double *dt1, const double *dt2,
                       const double *dt3);

/* One predictor + corrector step.  p/ps are updated in place (ghosts filled),
 * pflag gets the corrector's flags; returns dt for the next step
 * (min over blocks, clamped to 1.3*dt), or a negative value on allocation failure. */
double oracle_step(const oracle_params *par, double *p, double *ps, const oracle_metric *m,
                   double dt, int *pflag, long long *failcount);

/* inflow_check (GPU_bounds.h:1368-1400) on one primitive vector. */
void oracle_inflow_check(double *pr, const double *gcov, const double *gcon, int type);

/* ---- multi-level block sets (hamr_amr_oracle.c) ----
 * Blocks of uneven sizes on levels of ratio 2, each its own ghost-inclusive
 * grid in the layout above (p[b], ps[b], pflag[b]; metric m[b] in the layout
 * above with ncol = D1*D2 of the block).  The ghost fill is the library's
 * specification of the multi-level exchange (copies, 2x2x2 means, MC-limited
 * linear prolongation; see hamr_amr_oracle.c), found here by searching the
 * block list for the leaf that holds each position. */
typedef struct oracle_amr_block {
    int level, o[3], bs[3];   /* level, first interior cell and size in cells of the level */
} oracle_amr_block;

typedef struct oracle_amr {
    int nblk, nlev;
    int N[3][3];              /* cells of each level */
    const oracle_amr_block *blk;
    const oracle_metric *m;   /* per block */
    double gam, cour, dx0[3], rho_floor, u_floor;
    int reflux;               /* 1: flux and EMF correction at level jumps in the corrector */
    int riemann;              /* 0: HLL, 1: HLLC */
} oracle_amr;

/* Ghost fill of every block (p, ps, pflag: arrays of per-block pointers). */
void oracle_amr_bound(const oracle_amr *g, double **p, double **ps, int **pflag);
/* Only the exchange part of stage `stage` (1..3): the ghosts of internal faces. */
void oracle_amr_exchange(const oracle_amr *g, int stage, double **p, double **ps, int **pflag);
/* Flux correction at level jumps: each face of a block whose neighbour
 * across it is one level finer takes the mean of the four finer faces
 * (F1..F3: per-block flux arrays, dt the step's). */
void oracle_amr_reflux(const oracle_amr *g, double **F1, double **F2, double **F3, double dt);
/* EMF correction at level jumps: every edge EMF of a block boundary that a
 * one level finer leaf touches takes the mean of that leaf's two edges along
 * it (E: per-block corner EMF arrays, dt the step's). */
void oracle_amr_emf_correct(const oracle_amr *g, double **E, double dt);
/* The corrector's face-field update of every block from (p, ps) (ghosts
 * filled): fluxes and EMFs (corrected at level jumps when g->reflux), the polar
 * ring average and the CT update with dt from ps into psf (no ghost fill). */
void oracle_amr_ct(const oracle_amr *g, double **p, double **ps, double dt, double **psf);
/* Fluxes of every block from (p, ps) (ghosts filled), corrected at level
 * jumps when g->reflux (the corrector's fluxes of oracle_amr_step). */
void oracle_amr_fluxes(const oracle_amr *g, double **p, double **ps, double dt, double **F1, double **F2, double **F3);
/* One predictor + corrector step of every block (ghosts filled on input;
 * pflag required);
 * returns the next dt (min over blocks, clamped to 1.3 dt), < 0 on allocation failure. */
double oracle_amr_step(const oracle_amr *g, double **p, double **ps, double dt, int **pflag, long long *failcount);

/* Phase of a block of time level tl in sub-iteration s of a local-adaptive-
 * time-stepping cycle: 1 predictor (s % tl == tl - 1), 2 corrector (s % 2tl
 * == 2tl - 1, checked first), 0 rest. */
int oracle_las_phase(int tl, int s);
/* One local-adaptive-time-stepping cycle of 2 maxtl sub-iterations (time
 * levels tl[b]: powers of 2 up to maxtl; a finer block at a level jump no
 * slower than its coarser neighbour), ghosts filled on input, pflag required.
 * Returns the next dt (< 0 on a bad time level or allocation failure);
 * bdt_out (may be NULL) receives each block's harmonic Courant dt of the last
 * sub-iteration; Fint (may be NULL; 3 nblk arrays of NPR ncell, Fint[3 b + d]
 * the fluxes of direction d + 1 of block b) accumulates Dt F of every
 * corrector's (corrected) fluxes; Eint (may be NULL; nblk arrays of 3 ncell)
 * the same of the corrector's (corrected) edge EMFs. */
double oracle_amr_las_cycle(const oracle_amr *g, const int *tl, int maxtl, double **p, double **ps, double dt,
                            int **pflag, long long *failcount, double *bdt_out, double **Fint, double **Eint);
extern int oracle_las_plant;

/* ---- restart record of one block (hamr_rdump_oracle.c) ----
 * One block's ghost-inclusive arrays in the layout above (bs: interior size;
 * gdet with ncol = D1*D2).  The record holds NPR+4 doubles per cell. */
void oracle_rdump_block_write(const int bs[3], const double *p, const double *ps, const double *gdet, double *rec);
/* rec: the record of a block red[d] times finer along d; n2: this grid's total
 * x2 cells; spherical: the torus grid (B^2 theta-fraction ratio). */
void oracle_rdump_block_read(const int bs[3], const int red[3], int spherical, int n2, const double *rec,
                             const double *gdet, double *p, double *ps);
/* dump record of one block: 9 floats per interior cell (rho, u, u^0..u^3,
 * B^1..B^3); gcov/gcon at the cell centres, packed [col*10 + s], ncol = D1*D2. */
void oracle_dump_block(const int bs[3], const double *p, const double *gcov_cent, const double *gcon_cent, float *rec);

#ifdef __cplusplus
}
#endif
#endif
