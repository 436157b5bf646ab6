/*
 * hamr_hip.h -- C ABI of the MI355X GRMHD fluid step (libhamr_hip.so).
 *
 * A drop-in for the GPU launch layer that H-AMR's step_ch.c drives block by
 * block (GPU_main.cu of tp2872/HAMR_gamma_wind), plus a batched step over the
 * whole resident block set.  Plain C types only; every call returns 0 on
 * success and a negative code on error, with the message in hamr_last_error().
 * All device work is queued on the context's stream (asynchronous); calls that
 * return host data synchronise.  One host thread per context; not re-entrant.
 *
 * Conventions (as the reference's): NPR = 9 primitives RHO, UU, U1, U2, U3, B1,
 * B2, B3, KTOT (config.h:819-827); NG = 3 ghost zones; cell index
 * (i*(BS2+6) + j)*(BS3+6) + z with ghosts, z fastest (index_3D, main.c:355);
 * face fields ps[0..2] = B^1 at FACE1, B^2 at FACE2, B^3 at FACE3 (not
 * gdet-weighted); flag 0 = half (predictor) step reading p/ps and writing
 * ph/psh, flag 1 = full (corrector) step reading ph/psh and writing p/ps.
 */
#ifndef HAMR_HIP_H
#define HAMR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hamr_ctx hamr_ctx;

/* The run description.  Replaces the compile-time block geometry and physics
 * constants of config.h (BS_1..3, NB_1..3, GAMMA, BH_SPIN, COUR, ROUT;
 * config.h:102-105, :355-362, :532) with a runtime descriptor. */
typedef struct hamr_desc {
    int bs[3];          /* block interior cells per direction */
    int nb[3];          /* blocks per direction on the base level */
    double gam;         /* adiabatic index */
    double spin;        /* black-hole spin a */
    double cour;        /* Courant factor */
    double rin;         /* inner radius (<= 0: 0.9 (1 + sqrt(1 - a^2)), coord.c:364) */
    double rout;        /* outer radius */
    double rho_floor;   /* rho >= rho_floor r^-1.5 (0: off) */
    double u_floor;     /* u   >= u_floor   r^-2.5 (0: off) */
    int minkowski;      /* 1: flat, periodic unit box instead of the Kerr-Schild sphere */
    int device;         /* HIP device ordinal */
} hamr_desc;

/* Multi-level block set (no reference counterpart as a descriptor: the
 * reference builds its refined block tree at run time, AMR.c:168-757, with
 * uniform block sizes).  Level ratio 2; level l >= 1 covers
 * [box[l][k][0], box[l][k][1]) cells of level l along direction k (multiples of
 * 32); the base level is the grid of hamr_desc (nb * bs cells, a multiple of
 * 64x32x32).  Each level is tiled by 64x32x32 superblocks, each kept whole
 * (probability p_big), split into 32x16x16 blocks (p_mid) or into 16^3 blocks,
 * by a seeded choice. */
typedef struct hamr_amr_desc {
    int nlev;
    int box[3][3][2];
    unsigned long long seed;
    double p_big, p_mid;
} hamr_amr_desc;

/* Fishbone-Moncrief torus with a poloidal loop (init.c:808-1110, init_mag.c:6). */
typedef struct hamr_torus_desc {
    double rin, rmax, kappa;  /* inner edge, pressure maximum, polytropic constant */
    double beta_min;          /* field normalisation: min over cells of p_gas/p_mag */
    double aphi_cut;          /* A_phi ~ max(rho/rho_max - aphi_cut, 0) */
    double perturb;           /* u *= 1 + perturb (U[0,1) - 1/2), seeded */
    unsigned long long seed;
} hamr_torus_desc;

/* Array selectors of hamr_get_array / hamr_set_array (all blocks at once,
 * layout [block][component][cell] with ncell = (BS1+6)(BS2+6)(BS3+6) per slab). */
enum {
    HAMR_P = 0,     /* primitives, 9 components */
    HAMR_PS = 1,    /* face fields, 3 */
    HAMR_PH = 2,    /* half-step primitives, 9 */
    HAMR_PSH = 3,   /* half-step face fields, 3 */
    HAMR_F1 = 4,    /* fluxes through the lower x1 faces, 9 */
    HAMR_F2 = 5,
    HAMR_F3 = 6,
    HAMR_EMF = 7,   /* corner EMF E_1, E_2, E_3 */
    HAMR_ECENT = 8, /* cell-centre EMF, 3 */
    HAMR_U0 = 9,    /* conserved variables at the start of the step, 9 */
    HAMR_UNEW = 10  /* conserved variables after the last forward stage / update, 9 */
};

/* Arithmetic of the step (hamr_set_precision).  FP64: every stage in double,
 * the reference's arithmetic.  MIXED: reconstruction, fluxes, speeds and EMFs
 * in FP32 (F1-F3, EMF and ECENT are then float arrays on the device; the array
 * calls below convert them), while the metric and connection, the face-field
 * update, the conserved update (a compensated FP64 sum) and the inversions stay
 * FP64. */
enum { HAMR_PREC_FP64 = 0, HAMR_PREC_MIXED = 1 };

/* Riemann solver of the fluxes (hamr_set_riemann): HLL (fluxcalc2D2,
 * GPU_program1.cu:791-1068), or HLLC in the face's orthonormal frame
 * (fluxcalc2D_FT, GPU_hllc.h:3-247; the reference's HLLC = 1, config.h:480). */
enum { HAMR_RIEMANN_HLL = 0, HAMR_RIEMANN_HLLC = 1 };

/* ---- host-only set-up (no device needed; for drivers and tests) ---- */
/* Layout of a descriptor: D[3], cells per slab, padded slab length, metric
 * columns per slab and padded length, block and geometry counts. */
int hamr_desc_layout(const hamr_desc *desc, int *D, long long *ncell, long long *ksize, long long *ncol,
                     long long *kcol, int *nblk, int *ngeo);
/* Block table rows (17 ints per block, as hamr_block_row) and the
 * reference's block numbering (as hamr_block_order); -2 if that numbering
 * cannot be formed. */
int hamr_desc_block_table(const hamr_desc *desc, int *rows, int *order);
/* Grid coordinates startx[3], dx[3] (set_points / set_gridparam, coord.c:327-460). */
int hamr_desc_coords(const hamr_desc *desc, double *startx, double *dx);
/* Metric at code coordinates X[4] (gcov_func, gdet_func, gcon_func, conn_func,
 * metric.c:13-500): packed gcov[10], gcon[10], gdet, conn[40] = Gamma^lam of
 * pair s at lam*10+s, and Boyer-Lindquist r, theta (bl_coord, coord.c:111). */
int hamr_desc_metric_at(const hamr_desc *desc, const double *X, double *gcov, double *gcon, double *gdet,
                        double *conn, double *r_th);
/* The multi-level block set of `amr` without a device (as hamr_create_amr
 * builds it): info[17] as hamr_amr_info; rows (15 ints per block, as
 * hamr_block_layout), the copy/restriction entries of stages 1-3 one after
 * another (4 ints: ghost block, ghost cell, source block, source cell |
 * kind << 28 with kind 1 = cell, 2 = mean of the 2x2x2 cells from there) and
 * the prolongation entries (16 ints: ghost block and cell, coarse block and
 * cell, then (block, code) of the coarse cell's -x1 +x1 -x2 +x2 -x3 +x3
 * neighbours, kind 0 = none); any array may be NULL.  Returns -2 if the
 * description is rejected. */
int hamr_desc_amr(const hamr_desc *desc, const hamr_amr_desc *amr, long long *info, int *rows, int *ex, int *pro);
/* Geometry slab geo of that block set (device layout, kcol = its padded
 * D1*D2; hamr_geometry_info gives the sizes on a context) and its torus initial
 * state (unpadded whole-array layout, as hamr_get_array on a multi-level
 * context), without a device. */
int hamr_desc_amr_geometry(const hamr_desc *desc, const hamr_amr_desc *amr, int geo, double *gcov, double *gcon,
                           double *gdet, double *conn, double *rcen);
int hamr_desc_amr_torus(const hamr_desc *desc, const hamr_amr_desc *amr, const hamr_torus_desc *t, double *p,
                        double *ps);
/* One geometry slab in the device layout (as hamr_upload_metric takes it). */
int hamr_desc_geometry(const hamr_desc *desc, int geo, double *gcov, double *gcon, double *gdet, double *conn,
                       double *rcen);
/* Initial torus state, interior cells only, [block][component][ncell]. */
int hamr_desc_torus(const hamr_desc *desc, const hamr_torus_desc *t, double *p, double *ps);

/* GPU_init (GPU_main.cu:10) + set_AMR (AMR.c:473): allocate the resident block
 * set and the block table. */
int hamr_create(const hamr_desc *desc, hamr_ctx **out);
/* A context over the multi-level block set of `amr` on the torus grid of
 * `desc` (every block resident, one slab per block and array).  All entry
 * points below work on it; the ghost fill (hamr_boundprim) then also exchanges
 * across block sizes and levels: copies, restriction by 2x2x2 means
 * (packsendaverage1, GPU_program2.cu:154-197; the face field normal to the
 * exchange direction as the gdet-weighted mean of the four finer faces over
 * the coarse face's gdet) and MC-limited linear prolongation
 * (unpackreceivecoarse1, GPU_program2.cu:615-803).  No exchange writes the
 * normal face field of a block's own upper boundary face (its CT update owns
 * it; the reference's receive stores no normal component,
 * GPU_program2.cu:376-380).  Host layout
 * of whole arrays: the blocks' [component][cell] slabs one after another, each
 * with its own (bs+6)^3 cell count. */
int hamr_create_amr(const hamr_desc *desc, const hamr_amr_desc *amr, hamr_ctx **out);
/* info[17]: blocks, geometry slabs, padded cells, interior zones, exchange
 * entries (copy/restriction) of stages 1-3, prolongation entries of stages
 * 1-3, polar rings, 1 if multi-level, coarse faces at level jumps along x1,
 * x2, x3 (flux correction entries), coarse edges at level jumps (EMF
 * correction entries), the last stage-3 entries that the step's fill leaves
 * out (x3 ghost rows inside the x1 and x2 range of a block whose x3
 * neighbour has its level and shape: the step's update writes them from that
 * neighbour's cells, as a uniform grid's periodic x3 fill). */
int hamr_amr_info(const hamr_ctx *ctx, long long *info);
/* row[15]: level, origin o1 o2 o3 (cells of its level), bs1 bs2 bs3, geometry
 * slab, pole bits, kinds of faces -x1..+x3 (0 copy, 1 outflow, 2 polar mirror,
 * 3 exchange table). */
int hamr_block_layout(const hamr_ctx *ctx, int n, int *row);
/* info[6]: level, first column o1 o2 (cells of its level, interior origin),
 * D1, D2, kcol of geometry slab geo. */
int hamr_geometry_info(const hamr_ctx *ctx, int geo, long long *info);
int hamr_destroy(hamr_ctx *ctx);
const char *hamr_last_error(const hamr_ctx *ctx);

/* Select the arithmetic of every later stage call and step (no reference
 * counterpart: the reference's precision is fixed at compile time). */
int hamr_set_precision(hamr_ctx *ctx, int mode);
int hamr_get_precision(const hamr_ctx *ctx, int *mode);
/* Select the Riemann solver of every later flux stage and step (the
 * reference's HLLC / FRAME_TRANSFORM compile-time switch, config.h:480 and
 * :1434-1438).  Both solvers run in either arithmetic; the mixed step's HLLC
 * carries the frame energy without its rest-mass part (FP32 would cancel it). */
int hamr_set_riemann(hamr_ctx *ctx, int solver);
int hamr_get_riemann(const hamr_ctx *ctx, int *solver);
/* set_Mud (hllc.c:20-178): the orthonormal tetrad at a face of direction dir
 * (1..3) from the packed gcov/gcon there (pairs 00,01,02,03,11,12,13,22,23,33):
 * mud[mu*4 + a] = e_(a)^mu, mudi[a*4 + mu] = e^(a)_mu.  Host only. */
int hamr_face_tetrad(const double *gcov10, const double *gcon10, int dir, double *mud, double *mudi);
/* The device copy of geometry slab geo's tetrads (computed by
 * hamr_upload_metric / hamr_set_grid): tet[((dir-1)*32 + e)*kcol + col], e < 16
 * mud, e >= 16 mudi; 96*kcol doubles. */
int hamr_download_tetrads(hamr_ctx *ctx, int geo, double *tet);

/* Layout: D[3] = BS+6 extents, ncell, padded slab length, block and geometry counts. */
int hamr_layout(const hamr_ctx *ctx, int *D, long long *ncell, long long *ksize, int *nblk, int *ngeo);
/* Block table row (AMR_coord / neighbour / pole entries of block[n][*], AMR.c:168-757):
 * row[0] geometry slab, [1..3] block coordinates c1 c2 c3, [4] pole bits
 * (1: lower x2 face on the axis, 2: upper), [5..10] neighbour (or axis partner) of
 * faces -x1 +x1 -x2 +x2 -x3 +x3 (-1: physical outflow edge), [11..16] face kinds
 * (0 copy, 1 outflow, 2 polar mirror). */
int hamr_block_row(const hamr_ctx *ctx, int n, int *row17);
/* The reference's block numbering (AMR_set_coord's level-0 Z-order,
 * AMR.c:188-275; n_ord of a single-level grid): order[k] is this context's
 * block (natural index (c1*NB2 + c2)*NB3 + c3) that the reference numbers k,
 * the number its restart file rdump<k> carries (restart.c:41-42).  Also the
 * launch order.  On a multi-level set: the identity. */
int hamr_block_order(const hamr_ctx *ctx, int *order);
/* set_grid (main.c:380) + GPU_write_metric (GPU_main.cu:139): compute the
 * Kerr-Schild metric in code coordinates on the host and upload it. */
int hamr_set_grid(hamr_ctx *ctx);
/* Metric slab of geometry `geo` in the device layout: gcov/gcon[(pos*10+s)*kcol+col],
 * gdet[pos*kcol+col], conn[(lam*10+s)*kcol+col], rcen[col]; kcol = ncol padded to 64. */
int hamr_metric_sizes(const hamr_ctx *ctx, long long *ncol, long long *kcol);
int hamr_upload_metric(hamr_ctx *ctx, int geo, const double *gcov, const double *gcon, const double *gdet,
                       const double *conn, const double *rcen);
int hamr_download_metric(hamr_ctx *ctx, int geo, double *gcov, double *gcon, double *gdet, double *conn,
                         double *rcen);
/* init_torus (init.c:808) + set_mag (init_mag.c:6): initial state, uploaded
 * to p/ps (interior only; call hamr_boundprim(ctx, 1) to fill ghosts). */
int hamr_init_torus(hamr_ctx *ctx, const hamr_torus_desc *t);

/* GPU_write / GPU_read (GPU_main.cu:309, :2646): one block's state in the
 * reference's host layout (memory.c:27): p[cell][NPR] and ps[cell][NDIM] with
 * B^1..B^3 in components 1..3 (component 0 is not read, and not written back),
 * cells in ghost-inclusive index_3D order.  ps may be NULL. */
int hamr_upload(hamr_ctx *ctx, int n, const double *p_aos, const double *ps_aos);
int hamr_download(hamr_ctx *ctx, int n, double *p_aos, double *ps_aos);
/* Whole arrays in [block][component][cell] order (see HAMR_* selectors). */
int hamr_set_array(hamr_ctx *ctx, int which, const double *src);
int hamr_get_array(hamr_ctx *ctx, int which, double *dst);
/* The same with device pointers (same layout, device memory of the context's
 * device; for callers that keep the state in their own device buffers). */
int hamr_set_array_dev(hamr_ctx *ctx, int which, const double *dsrc);
int hamr_get_array_dev(hamr_ctx *ctx, int which, double *ddst);

/* ---- restart records (the rdump files of restart.c) ----
 * Record of one block: its ghost-inclusive cells in index_3D order (z fastest),
 * NPR+4 doubles per cell: p[0..8], 0 (face-field component 0, never read
 * back), then B^1 gdet(FACE1), B^2 gdet(FACE2), B^3 gdet(FACE3), gdet of the
 * cell's (i, j) column.  This is the content of the file rdump<n>
 * (MPI_File_iwrite of array_rdump, restart.c:64-66, raw doubles, no header). */
/* Doubles of block n's record when its block is red_d times finer along d
 * (red = 1: the block itself; restart.c:181 npr_file and :233 loop bounds). */
int hamr_rdump_block_size(const hamr_ctx *ctx, int n, int red1, int red2, int red3, long long *count);
/* rdump_block_write (restart.c:52-67): block n's record of the current state
 * into rec (host memory). */
int hamr_rdump_block_write(hamr_ctx *ctx, int n, double *rec);
/* rdump_block_read (restart.c:178-269): block n's state from a record whose
 * block is red_d (1..64) times finer along d (BS_read / BS, restart.c:228-232).
 * Sets p and the face field of the cells [NG, NG+BS] of each direction (the
 * interior and the first upper ghost row, restart.c:244) to the means of the
 * fine cells and faces they cover; the other ghosts are left as they were
 * (restart_read bounds them afterwards: hamr_boundprim(ctx, 1)). */
int hamr_rdump_block_read(hamr_ctx *ctx, int n, const double *rec, int red1, int red2, int red3);
/* The records of all blocks back to back in block order, in device memory of
 * the context's device: off[n] (n = 0..nblk, nblk+1 entries) is the first
 * double of block n's record, off[nblk] the total.  The write packs every block
 * in one launch; the read takes records of the same block sizes (red = 1). */
int hamr_rdump_offsets(const hamr_ctx *ctx, long long *off);
int hamr_rdump_write_dev(hamr_ctx *ctx, double *drec);
int hamr_rdump_read_dev(hamr_ctx *ctx, const double *drec);

/* ---- dump records (the new_dump files of dump.c) ----
 * dump_block (dump.c:276-356, ideal gas, no optional physics): a block's
 * interior cells in index_3D order, NPRDUMP = 9 floats each: rho, u, u^0..u^3
 * (ucon_calc at the cell centre, phys.c:309) and the cell-centred B^1..B^3.
 * The file dumps<k>/new_dump<rank> is the records of the rank's blocks in its
 * block order (dump_new, dump.c:4-66). */
int hamr_dump_block_size(const hamr_ctx *ctx, int n, long long *count);
int hamr_dump_block_write(hamr_ctx *ctx, int n, float *rec);
/* All blocks back to back in block order in device memory: off[n] (nblk+1
 * entries) is the first float of block n's record, off[nblk] the total. */
int hamr_dump_offsets(const hamr_ctx *ctx, long long *off);
int hamr_dump_write_dev(hamr_ctx *ctx, float *drec);

/* ---- per-block stages, the call sequence of advance_GPU (step_ch.c:997-1336) ---- */
/* GPU_fluxcalc2D (GPU_main.cu:560): PPM + HLL flux through the lower dir-faces
 * of block n; records the block's Courant dt for hamr_block_dt. */
int hamr_fluxcalc(hamr_ctx *ctx, int dir, int flag, int n);
/* fluxcalc_GPU (GPU_main.cu:1091): the block's minimum cour*dx/ctop of the last
 * fluxcalc in direction dir (1e9 if none); blocking; resets the minimum. */
int hamr_block_dt(hamr_ctx *ctx, int n, int dir, double *out);
/* flux_send1-3 / flux_rec1-3 with calc_corr = 1 (boundcomF.c:5-1315; called
 * between the fluxes and consttransport1, step_ch.c:1142-1196): after every
 * block's fluxes of the full step (flag 1), each face of a block whose
 * neighbour across it is one level finer takes ((0.25 dt) (f00 + f01 + f10 +
 * f11)) / dt of the four finer faces that tile it.  flag 0 (the half step) and
 * single-level sets: nothing. */
int hamr_flux_correct(hamr_ctx *ctx, int flag, double dt);
/* EMF correction at level jumps: E_send1-3, E*_send_corn and E_rec1-3 /
 * E*_receive_corn with calc_corr = 1 inside GPU_consttransport_bound
 * (GPU_main.cu:1144-1330; boundcomE.c:5-2926, GPU_program2.cu:1875-2484).
 * With flag = 1 (the full step) every edge EMF of a block boundary that a
 * one level finer neighbour tiles takes the mean of the two finer edges; with
 * flag = 0 nothing happens.  Call after hamr_ct_bound, before hamr_ct_update.
 * A single-level set has no such edges. */
int hamr_emf_correct(hamr_ctx *ctx, int flag, double dt);
/* GPU_consttransport1 + GPU_consttransport2 (GPU_main.cu:1365, :1393). */
int hamr_ct_emf(hamr_ctx *ctx, int flag, int n);
/* GPU_consttransport_bound (GPU_main.cu:1144): polar E1 average over each ring
 * of phi-blocks; dt is the step's dt (the reference's scale factor). */
int hamr_ct_bound(hamr_ctx *ctx, double dt);
/* GPU_consttransport3 (GPU_main.cu:1535): face-B update with step Dt. */
int hamr_ct_update(hamr_ctx *ctx, int flag, double Dt, int n);
/* GPU_fixup (GPU_main.cu:1740): conserved update + Noble cascade + floors. */
int hamr_update(hamr_ctx *ctx, int flag, int n, double Dt);
/* The inversion half of GPU_fixup alone (GPU_fixup.h:349-425; the cascade
 * of step_ch.c:632-651): the Noble 2D inversion of every interior cell of
 * every block from the conserved variables in HAMR_UNEW, the entropy-based 1D
 * v^2 and the 1D inversions where it fails, the floors, the flags and the
 * failimage counts.  flag 0: guess p, result in ph; flag 1: guess ph, result
 * in p.  For callers that form U themselves, and to drive the failure paths. */
int hamr_invert(hamr_ctx *ctx, int flag);
/* GPU_boundprim (GPU_main.cu:2113): ghost fill of all blocks; flag 0 fills
 * ph/psh, flag 1 fills p/ps. */
int hamr_boundprim(hamr_ctx *ctx, int flag);
/* pflag (ghost-inclusive, per cell) and failimage counts of block n:
 * [0] 2D failures, [1] 1D-v^2 failures, [2] 1D failures, [3] floor hits. */
int hamr_flags(hamr_ctx *ctx, int n, int *pflag, int *failimage);
/* Zero the failimage counters of all blocks. */
int hamr_clear_flags(hamr_ctx *ctx);

/* ---- batched step (GPU_step_ch, step_ch.c:916, in single-timelevel form) ---- */
/* Courant dt of the current state p/ps (per-block harmonic, min over blocks). */
int hamr_estimate_dt(hamr_ctx *ctx, double *dt);
/* nsteps predictor+corrector steps over all blocks with the dt kept on the
 * device (next dt = min over blocks, clamped to 1.3 dt).  *dt_inout: the first
 * step's dt on input (<= 0: keep the device value), the next dt on output. */
int hamr_step(hamr_ctx *ctx, int nsteps, double *dt_inout);
/* Capture one step into a HIP graph and replay it nsteps times (same result as
 * hamr_step). */
int hamr_step_graph(hamr_ctx *ctx, int nsteps, double *dt_inout);
int hamr_sync(hamr_ctx *ctx);

/* ---- local adaptive time stepping (multi-level sets; LAS.c, advance_GPU) ----
 * Replaces the timelevel machinery of GPU_step_ch / advance_GPU
 * (step_ch.c:916-1336, PRESTEP = PRESTEP2 = 0, AMR_SWITCHTIMELEVEL =
 * AMR_MAXTIMELEVEL), set_timelevel (LAS.c:5-236), the flux and EMF
 * corrections across time levels (boundcomF.c / boundcomE.c with calc_corr 1,
 * GPU_boundcomF.cu:5-330, GPU_program2.cu:1168-2704) and the ghost exchange
 * between blocks that step together (unpackreceive1-3, GPU_program2.cu:292-390).
 *
 * Time levels of every block (nblk values, powers of 2 from 1 to 16, equal over
 * each polar ring; a finer block at a level jump may not have a larger level
 * than its coarser neighbour).  Builds the cycle's work lists and correction
 * tables; <0 and a message on a violation.  Multi-level contexts only. */
int hamr_set_timelevels(hamr_ctx *ctx, const int *tl);
/* The current time levels (1 for every block until set). */
int hamr_get_timelevels(const hamr_ctx *ctx, int *tl);
/* Per-block Courant dt (the harmonic 1/(1/dt1 + 1/dt2 + 1/dt3), bdt[][0] of
 * step_ch.c:1323) recorded by the last corrector of hamr_step,
 * hamr_step_graph, hamr_step_las or hamr_estimate_dt. */
int hamr_block_dts(hamr_ctx *ctx, double *bdt);
/* set_timelevel's rule on those block dts and the current dt, then
 * hamr_set_timelevels: block n gets the largest 2^i <= maxtl with bdt[n]/dt >=
 * 1.05 2^i (else 1), each polar ring the minimum of its blocks; a finer block
 * at a level jump is lowered to its coarser neighbour's level (with the ring
 * minimum again until nothing changes).  tl_out (nblk values, may be NULL)
 * receives the result. */
int hamr_timelevels_from_dt(hamr_ctx *ctx, int maxtl, int *tl_out);
/* The same rule on plain arrays (no context): ring is nring x ring_len block
 * ids, pairs npair (coarse, finer) block pairs. */
int hamr_las_rule(int nblk, const double *bdt, double dt, int maxtl, int nring, int ring_len, const int *ring,
                  int npair, const int *pairs, int *tl);
/* ncycles cycles of 2 maxtl sub-iterations (maxtl: a power of 2, at least the
 * largest time level, at most 16), each advancing every block by maxtl dt: a
 * block of level tl takes its predictor (tl dt / 2) at sub-iterations s with
 * s % tl == tl - 1 and its corrector (tl dt) at s % 2tl == 2tl - 1.  dt is kept
 * on the device as in hamr_step and is updated at each cycle's end from the
 * Courant minima of its last sub-iteration. */
int hamr_step_las(hamr_ctx *ctx, int maxtl, int ncycles, double *dt_inout);
/* Capture one cycle into a HIP graph (again when maxtl, the time levels, the
 * arithmetic or the launch shapes change) and replay it ncycles times (same
 * result as hamr_step_las). */
int hamr_step_las_graph(hamr_ctx *ctx, int maxtl, int ncycles, double *dt_inout);
/* info[0..2]: flux-correction entries per direction, [3]: same-level EMF
 * entries, [4]: level-jump EMF entries, [5]: the largest time level (0 until set). */
int hamr_las_info(const hamr_ctx *ctx, long long *info);
/* Average device time (ms, HIP events on the context stream) of one launch of
 * a stage over all blocks, with predictor-step operands (p and ps are not
 * modified): 0-2 flux x1..x3, 3 EMF, 4 polar average, 5 face-B update,
 * 6 conserved update + inversion, 7 ghost fill (3 directions) as the step
 * runs it after the update (the x3 ghost rows that the update writes left
 * out). */
int hamr_time_stage(hamr_ctx *ctx, int stage, int nrep, double *avg_ms);
/* ---- launch autotuner (no reference counterpart) ----
 * Every kernel of the step walks a work list of (block, tile) items.  For each
 * kernel (HAMR_TUNE_*) the tuner times, on the current state, the tile sizes
 * 64/128/256 threads, grids of one workgroup per item or 1024/2048/4096
 * persistent workgroups, and item orders plain or XCD-contiguous, and keeps the
 * fastest (nrep timed launches each).  Results do not depend on the choice.
 * The step graph is re-captured after a change. */
enum {
    HAMR_TUNE_FLUX1 = 0, HAMR_TUNE_FLUX2, HAMR_TUNE_FLUX3, HAMR_TUNE_ECENT, HAMR_TUNE_ECORNER, HAMR_TUNE_CT,
    HAMR_TUNE_CONS, HAMR_TUNE_INVERT, HAMR_TUNE_FALLBACK, HAMR_TUNE_BOUND1, HAMR_TUNE_BOUND2, HAMR_TUNE_BOUND3,
    HAMR_TUNE_EXCHANGE, HAMR_TUNE_N
};
int hamr_autotune(hamr_ctx *ctx, int nrep);
/* cfg[3]: threads per workgroup, workgroups (0: one per item), flags; ms:
 * tuned time (-1: not tuned).  Flags bit 0: XCD-contiguous item order; bit 1,
 * the slot's second form: for HAMR_TUNE_FLUX1-3 line-staged tiles of 256
 * threads (each cell's slope, interface and limited parabola formed once per
 * tile) instead of one thread per face; for HAMR_TUNE_CONS the conserved
 * update and the 2D inversion in two kernels (k_cons, then HAMR_TUNE_INVERT's
 * k_invert) instead of one (k_update).  For HAMR_TUNE_CONS, workgroups = -1
 * selects the two-kernel form with the update's x1-marching k_cons (one thread
 * per (j, z) walking i).  The tuner tries every form; all give the same result
 * bit for bit. */
int hamr_tune_get(const hamr_ctx *ctx, int slot, int *cfg, double *ms);
int hamr_tune_set(hamr_ctx *ctx, int slot, const int *cfg);
/* Average device time (ms, HIP events on the context stream) of one launch of
 * kernel `slot` over all blocks with its tuned shape and predictor-step
 * operands (HAMR_TUNE_EXCHANGE: the three exchange stages of a multi-level set).
 * Like hamr_autotune it keeps p/ps, the flags and the failure counters of the
 * last step, and leaves p/ps and ph/psh with their ghosts filled; the
 * intermediate arrays (fluxes, EMFs, U0, Un, ph, psh) hold the timed launches'
 * results. */
int hamr_time_kernel(hamr_ctx *ctx, int slot, int nrep, double *avg_ms);

/* ---- one forward stage and its adjoint (no reference counterpart) ----
 * Forward stage: the predictor's flux, EMF, face-field and conserved updates
 * with the full step Dt, from p/ps (ghosts filled first):
 *   U' = U(P) - Dt sum_d dF_d/dx_d + Dt S(P) (HAMR_UNEW), U'[B] = the face
 *   average of the updated face field (HAMR_PSH).
 * Adjoint: for L = sum over interior cells of w . U', the gradients dL/dP
 * (HAMR_P layout) and dL/d(face field) (HAMR_PS layout), FP64, recomputing the
 * reconstructed states; the limiter and upwind selections are those of the
 * forward stage.  Supported for a flat periodic box of one block in FP64
 * (HAMR_PREC_FP64, minkowski = 1, nb = 1, 1, 1). */
int hamr_stage_forward(hamr_ctx *ctx, double Dt);
int hamr_stage_adjoint(hamr_ctx *ctx, double Dt, const double *w, double *grad_p, double *grad_ps);
int hamr_stage_adjoint_dev(hamr_ctx *ctx, double Dt, const double *w_dev, double *grad_p_dev, double *grad_ps_dev);

/* hamr_stage_adjoint_dev without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, training or serving loops that forbid hipMalloc).  The
 * adjoint's work arrays (57 doubles per cell of the block's padded slab) live
 * in a file-scope device array of HAMR_NOALLOC_SCRATCH_BYTES bytes that each
 * GPU the library runs on holds until the process ends; a box whose work
 * arrays do not fit is refused (-1, hamr_last_error says so): at 128 MiB,
 * boxes of up to 294,337 ghost-inclusive cells, e.g. 60^3 interior zones.
 * Inputs and outputs as hamr_stage_adjoint_dev (device pointers).  The call
 * makes no hipMalloc/hipFree or other allocating call; it clears the
 * accumulators it needs with a fill kernel and synchronises the context's
 * stream before returning.  All calls on one GPU share that GPU's array, so
 * calls that overlap on one GPU (several contexts on several host threads or
 * streams) race; calls in sequence do not. */
#define HAMR_NOALLOC_SCRATCH_BYTES (128ll << 20)
int hamr_stage_adjoint_dev_noalloc(hamr_ctx *ctx, double Dt, const double *w_dev, double *grad_p_dev,
                                   double *grad_ps_dev);

/* ---- libhamr_host.so: the reference's host-side step in C ----
 * advance_GPU / GPU_step_ch (step_ch.c:916-1336) over the per-block entry
 * points above for blocks 0..nblk-1 (every block at timelevel 1), each block's
 * Courant dt read back to the host as fluxcalc_GPU does; *ndt_out: the next dt
 * (min over blocks of the per-block harmonic, clamped to 1.3 dt). */
int hamr_host_advance(hamr_ctx *ctx, int nblk, double dt, double *ndt_out);

/* Streaming copy of `bytes` bytes (read once, written once), nrep launches:
 * the known-size kernel that calibrates HBM byte counters; average ms. */
int hamr_copy_probe(hamr_ctx *ctx, long long bytes, int nrep, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
