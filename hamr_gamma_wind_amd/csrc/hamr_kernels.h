// hamr_kernels.h -- kernel parameters and launchers of the gfx950 step.
#pragma once

#include <hip/hip_runtime.h>

#include "hamr_grid.h"
#include "hamr_phys.h"

namespace hamr {

// Passed by value to every kernel.  Everything that depends on the block (its
// size, slab offsets, metric slab, spacing) is read from the block table.
struct KParams {
    double gam, cour;
    double rho_floor, u_floor;
    const BlockDev *blk;
    const double *gcov, *gcon, *gdet, *conn, *rcen;
    const double *tet;  // face tetrads: tet[moff*96 + ((dir-1)*32 + e)*kcol + col] (e < 16: e_(a)^mu, else e^(a)_mu)
    // The metric rounded to FP32 for the mixed step's kernels, with 1 - alpha
    // formed in FP64 first: g32[moff*NPOS*G32N + (pos*G32N + s)*kcol + col],
    // s = 0-9 gcov, 10-19 gcon, 20 gdet, 21 1 - 1/sqrt(-g^00).
    const float *g32;
    int riemann;        // HAMR_RIEMANN_HLL (0) or HAMR_RIEMANN_HLLC (1), in either arithmetic
    const PushEntry *push3;  // push table (rows from BlockDev::push3)
    int tiled_push;          // 1: a multi-level set (per-tile push entries); 0: every face one entry
    // Local adaptive time stepping: the sub-iteration s of a cycle (-1 outside
    // one) and the phase a launch serves (LAS_HALF or LAS_FULL).  In a cycle a
    // kernel scales its step by the block's time level (las_phase).
    int las_s;
    int las_want;
    // The state arrays (p, ph and their face fields): in a cycle the update
    // writes a receiver's pushed ghosts into both pairs, as the exchange does.
    double *xp, *xph, *xps, *xpsh;
};

// Phase of block B in sub-iteration s of a cycle (advance_GPU, step_ch.c:1003-1010,
// with PRESTEP = PRESTEP2 = 0): a block of time level tl takes its predictor
// when s % tl == tl - 1 and its corrector when s % 2tl == 2tl - 1.
enum { LAS_IDLE = 0, LAS_HALF = 1, LAS_FULL = 2 };
__host__ __device__ inline int las_phase(int tl, int s)
{
    if (tl < 1) tl = 1;
    if (s % (2 * tl) == 2 * tl - 1) return LAS_FULL;
    if (s % tl == tl - 1) return LAS_HALF;
    return LAS_IDLE;
}

// Work list of one launch: `items` are (block, tile) pairs, a tile being
// blockDim.x consecutive threads of the block's iteration space.  A grid of
// any size walks the list with a stride of gridDim.x (a persistent launch when
// the grid is smaller than the list).  With swz = 1 the item order is permuted
// so that the workgroups one XCD runs (ids congruent mod 8) take one contiguous
// range of the list: the tiles of a block, whose stencils share halo data, then
// meet in the same L2.
struct WorkList {
    const int2 *items;
    int n;
    int swz;
};

// One launch configuration: the work list built for `threads` per tile, the
// workgroup count (<= wl.n) and the LDS the update kernel stages per tile.
struct Launch {
    WorkList wl;
    int threads;
    int grid;
    int lds_cols;  // LDS a tile stages: metric columns (k_cons), or values per staged array (line-staged flux)
    int line = 0;  // flux launches: 1 = line-staged tiles (KC_LINE lists), 0 = one thread per face (KC_FLUX lists)
};

constexpr int G32N = 22;  // FP32 metric values per column and position (KParams::g32)

// Tiles of the line-staged flux kernel (kLineT threads per workgroup) of
// direction dir (1..3) for a block of b1 x b2 x b3 cells.  A line (one
// transverse position of the face range, both transverse directions over
// [NG-1, NG+BS]) has b_dir + 1 faces, cut into nseg segments of S faces (the
// last one possibly shorter), at most kLineSeg each.  The lines form n1 rows
// (the slower transverse direction) of n2 lines; a tile is one segment on L =
// kLineT / (S + 3) consecutive lines of one row (tpr tiles per row), one
// thread per cell of the S + 3 cells whose slopes the segment's faces need.
constexpr int kLineT = 256;
constexpr int kLineSeg = 33;
struct LineTiles {
    int S, nseg, L, n1, n2, tpr, ntiles;
};
__host__ __device__ inline LineTiles line_tiles(int b1, int b2, int b3, int dir)
{
    LineTiles x;
    const int nf = (dir == 1 ? b1 : dir == 2 ? b2 : b3) + 1;
    x.nseg = (nf + kLineSeg - 1) / kLineSeg;
    x.S = (nf + x.nseg - 1) / x.nseg;
    x.L = kLineT / (x.S + 3);
    x.n1 = (dir == 1 ? b2 : b1) + 2;
    x.n2 = (dir == 3 ? b2 : b3) + 2;
    x.tpr = (x.n2 + x.L - 1) / x.L;
    x.ntiles = x.nseg * x.n1 * x.tpr;
    return x;
}

// Kernel classes: each has its own iteration space per block.
enum KClass : int {
    KC_FLUX1 = 0, KC_FLUX2, KC_FLUX3,  // faces of one direction (+ one transverse ghost layer)
    KC_ECENT,                          // cells [NG-1, NG+BS]
    KC_CORNER,                         // corners/faces [NG, NG+BS]   (corner EMF, face-B update)
    KC_INTERIOR,                       // interior cells               (update, inversion, fallback)
    KC_BOUND1, KC_BOUND2, KC_BOUND3,   // ghost layers of one direction
    KC_LINE1, KC_LINE2, KC_LINE3,      // line-staged flux tiles of one direction (line_tiles)
    KC_N
};
// Threads a block needs in a kernel class.
int64_t class_work(const BlockDev &B, int cls);
// Metric columns a tile of `threads` interior cells of a block with bs3 cells
// along z can touch.
inline int cons_max_cols(int bs3, int threads) { return (threads - 1) / bs3 + 2; }

// `mixed` selects the arithmetic of reconstruction, fluxes and EMFs: 0 = FP64
// (F, Ec, E are double arrays), 1 = FP32 (float arrays with the same slab layout).
void launch_flux(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, void *F,
                 unsigned long long *dtmin, int mixed, hipStream_t s);
void launch_emf(const KParams &P, const Launch &centre, const Launch &corner, const double *p, const void *F1,
                const void *F2, const void *F3, void *Ec, void *E, int mixed, hipStream_t s);
// Its two kernels one at a time (for timing and tuning).
void launch_emf_centre(const KParams &P, const Launch &L, const double *p, void *Ec, int mixed, hipStream_t s);
void launch_emf_corner(const KParams &P, const Launch &L, const void *Ec, const void *F1, const void *F2,
                       const void *F3, void *E, int mixed, hipStream_t s);
// Polar E1 ring average.  ring[r*len + b] lists the blocks of ring r in phi
// order, rside[r] its side (0: lower x2 face is the axis, 1: upper); ni_max is
// the largest bs1 + 1 over the rings.
void launch_pole(const KParams &P, const int *ring, const int *rside, int nring, int len, int ni_max, double dt_val,
                 const double *dt_ptr, void *E, int mixed, hipStream_t s);
void launch_ct_update(const KParams &P, const Launch &L, double Dt_val, const double *dt_ptr, double dt_mult,
                      const double *psi, double *psf, const void *E, int mixed, hipStream_t s);
void launch_update(const KParams &P, const Launch &L, int full, double Dt_val, const double *dt_ptr, double dt_mult,
                   const double *pi, const double *pb, double *pf, double *U0, double *Unew, const double *psf,
                   const void *F1, const void *F2, const void *F3, int *pflag, int *failcnt, int mixed,
                   hipStream_t s);
// Conserved update only (k_cons, predictor form with step Dt): U(p) -> U0,
// U(p) - Dt div F + Dt S(p) with the face average of psf -> Unew.
void launch_cons(const KParams &P, const Launch &L, double Dt, const double *p, double *U0, double *Unew,
                 const double *psf, const void *F1, const void *F2, const void *F3, int mixed, hipStream_t s);
// k_cons marching along x1 (one thread per (j, z) of each block's plane; grid
// y = blocks, at most 65535): the same cells and arithmetic as part 0 below.
// max_plane / min_bs3: the largest bs2 bs3 and the smallest bs3 of the blocks.
void launch_cons_march(const KParams &P, int nblk, int max_plane, int min_bs3, int full, double Dt_val,
                       const double *dt_ptr, double dt_mult, const double *pi, const double *pb, double *U0,
                       double *Unew, const double *psf, const void *F1, const void *F2, const void *F3, int mixed,
                       hipStream_t s);
// The kernels of the update one at a time: part 0 k_cons, 1 k_invert, 2
// k_fallback, 3 k_update (k_cons and k_invert in one kernel).  launch_update
// runs parts 3 and 2.
void launch_update_part(const KParams &P, int part, const Launch &L, int full, double Dt_val, const double *dt_ptr,
                        double dt_mult, const double *pi, const double *pb, double *pf, double *U0, double *Unew,
                        const double *psf, const void *F1, const void *F2, const void *F3, int *pflag, int *failcnt,
                        int mixed, hipStream_t s);
// Adjoint of one forward stage of a single periodic flat block (FP64): from the
// weights w of U' (9 slabs) to lP (9 slabs, dL/dP) and lps (3 slabs, dL/d face
// field); F1-F3 are the stage's fluxes; lpsf, lE, lEc, lF1-lF3 are scratch.
void launch_adjoint(const KParams &P, const BlockDev &B, double Dt, const double *p, const double *ps,
                    const double *F1, const double *F2, const double *F3, const double *w, double *lP, double *lps,
                    double *lpsf, double *lE, double *lEc, double *lF1, double *lF2, double *lF3, hipStream_t s);

// Ghost fill of one direction over the faces whose kind is FK_COPY, FK_OUTFLOW
// or FK_POLE (FK_AMR faces are left to the exchange tables).  With
// outflow_only = 1 only radial outflow faces are filled, and not at the ghost
// rows of a polar face (the final pass of the multi-level fill).
// With x3_rows_done = 1 the x3 fill skips the rows inside the block's x1 and
// x2 range whose ghosts the update (k_invert, k_fallback) and face-field
// (k_ct_update) kernels of the step wrote from the periodic neighbour.
void launch_bound(const KParams &P, int dir, const Launch &L, double *p, double *ps, int *pflag, int outflow_only,
                  hipStream_t s, int x3_rows_done = 0);
// Flux correction at level jumps along direction dir (1..3): each entry's
// coarse face takes the mean of its four finer faces (F: the direction's flux
// array, FP32 when mixed); dt from dt_ptr if non-null.
void launch_amr_reflux(const KParams &P, const AmrReflux *e, int64_t n, int dir, void *F, double dt_val,
                       const double *dt_ptr, int mixed, hipStream_t s);
// EMF correction at level jumps: each entry's coarse edge EMF takes the mean
// of its two finer edges (E: the corner EMF array, FP32 when mixed).
void launch_amr_emf(const KParams &P, const AmrEmf *e, int64_t n, void *E, double dt_val, const double *dt_ptr,
                    int mixed, hipStream_t s);
// Restart records (rdump files): RREC doubles per ghost-inclusive cell.  Item
// q of a launch is block blist[q], whose record starts at out + roff[q] -
// roff_host[first]; max_cells bounds the items' cell counts (pack: D1 D2 D3;
// unpack: (bs1+1)(bs2+1)(bs3+1)).
constexpr int RREC = NPR + 4;
struct RdumpRead {
    int red[3];            // the record's block is red[d] times finer along direction d
    int spherical;         // B^2 scaled by fractheta / fractheta_old
    double fractheta, fractheta_old;
};
void launch_rdump_pack(const KParams &P, const int *blist, const int64_t *roff, const int64_t *roff_host, int first,
                       int count, int64_t max_cells, const double *p, const double *ps, double *out, hipStream_t s);
// max_e3 / max_rows: the largest bs3 + 1 and (bs1 + 1)(bs2 + 1) of the items
// (records of the blocks' own size take a staged kernel when max_e3 <= 256).
void launch_rdump_unpack(const KParams &P, const int *blist, const int64_t *roff, const int64_t *roff_host, int first,
                         int count, int64_t max_cells, const RdumpRead &R, const double *rec, double *p, double *ps,
                         int max_e3, int64_t max_rows, hipStream_t s);
// Dump records (new_dump files): DREC floats per interior cell; item q's
// record starts at out + doff[q] - doff_host[first]; max_cells bounds the items'
// interior cell counts.
constexpr int DREC = 9;
void launch_dump_pack(const KParams &P, const int *blist, const int64_t *doff, const int64_t *doff_host, int first,
                      int count, int64_t max_cells, const double *p, float *out, hipStream_t s);
// The FP32 metric (KParams::g32) of ncol columns of one geometry slab from its
// FP64 metric (the slab's pointers into gcov, gcon, gdet and g32).
void launch_geom32(const double *gcov, const double *gcon, const double *gdet, float *g32, int64_t kcol,
                   hipStream_t s);
// dt of the next step from the Courant minima (bdt, when non-null, receives
// each block's harmonic 1/(1/dt1 + 1/dt2 + 1/dt3)); the minima are reset.
void launch_dt_finalize(int nblk, unsigned long long *dtmin, double *dt, hipStream_t s, double *bdt = nullptr);
// Corrections of a local-adaptive-time-stepping cycle at sub-iteration P.las_s
// (LasCorr entries of one direction of fluxes, dir 1..3, or of EMFs, dir 0):
// stage 0 adds every sender's values at its corrector into the accumulators
// (restarting them at the first corrector after the receiver's last), stage 1
// writes acc / (dt tl) into every receiver at its corrector.
void launch_las_corr(const KParams &P, const LasCorr *e, int64_t n, int dir, int stage, void *FE, double *acc,
                     const double *dt_ptr, int mixed, hipStream_t s);
void launch_dt_reset(int nblk, unsigned long long *dtmin, hipStream_t s);
void launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t s);

// Multi-level ghost exchange of stage dir (1..3): copy/restriction entries,
// then prolongation entries (they write disjoint ghost cells and read only
// interior cells and ghosts of earlier stages).  grid <= 0: one thread per entry.
// In a local-time-stepping cycle (P.las_s >= 0) ph/psh are the half-step
// arrays: ghosts are exchanged between blocks that both step in the
// sub-iteration, read from each source's arrays of its phase and written into
// both array pairs of the ghost's block.  Outside a cycle ph = psh = nullptr.
void launch_amr_exchange(const KParams &P, int dir, const AmrEntry *e, int64_t n, const AmrProlong *pe, int64_t npro,
                         double *p, double *ps, int *pflag, int threads, int grid, hipStream_t s,
                         double *ph = nullptr, double *psh = nullptr);

}  // namespace hamr
