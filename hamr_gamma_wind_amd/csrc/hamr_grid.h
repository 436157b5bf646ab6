// hamr_grid.h -- block table, device layout and host set-up of the step.
//
// Layout in HBM (all blocks resident, one slab per block):
//   cell(i,j,z) = (i*D2 + j)*D3 + z, D = BS + 2*NG, z fastest (the reference's
//   index_3D order); each component of each block is a slab of `ksize` cells
//   (padded to a multiple of 64 doubles = 512 B), components of a block are
//   contiguous:  state[((blk*ncomp) + k)*ksize + cell].
//   Metric: one geometry slab per distinct (c1, c2) block column (the metric is
//   axisymmetric), columns col(i,j) = i*D2 + j padded to `kcol`:
//   gcov/gcon[((geo*NPOS + pos)*10 + s)*kcol + col], gdet[(geo*NPOS + pos)*kcol + col],
//   conn[(geo*40 + lam*10 + s)*kcol + col], rcen[geo*kcol + col].
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hamr {

constexpr int NG = 3;

// Face kinds of the ghost fill.  FK_AMR: the face's ghosts come from the
// multi-level exchange tables (neighbours of other sizes or levels).
enum FaceKind : int { FK_COPY = 0, FK_OUTFLOW = 1, FK_POLE = 2, FK_AMR = 3 };

// Per-block record used by the kernels (plain data, 168 bytes).  Component k of
// an array with ncomp components starts at ncomp*off + k*ks for this block.
struct BlockDev {
    int geo;        // geometry slab
    int c1, c2, c3; // block coordinates (uniform grid: block index; multi-level: origin / 16)
    int pole;       // bit0: lower x2 face is the polar axis, bit1: upper x2 face
    int nbr[6];     // source block for faces -x1,+x1,-x2,+x2,-x3,+x3 (-1: none)
    int kind[6];    // FaceKind of each face
    int level;      // refinement level (0: base)
    int o[3];       // first interior cell, in cells of the block's level
    int bs[3];      // interior cells
    int D[3];       // bs + 2 NG
    int ks;         // cells per component slab (padded to 64)
    int kcol;       // columns per metric slab (padded to 64)
    int tl;         // time level of local adaptive time stepping (a power of 2; 0 until set)
    int push3[2];   // lower / upper x3 face: first entry of the face's row of the push table (-1: none),
                    // | kPushWhole when one entry serves the whole face
    int64_t off;    // cells per component of all earlier blocks' slabs
    int64_t moff;   // metric columns of all earlier geometry slabs
    double dx[3];   // code-coordinate spacing of the block's level
};
static_assert(sizeof(BlockDev) == 168, "BlockDev layout");

// Push table: the update and face-field kernels write the NG lowest and
// highest cells of a block's z-rows also into the x3 ghost cells of the block
// across that x3 face that holds them (same level; the ghost rows inside that
// block's x1 and x2 range), so that the fill right after the update skips
// them.  A face's row of the table (from BlockDev::push3) has one entry per
// kPushTile x kPushTile tile of the block's (i, j) columns, tile (t1, t2) at
// t1 * ceil(bs2 / kPushTile) + t2: the receiving block, or -1 (its ghosts come
// from the exchange).  The receiver's ghost cell of the block's cell (i, j, z)
// is (i + o1 - o1', j + o2 - o2', z - bs3) across the upper face and
// (i + o1 - o1', j + o2 - o2', z + bs3') across the lower one.
constexpr int kPushTile = 16;
// The push table of a uniform grid (every x3 face: the periodic neighbour);
// sets each block's push3 offsets.
std::vector<int> uniform_push_table(std::vector<BlockDev> &t);
// Device form of a push-table entry: the receiver and where its ghost cell is.
// With the receiver's D2 and D3 equal to the sender's the ghost cell is the
// sender's cell index + delta; otherwise delta is kPushGeneral and the index is
// formed from the two blocks' rows.
constexpr int kPushGeneral = -2147483647 - 1;
// BlockDev::push3 flag: every tile of the face has the same entry (the row's
// first), so a launch reads it at one address for the whole face.
constexpr int kPushWhole = 1 << 30;
struct PushEntry {
    int m;        // receiving block (-1: none)
    int delta;
    int ks;       // its component stride
    int pad;
    int64_t off;  // its slab offset
};
static_assert(sizeof(PushEntry) == 24, "PushEntry layout");
// The device table of the receivers' table tab; sets kPushWhole in the blocks' push3.
std::vector<PushEntry> push_entries(std::vector<BlockDev> &blk, const std::vector<int> &tab);

// Multi-level exchange table entries (one per ghost cell).  A source code is
// cell | kind << 28 with kind AS_CELL (that cell) or AS_AVG8 (the mean of the
// 2x2x2 cells whose lowest corner is that cell); AS_NONE marks a missing
// neighbour (no slope along that direction).
enum AmrSrcKind : int { AS_NONE = 0, AS_CELL = 1, AS_AVG8 = 2 };
struct AmrEntry {      // copy or restriction
    int dblk, dcell;   // ghost cell
    int sblk, scode;   // source
};
struct AmrProlong {    // prolongation from the coarse cell (cblk, ccell)
    int dblk, dcell;
    int cblk, ccell;
    int nb[6][2];      // (block, code) of the coarse cell's -x1,+x1,-x2,+x2,-x3,+x3 neighbours
};
// Flux correction at a level jump: the face cf (cell index of the face's upper
// cell) of block cb along one direction takes the mean of the four faces of
// the one level finer block fb that tile it: ff, ff + sb, ff + sa, ff + sa + sb,
// with sa, sb fb's strides of the two transverse directions (the slower first).
struct AmrReflux {
    int cb, cf;
    int fb, ff;
};
// EMF correction at a level jump: the edge EMF component comp (0..2: E1..E3)
// at corner cell cc & 0x0FFFFFFF of block cb (comp = cc >> 28) takes the mean of
// the two edges of the one level finer block fb that tile it: corners fc and fc
// + s, s fb's cell stride along the edge.
struct AmrEmf {
    int cb, cc;
    int fb, fc;
};
// Flux or EMF correction of a local-adaptive-time-stepping cycle: the value at
// cell rcell of block rb (the receiver; for an EMF, rcell & 0x0FFFFFFF with the
// component in bits 28-29, as AmrEmf) is replaced, at the receiver's corrector,
// by the time-weighted sum of the sender sb's values over the sender's
// correctors since the receiver's last one, divided by the receiver's step.
// The sender's value at scell & 0x3FFFFFFF is one face / edge (a same-level
// neighbour of a smaller time level), or with bit 30 (LAS_MEAN) the mean of the
// four faces / two edges of a one level finer block that tile the receiver's
// (the strides of AmrReflux / AmrEmf).  acc: one accumulator per entry and
// component.
constexpr int LAS_MEAN = 1 << 30;
struct LasCorr {
    int rb, rcell;
    int sb, scell;
};
static_assert(sizeof(AmrEntry) == 16 && sizeof(AmrProlong) == 64 && sizeof(AmrReflux) == 16 &&
                  sizeof(AmrEmf) == 16 && sizeof(LasCorr) == 16,
              "table entry layout");

// Static description of a run (mirrors the reference's compile-time config).
struct GridDesc {
    int bs[3];      // block interior size
    int nb[3];      // blocks per direction
    double gam;     // adiabatic index
    double a;       // black-hole spin
    double cour;    // Courant factor
    double rin;     // inner radius of the grid (<= 0: 0.9 (1 + sqrt(1-a^2)))
    double rout;    // outer radius
    double rho_floor, u_floor;  // floors rho_floor r^-1.5, u_floor r^-2.5
    int minkowski;  // 1: flat periodic box (unit cube per block grid), else Kerr-Schild torus grid
};

struct Layout {
    int D[3];
    int64_t ncell;  // D1*D2*D3
    int64_t ksize;  // padded slab length (cells)
    int64_t ncol;   // D1*D2
    int64_t kcol;   // padded metric slab length (columns)
    int nblk, ngeo;
};

Layout make_layout(const GridDesc &d);

// Block table: natural order n = (c1*NB2 + c2)*NB3 + c3 on the base level.
// Neighbours: x1 faces outflow at the radial ends, x2 faces at the axis use the
// pi-shifted partner block (c3 + NB3/2) with the polar mirror, x3 periodic.
std::vector<BlockDev> build_block_table(const GridDesc &d, int *ngeo_out);
// The reference's block numbering (AMR_set_coord's level-0 Z-order,
// AMR.c:188-275): element k is the natural index of the block the reference
// numbers k; also the launch order.  Empty if the numbering is inconsistent.
std::vector<int> zorder(const GridDesc &d);

// Host metric generation for one geometry slab (block column c1, c2).
// Writes into host arrays laid out as one slab of the device layout.
struct GridCoords {
    double startx[3], dx[3];
    double rin, rout, fractheta;
};
GridCoords grid_coords(const GridDesc &d);
void bl_coord(const GridDesc &d, const double *X, double *r, double *th);
void metric_code(const GridDesc &d, const double *X, double gcov[4][4]);
double gdet_of(const double g[4][4]);
int invert4(const double a[4][4], double inv[4][4]);
void conn_code(const GridDesc &d, const double *X, const double gcon[4][4], double conn[4][4][4]);
// Code coordinates of position `pos` of global cell (i, j) (interior-relative,
// may be a ghost index), following the reference's coord() mirroring.
void coord_of(const GridDesc &d, const GridCoords &gc, int i, int j, int pos, double X[4]);
// The same on refinement level `level` (spacing halved per level; i, j in the
// level's cells).
void coord_of_level(const GridDesc &d, const GridCoords &gc, int level, int i, int j, int pos, double X[4]);
void fill_geometry(const GridDesc &d, const Layout &L, int c1, int c2, double *gcov, double *gcon,
                   double *gdet, double *conn, double *rcen);
// One geometry slab of D1 x D2 columns (padded to kcol) whose first interior
// column is cell (o1, o2) of level `level`.
void fill_geometry_level(const GridDesc &d, int level, int o1, int o2, int D1, int D2, int64_t kcol, double *gcov,
                         double *gcon, double *gdet, double *conn, double *rcen);

// Orthonormal tetrad at a face of direction dir (1..3) from the packed gcov /
// gcon there (pairs 00,01,02,03,11,12,13,22,23,33), set_Mud (hllc.c:20-178):
// mud[mu*4 + a] = e_(a)^mu (e_(0) the normal observer, e_(dir) in the span of
// the normal and grad x^dir), mudi[a*4 + mu] = e^(a)_mu, its inverse.
void face_tetrad(const double *gcov, const double *gcon, int dir, double *mud, double *mudi);
// The tetrads of one geometry slab in the device layout tet[(t*32 + e)*kcol +
// col] (t = dir - 1, e < 16: mud, e >= 16: mudi) from the slab's packed metric
// gcov/gcon[(pos*10 + s)*kcol + col]; columns >= ncol (padding) are zero.
void fill_tetrads(const double *gcov, const double *gcon, int64_t ncol, int64_t kcol, double *tet);

// Torus initial state on the whole grid, written per block in the device layout
// (p: NPR components, ps: 3).  Returns an empty string or an error message.
struct TorusDesc {
    double rin, rmax, kappa;   // Fishbone-Moncrief torus (r_in, r at pressure max, K)
    double beta_min;           // field normalisation: min over cells of p_gas / p_mag
    double aphi_cut;           // A_phi ~ max(rho/rho_max - aphi_cut, 0)
    double perturb;            // u *= 1 + perturb (U[0,1) - 1/2)
    uint64_t seed;
};
// Works for any block table (uniform or multi-level): per-block sizes, levels,
// slab and metric offsets come from the rows.
std::string init_torus(const GridDesc &d, const std::vector<BlockDev> &blocks, const TorusDesc &t,
                       const double *gcov, const double *gcon, const double *gdet, double *p, double *ps);

uint64_t hash64(uint64_t x);

}  // namespace hamr
