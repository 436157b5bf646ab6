// hamr_amr.h -- multi-level block sets (level ratio 2, uneven block sizes):
// description, layout and ghost-exchange tables built on the host.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "hamr_grid.h"

namespace hamr {

constexpr int kAmrMaxLevels = 3;

struct AmrDesc {
    int nlev;                       // refinement levels (1..3)
    int box[kAmrMaxLevels][3][2];   // level l >= 1 covers [box[l][k][0], box[l][k][1]) cells of level l along k
    uint64_t seed;                  // seeded choice of block sizes
    double p_big, p_mid;            // a superblock stays one 64x32x32 block / splits into 32x16x16 (else 16^3)
};

// One geometry slab: D1 x D2 columns of level `level` from cell (o1, o2), padded to kcol.
struct AmrGeo {
    int level, o1, o2, D1, D2, kcol;
    int64_t moff;
};

struct AmrGrid {
    std::vector<BlockDev> blocks;   // launch order
    std::vector<AmrGeo> geos;
    int64_t cells = 0;              // sum of the blocks' padded slab lengths
    int64_t cols = 0;               // sum of the geometry slabs' padded column counts
    std::vector<AmrEntry> ex[3];    // per stage: copies and restrictions
    int64_t npush3 = 0;             // the last npush3 entries of stage 3: ghost rows that the update writes
    std::vector<int> push3tab;      // push table (BlockDev::push3 rows)
    std::vector<AmrProlong> pro[3]; // per stage: prolongations
    std::vector<AmrReflux> rfx[3];  // per direction: coarse faces at level jumps
    std::vector<AmrEmf> emf;        // coarse edges at level jumps
    std::vector<int> ring, rside;   // polar rings (ring_len blocks each) and their sides
    int ring_len = 0, ni_max = 0;
    int nlev = 0;
    int N[kAmrMaxLevels][3] = {};   // cells of each level
};

// The base grid's coordinates, physics and superblock count come from d
// (d.nb * d.bs = cells of level 0, a multiple of 64x32x32).  Returns "" or an error.
std::string build_amr(const GridDesc &d, const AmrDesc &a, AmrGrid &g);
// Range checks of every table index (run by build_amr).
std::string validate_amr(const AmrGrid &g);

// Local adaptive time stepping on a multi-level set (time level tl[n] of block
// n, powers of 2 up to kMaxTimeLevel).
constexpr int kMaxTimeLevel = 16;
struct LasTables {
    std::vector<LasCorr> flux[3];   // per direction: level jumps (g.rfx) and same-level faces, tl_s < tl_r
    std::vector<LasCorr> emf_same;  // same-level edges taking the fastest neighbour's EMF
    std::vector<LasCorr> emf_jump;  // level-jump edges (g.emf)
};
// The time levels are checked first: powers of 2 up to kMaxTimeLevel, equal
// over each polar ring, and no finer block at a level jump slower than its
// coarser neighbour.  Returns "" or an error.
std::string build_las_tables(const AmrGrid &g, const std::vector<int> &tl, LasTables &t);
// The time-level rule of set_timelevel (LAS.c:5-144, PRESTEP = 0): block n gets
// the largest 2^i <= maxtl with bdt[n] / dt >= 1.05 2^i (else 1), each polar
// ring the minimum of its blocks; then, for the level-jump rule above, a finer
// block at a jump takes at most its coarser neighbour's level (pairs: (coarse,
// finer) block pairs), repeated with the ring minimum until nothing changes.
void las_rule(int nblk, const double *bdt, double dt, int maxtl, const std::vector<int> &ring, int ring_len,
              const std::vector<std::pair<int, int>> &pairs, int *tl);
// (coarse, finer) block pairs of the level-jump corrections of g, sorted and unique.
std::vector<std::pair<int, int>> jump_pairs(const AmrGrid &g);

}  // namespace hamr
