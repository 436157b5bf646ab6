// hamr_amr.h -- multi-level block sets (level ratio 2, uneven block sizes):
// description, layout and ghost-exchange tables built on the host.
#pragma once

#include <cstdint>
#include <string>
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

}  // namespace hamr
