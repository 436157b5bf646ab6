// hamr_amr.cpp -- multi-level block sets of uneven block sizes: the block
// table, slab and metric layout, pole rings and the ghost-exchange tables.
//
// Levels have ratio 2 in every direction.  Level l covers the box box[l]
// (level 0: the whole grid); the leaves of level l are box[l] minus
// box[l+1].  Each level is tiled by 64x32x32 superblocks; a superblock (or the
// part of it that is a leaf of its level) is split into blocks of 64x32x32,
// 32x16x16 or 16^3 cells by a seeded choice.  The blocks of one polar ring (same
// level, radial range and size, all phi) get the same choice, so that the
// ring's polar E1 average and the pi-shifted partner exist as in the uniform
// grid.
//
// Ghost exchange (this library's own specification of the multi-level ghost fill;
// the reference exchanges through boundcomP.c / GPU_program2.cu packs):
//   stage d = 1, 2, 3 fills the ghosts along d over the extent already filled
//   along the earlier directions and the interior along the later ones (the
//   order of bound_prim, bounds.c:11-105).  A ghost cell at global position Y
//   of its block's level l:
//   - along d beyond the radial boundary or the polar axis: the outflow and
//     polar-mirror kernels (not in the tables);
//   - in a polar ghost row (x2 beyond the axis, stage 3): a copy of the
//     same-level leaf that holds the row's clamped position, at Y itself (that
//     leaf's mirror ghost);
//   - beyond the block along an earlier direction (a corner ghost) when the leaf
//     across the face that holds the position clamped into the block's range
//     is on level l: a copy of that leaf's storage cell at Y (its ghost of the
//     earlier stage), so that face fields there are the neighbour's own, as the
//     reference's face exchanges carry the neighbour's ghost rows;
//   - otherwise beyond the radial boundary along an earlier direction: left to
//     the final outflow pass, which applies the outflow rule to the block's own
//     edge cell;
//   - otherwise (x3 periodic wrap): the leaf L containing Y gives a copy
//     (level l), the mean of the 2x2x2 children (level l+1), or the
//     MC-limited linear prolongation of its cell (level l-1), whose neighbour
//     values are the cells of that coarse level, the means of finer children,
//     or the cell of a coarser leaf; a neighbour beyond the radial boundary or
//     the axis gives no slope along that direction.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "hamr_amr.h"
#include "hamr_phys.h"

namespace hamr {

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

static uint64_t spread3(uint64_t v)
{
    uint64_t r = 0;
    for (int b = 0; b < 21; b++) r |= ((v >> b) & 1ULL) << (3 * b);
    return r;
}

namespace {

constexpr int UNIT = 16;                   // smallest block edge
constexpr int SB[3] = {64, 32, 32};        // superblock

struct Level {
    int N[3];                              // cells of the level
    int U[3];                              // units (16^3) of the level
    std::vector<int> own;                  // unit -> leaf block id, or kCoarse / kRefined
    int uid(int a, int b, int c) const { return (a * U[1] + b) * U[2] + c; }
};
constexpr int kCoarse = -1, kRefined = -2;

struct Builder {
    const GridDesc &d;
    const AmrDesc &a;
    std::vector<Level> lev;
    std::vector<BlockDev> blk;
    std::string err;

    Builder(const GridDesc &d_, const AmrDesc &a_) : d(d_), a(a_) {}

    bool in_box(int l, int u1, int u2, int u3) const
    {
        // unit (u1, u2, u3) of level l inside box[l] (l = 0: the whole grid)
        if (l == 0) return true;
        const int u[3] = {u1, u2, u3};
        for (int k = 0; k < 3; k++)
            if (u[k] * UNIT < a.box[l][k][0] || u[k] * UNIT >= a.box[l][k][1]) return false;
        return true;
    }
    bool leaf_unit(int l, int u1, int u2, int u3) const
    {
        return in_box(l, u1, u2, u3) && !(l + 1 < a.nlev && in_box(l + 1, 2 * u1, 2 * u2, 2 * u3));
    }
    uint64_t choice(int l, int s1, int s2, int s3, int extra) const
    {
        return hash64(a.seed ^ hash64(((((uint64_t)l * 1024 + s1) * 1024 + s2) * 1024 + s3) * 64 + extra));
    }
    int pick(uint64_t h) const
    {
        const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
        return u < a.p_big ? 0 : (u < a.p_big + a.p_mid ? 1 : 2);
    }
    void add(int l, int o1, int o2, int o3, int b1, int b2, int b3)
    {
        BlockDev B;
        std::memset(&B, 0, sizeof B);
        B.level = l;
        B.o[0] = o1; B.o[1] = o2; B.o[2] = o3;
        B.bs[0] = b1; B.bs[1] = b2; B.bs[2] = b3;
        B.c1 = o1 / UNIT; B.c2 = o2 / UNIT; B.c3 = o3 / UNIT;
        blk.push_back(B);
    }

    bool build_levels()
    {
        const int N0[3] = {d.nb[0] * d.bs[0], d.nb[1] * d.bs[1], d.nb[2] * d.bs[2]};
        if (a.nlev < 1 || a.nlev > kAmrMaxLevels) { err = "amr: 1 to 3 levels"; return false; }
        for (int k = 0; k < 3; k++)
            if (N0[k] % SB[k]) { err = "amr: base grid must be a whole number of 64x32x32 superblocks"; return false; }
        lev.resize(a.nlev);
        for (int l = 0; l < a.nlev; l++) {
            Level &L = lev[l];
            for (int k = 0; k < 3; k++) {
                L.N[k] = N0[k] << l;
                L.U[k] = L.N[k] / UNIT;
            }
            L.own.assign((size_t)L.U[0] * L.U[1] * L.U[2], kCoarse);
            if (l >= 1)
                for (int k = 0; k < 3; k++) {
                    const int lo = a.box[l][k][0], hi = a.box[l][k][1];
                    // a refined box is a whole number of units of its own level and of the level above
                    if (lo < 0 || hi > L.N[k] || lo >= hi || lo % (2 * UNIT) || hi % (2 * UNIT)) {
                        err = "amr: refined boxes must be aligned to 32 cells of their level and inside the grid";
                        return false;
                    }
                }
            if (l >= 2)
                for (int k = 0; k < 3; k++)
                    if (a.box[l][k][0] < 2 * a.box[l - 1][k][0] || a.box[l][k][1] > 2 * a.box[l - 1][k][1]) {
                        err = "amr: refined boxes must nest";
                        return false;
                    }
        }
        return true;
    }

    void tile()
    {
        for (int l = 0; l < a.nlev; l++) {
            const Level &L = lev[l];
            const int ns[3] = {L.N[0] / SB[0], L.N[1] / SB[1], L.N[2] / SB[2]};
            for (int s1 = 0; s1 < ns[0]; s1++)
                for (int s2 = 0; s2 < ns[1]; s2++)
                    for (int s3 = 0; s3 < ns[2]; s3++) {
                        const int u0[3] = {s1 * 4, s2 * 2, s3 * 2};
                        int cnt = 0;
                        for (int x = 0; x < 4; x++)
                            for (int y = 0; y < 2; y++)
                                for (int z = 0; z < 2; z++) cnt += leaf_unit(l, u0[0] + x, u0[1] + y, u0[2] + z);
                        if (!cnt) continue;
                        // superblocks on the polar axis choose per ring (no phi index)
                        const bool polar = (s2 == 0 || s2 == ns[1] - 1) && !d.minkowski;
                        const int c3 = polar ? 0 : s3;
                        if (cnt == 16) {
                            const int k = pick(choice(l, s1, s2, c3, 0));
                            if (k == 0 && !(polar && ns[2] == 1)) {
                                add(l, s1 * 64, s2 * 32, s3 * 32, 64, 32, 32);
                                continue;
                            }
                            if (k <= 1) {
                                for (int x = 0; x < 2; x++)
                                    for (int y = 0; y < 2; y++)
                                        for (int z = 0; z < 2; z++)
                                            add(l, s1 * 64 + 32 * x, s2 * 32 + 16 * y, s3 * 32 + 16 * z, 32, 16, 16);
                                continue;
                            }
                        }
                        // partial superblock (or split to 16^3): pairs of units along x1 where both are leaves
                        for (int x = 0; x < 4; x += 2)
                            for (int y = 0; y < 2; y++)
                                for (int z = 0; z < 2; z++) {
                                    const int a1 = u0[0] + x, a2 = u0[1] + y, a3 = u0[2] + z;
                                    const bool e0 = leaf_unit(l, a1, a2, a3), e1 = leaf_unit(l, a1 + 1, a2, a3);
                                    const bool pair = e0 && e1 && cnt < 16 &&
                                                      pick(choice(l, s1, s2, c3, 1 + x * 4 + y * 2 + (polar ? 0 : z))) <= 1;
                                    if (pair) add(l, a1 * UNIT, a2 * UNIT, a3 * UNIT, 32, 16, 16);
                                    else {
                                        if (e0) add(l, a1 * UNIT, a2 * UNIT, a3 * UNIT, 16, 16, 16);
                                        if (e1) add(l, (a1 + 1) * UNIT, a2 * UNIT, a3 * UNIT, 16, 16, 16);
                                    }
                                }
                    }
        }
    }

    void order_and_own()
    {
        // launch order: level, then Morton order of the origin's units
        std::stable_sort(blk.begin(), blk.end(), [](const BlockDev &x, const BlockDev &y) {
            if (x.level != y.level) return x.level < y.level;
            const uint64_t kx = (spread3(x.o[0] / UNIT) << 2) | (spread3(x.o[1] / UNIT) << 1) | spread3(x.o[2] / UNIT);
            const uint64_t ky = (spread3(y.o[0] / UNIT) << 2) | (spread3(y.o[1] / UNIT) << 1) | spread3(y.o[2] / UNIT);
            return kx < ky;
        });
        for (int l = 0; l < a.nlev; l++) {
            Level &L = lev[l];
            for (int u1 = 0; u1 < L.U[0]; u1++)
                for (int u2 = 0; u2 < L.U[1]; u2++)
                    for (int u3 = 0; u3 < L.U[2]; u3++)
                        if (!leaf_unit(l, u1, u2, u3) && in_box(l, u1, u2, u3)) L.own[L.uid(u1, u2, u3)] = kRefined;
        }
        for (int n = 0; n < (int)blk.size(); n++) {
            const BlockDev &B = blk[n];
            Level &L = lev[B.level];
            for (int u1 = B.o[0] / UNIT; u1 < (B.o[0] + B.bs[0]) / UNIT; u1++)
                for (int u2 = B.o[1] / UNIT; u2 < (B.o[1] + B.bs[1]) / UNIT; u2++)
                    for (int u3 = B.o[2] / UNIT; u3 < (B.o[2] + B.bs[2]) / UNIT; u3++) L.own[L.uid(u1, u2, u3)] = n;
        }
    }

    // leaf holding cell Y of level l (Y inside the grid); returns block, sets its level
    int find(int l, const int Y[3], int *lc)
    {
        const Level &L = lev[l];
        const int o = L.own[L.uid(Y[0] / UNIT, Y[1] / UNIT, Y[2] / UNIT)];
        if (o >= 0) { *lc = l; return o; }
        if (o == kRefined) {
            if (l + 1 >= a.nlev) return -1;
            const Level &F = lev[l + 1];
            const int f = F.own[F.uid(2 * Y[0] / UNIT, 2 * Y[1] / UNIT, 2 * Y[2] / UNIT)];
            *lc = l + 1;
            return f;  // -1/-2: more than one level apart (checked by the caller)
        }
        if (l == 0) return -1;
        const Level &C = lev[l - 1];
        const int c = C.own[C.uid((Y[0] >> 1) / UNIT, (Y[1] >> 1) / UNIT, (Y[2] >> 1) / UNIT)];
        *lc = l - 1;
        if (c < 0 && l >= 2) {
            const Level &C2 = lev[l - 2];
            *lc = l - 2;
            return C2.own[C2.uid((Y[0] >> 2) / UNIT, (Y[1] >> 2) / UNIT, (Y[2] >> 2) / UNIT)];
        }
        return c;
    }
};

int cell_in(const BlockDev &B, const int Y[3])
{
    const int i = Y[0] - B.o[0] + NG, j = Y[1] - B.o[1] + NG, z = Y[2] - B.o[2] + NG;
    return (i * B.D[1] + j) * B.D[2] + z;
}

bool interior_of(const BlockDev &B, const int Y[3])
{
    for (int k = 0; k < 3; k++)
        if (Y[k] < B.o[k] || Y[k] >= B.o[k] + B.bs[k]) return false;
    return true;
}

bool storage_of(const BlockDev &B, const int Y[3])
{
    for (int k = 0; k < 3; k++)
        if (Y[k] < B.o[k] - NG || Y[k] >= B.o[k] + B.bs[k] + NG) return false;
    return true;
}

}  // namespace

std::string build_amr(const GridDesc &d, const AmrDesc &a, AmrGrid &g)
{
    if (d.minkowski) return "amr: the multi-level grid is the torus grid";
    Builder b(d, a);
    if (!b.build_levels()) return b.err;
    b.tile();
    if (b.blk.empty()) return "amr: no blocks";
    b.order_and_own();
    const GridCoords gc = grid_coords(d);
    const int nblk = (int)b.blk.size();
    // layout: slabs in launch order, one geometry slab per distinct (level, o1, o2, bs1, bs2)
    std::map<std::tuple<int, int, int, int, int>, int> geo;
    g.geos.clear();
    int64_t off = 0, moff = 0;
    for (int n = 0; n < nblk; n++) {
        BlockDev &B = b.blk[n];
        for (int k = 0; k < 3; k++) {
            B.D[k] = B.bs[k] + 2 * NG;
            B.dx[k] = gc.dx[k] * std::ldexp(1.0, -B.level);
        }
        const int64_t ncell = (int64_t)B.D[0] * B.D[1] * B.D[2];
        B.ks = (int)round_up(ncell, 64);
        B.kcol = (int)round_up((int64_t)B.D[0] * B.D[1], 64);
        B.off = off;
        off += B.ks;
        auto key = std::make_tuple(B.level, B.o[0], B.o[1], B.bs[0], B.bs[1]);
        auto it = geo.find(key);
        if (it == geo.end()) {
            AmrGeo G{B.level, B.o[0], B.o[1], B.D[0], B.D[1], B.kcol, moff};
            it = geo.emplace(key, (int)g.geos.size()).first;
            g.geos.push_back(G);
            moff += B.kcol;
        }
        B.geo = it->second;
        B.moff = g.geos[B.geo].moff;
    }
    g.cells = off;
    g.cols = moff;
    // faces: physical boundaries and pole partners; every other face is exchanged
    std::map<std::tuple<int, int, int, int, int, int, int>, int> where;  // (level, o, bs) -> block
    for (int n = 0; n < nblk; n++) {
        const BlockDev &B = b.blk[n];
        where[std::make_tuple(B.level, B.o[0], B.o[1], B.o[2], B.bs[0], B.bs[1], B.bs[2])] = n;
    }
    for (int n = 0; n < nblk; n++) {
        BlockDev &B = b.blk[n];
        const Level &L = b.lev[B.level];
        for (int f = 0; f < 6; f++) {
            B.nbr[f] = -1;
            B.kind[f] = FK_AMR;
        }
        if (!d.minkowski) {
            if (B.o[0] == 0) B.kind[0] = FK_OUTFLOW;
            if (B.o[0] + B.bs[0] == L.N[0]) B.kind[1] = FK_OUTFLOW;
            for (int side = 0; side < 2; side++) {
                const bool axis = side ? (B.o[1] + B.bs[1] == L.N[1]) : (B.o[1] == 0);
                if (!axis) continue;
                const int o3 = (B.o[2] + L.N[2] / 2) % L.N[2];
                auto it = where.find(std::make_tuple(B.level, B.o[0], B.o[1], o3, B.bs[0], B.bs[1], B.bs[2]));
                if (it == where.end()) return "amr: a polar block has no pi-shifted partner of its shape";
                B.kind[2 + side] = FK_POLE;
                B.nbr[2 + side] = it->second;
                B.pole |= 1 << side;
            }
        }
    }
    // push table: for each x3 face and each 16x16 tile of the block's columns,
    // the same-level leaf across the face that holds the tile (its ghost rows
    // there then copy this block's cells; the update writes them)
    g.push3tab.clear();
    for (int n = 0; n < nblk; n++) {
        BlockDev &B = b.blk[n];
        const int l = B.level;
        const Level &L = b.lev[l];
        const int t1n = (B.bs[0] + kPushTile - 1) / kPushTile, t2n = (B.bs[1] + kPushTile - 1) / kPushTile;
        for (int side = 0; side < 2; side++) {
            B.push3[side] = (int)g.push3tab.size();
            const int z = ((B.o[2] + (side ? B.bs[2] : -1)) % L.N[2] + L.N[2]) % L.N[2];
            bool any = false;
            for (int t1 = 0; t1 < t1n; t1++)
                for (int t2 = 0; t2 < t2n; t2++) {
                    const int Y0[3] = {B.o[0] + t1 * kPushTile, B.o[1] + t2 * kPushTile, z};
                    const int Y1[3] = {std::min(Y0[0] + kPushTile, B.o[0] + B.bs[0]) - 1,
                                       std::min(Y0[1] + kPushTile, B.o[1] + B.bs[1]) - 1, z};
                    int l0, l1, m = b.find(l, Y0, &l0), m1 = b.find(l, Y1, &l1);
                    if (m < 0 || l0 != l || m1 != m || l1 != l || B.bs[2] < NG) m = -1;
                    any = any || m >= 0;
                    g.push3tab.push_back(m);
                }
            if (!any) {
                g.push3tab.resize(B.push3[side]);
                B.push3[side] = -1;
            }
        }
    }
    // pole rings: same side, level, radial range and size, in phi order
    g.ring.clear();
    g.rside.clear();
    g.ring_len = 0;
    g.ni_max = 0;
    {
        std::map<std::tuple<int, int, int, int, int, int>, std::vector<std::pair<int, int>>> rings;
        for (int n = 0; n < nblk; n++) {
            const BlockDev &B = b.blk[n];
            for (int side = 0; side < 2; side++)
                if (B.pole & (1 << side))
                    rings[std::make_tuple(side, B.level, B.o[0], B.bs[0], B.bs[1], B.bs[2])].push_back({B.o[2], n});
        }
        for (auto &r : rings) {
            auto v = r.second;
            std::sort(v.begin(), v.end());
            const int side = std::get<0>(r.first);
            if (g.ring_len == 0) g.ring_len = (int)v.size();
            if ((int)v.size() != g.ring_len) return "amr: polar rings of different lengths";
            int tot = 0;
            for (auto &e : v) tot += b.blk[e.second].bs[2];
            if (tot != b.lev[b.blk[v[0].second].level].N[2]) return "amr: a polar ring does not close";
            for (auto &e : v) g.ring.push_back(e.second);
            g.rside.push_back(side);
            g.ni_max = std::max(g.ni_max, b.blk[v[0].second].bs[0] + 1);
        }
    }
    // exchange tables
    for (int s = 0; s < 3; s++) {
        g.ex[s].clear();
        g.pro[s].clear();
    }
    for (int n = 0; n < nblk; n++) {
        const BlockDev &B = b.blk[n];
        const int l = B.level;
        const Level &L = b.lev[l];
        for (int d0 = 0; d0 < 3; d0++) {
            const size_t ex0 = g.ex[d0].size(), pro0 = g.pro[d0].size();
            int lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                if (k < d0) { lo[k] = 0; hi[k] = B.D[k]; }
                else { lo[k] = NG; hi[k] = NG + B.bs[k]; }
            }
            for (int side = 0; side < 2; side++) {
                if (B.kind[2 * d0 + side] != FK_AMR) continue;
                lo[d0] = side ? NG + B.bs[d0] : 0;
                hi[d0] = side ? B.D[d0] : NG;
                for (int i = lo[0]; i < hi[0]; i++)
                    for (int j = lo[1]; j < hi[1]; j++)
                        for (int z = lo[2]; z < hi[2]; z++) {
                            const int dcell = (i * B.D[1] + j) * B.D[2] + z;
                            int Y[3] = {B.o[0] + i - NG, B.o[1] + j - NG, B.o[2] + z - NG};
                            Y[2] = ((Y[2] % L.N[2]) + L.N[2]) % L.N[2];
                            const bool out1 = Y[0] < 0 || Y[0] >= L.N[0];
                            const bool out2 = Y[1] < 0 || Y[1] >= L.N[1];
                            if (d.minkowski && (out1 || out2)) return "amr: flat boxes are periodic in x3 only";
                            if (out2) {
                                // polar ghost row: the same-level leaf at the clamped position holds it
                                int Yc[3] = {std::min(std::max(Y[0], 0), L.N[0] - 1), Y[1] < 0 ? 0 : L.N[1] - 1, Y[2]};
                                int lc, c = b.find(l, Yc, &lc);
                                if (c < 0 || lc != l || !storage_of(b.blk[c], Y))
                                    return "amr: a polar ghost row borders another level";
                                g.ex[d0].push_back({n, dcell, c, (AS_CELL << 28) | cell_in(b.blk[c], Y)});
                                continue;
                            }
                            // a corner ghost (beyond the block along an earlier
                            // direction): the same-level leaf across this face that
                            // holds the clamped position has it in its storage (its
                            // ghost of the earlier stage), face fields included
                            bool corner = false;
                            int Yk[3] = {Y[0], Y[1], Y[2]};
                            const int loc[3] = {i, j, z};
                            for (int e = 0; e < d0; e++)
                                if (loc[e] < NG || loc[e] >= NG + B.bs[e]) {
                                    corner = true;
                                    Yk[e] = std::min(std::max(Y[e], B.o[e]), B.o[e] + B.bs[e] - 1);
                                }
                            if (corner) {
                                int lk, k = b.find(l, Yk, &lk);
                                if (k >= 0 && lk == l) {
                                    if (!storage_of(b.blk[k], Y)) return "amr: corner ghost outside its source";
                                    g.ex[d0].push_back({n, dcell, k, (AS_CELL << 28) | cell_in(b.blk[k], Y)});
                                    continue;
                                }
                            }
                            if (out1) continue;  // final outflow pass
                            int lc, c = b.find(l, Y, &lc);
                            if (c < 0) return "amr: neighbouring blocks more than one level apart";
                            const BlockDev &C = b.blk[c];
                            if (lc == l) {
                                g.ex[d0].push_back({n, dcell, c, (AS_CELL << 28) | cell_in(C, Y)});
                            } else if (lc == l + 1) {
                                const int Yf[3] = {2 * Y[0], 2 * Y[1], 2 * Y[2]};
                                const int Yf1[3] = {2 * Y[0] + 1, 2 * Y[1] + 1, 2 * Y[2] + 1};
                                if (!interior_of(C, Yf) || !interior_of(C, Yf1)) return "amr: children split across blocks";
                                g.ex[d0].push_back({n, dcell, c, (AS_AVG8 << 28) | cell_in(C, Yf)});
                            } else if (lc == l - 1) {
                                const Level &Lc = b.lev[l - 1];
                                const int Pc[3] = {Y[0] >> 1, Y[1] >> 1, Y[2] >> 1};
                                if (!interior_of(C, Pc)) return "amr: coarse cell outside its leaf";
                                AmrProlong E;
                                E.dblk = n;
                                E.dcell = dcell;
                                E.cblk = c;
                                E.ccell = cell_in(C, Pc);
                                for (int dd = 0; dd < 3; dd++)
                                    for (int sg = 0; sg < 2; sg++) {
                                        int Q[3] = {Pc[0], Pc[1], Pc[2]};
                                        Q[dd] += sg ? 1 : -1;
                                        int *slot = E.nb[2 * dd + sg];
                                        if (dd == 2) Q[2] = ((Q[2] % Lc.N[2]) + Lc.N[2]) % Lc.N[2];
                                        if (Q[dd] < 0 || Q[dd] >= Lc.N[dd]) { slot[0] = 0; slot[1] = AS_NONE << 28; continue; }
                                        if (interior_of(C, Q)) { slot[0] = c; slot[1] = (AS_CELL << 28) | cell_in(C, Q); continue; }
                                        int lq, q = b.find(l - 1, Q, &lq);
                                        if (q < 0) return "amr: slope neighbour not found";
                                        const BlockDev &S = b.blk[q];
                                        if (lq == l - 1) { slot[0] = q; slot[1] = (AS_CELL << 28) | cell_in(S, Q); }
                                        else if (lq == l) {
                                            const int Qf[3] = {2 * Q[0], 2 * Q[1], 2 * Q[2]};
                                            const int Qf1[3] = {2 * Q[0] + 1, 2 * Q[1] + 1, 2 * Q[2] + 1};
                                            if (!interior_of(S, Qf) || !interior_of(S, Qf1)) return "amr: children split across blocks";
                                            slot[0] = q;
                                            slot[1] = (AS_AVG8 << 28) | cell_in(S, Qf);
                                        } else {
                                            const int sh = (l - 1) - lq;
                                            const int Qc[3] = {Q[0] >> sh, Q[1] >> sh, Q[2] >> sh};
                                            if (!interior_of(S, Qc)) return "amr: coarse neighbour outside its leaf";
                                            slot[0] = q;
                                            slot[1] = (AS_CELL << 28) | cell_in(S, Qc);
                                        }
                                    }
                                g.pro[d0].push_back(E);
                            } else {
                                return "amr: neighbouring blocks more than one level apart";
                            }
                        }
            }
            // the block's entries of this stage in memory order of the ghost cells
            // (the upper ghosts of one row next to the lower ghosts of the next)
            std::sort(g.ex[d0].begin() + ex0, g.ex[d0].end(),
                      [](const AmrEntry &x, const AmrEntry &y) { return x.dcell < y.dcell; });
            std::sort(g.pro[d0].begin() + pro0, g.pro[d0].end(),
                      [](const AmrProlong &x, const AmrProlong &y) { return x.dcell < y.dcell; });
        }
    }
    // stage 3: the ghost rows that the update writes (from the push table)
    // go last, so that a fill right after the update launches only the entries
    // before them; each must copy the cell that the update writes there
    {
        // the cell of block s that the push table sends to ghost cell dc of
        // block r, or -1
        auto pushed_from = [&](int r, int dc, int s) {
            const BlockDev &R = b.blk[r], &S = b.blk[s];
            const int i = dc / (R.D[1] * R.D[2]), j = (dc / R.D[2]) % R.D[1], z = dc % R.D[2];
            if (i < NG || i >= NG + R.bs[0] || j < NG || j >= NG + R.bs[1] || (z >= NG && z < NG + R.bs[2]))
                return -1;
            const int side = z < NG ? 1 : 0;  // the sender's face: its upper one for the receiver's lower ghosts
            if (S.level != R.level || S.push3[side] < 0) return -1;
            const int is = i + R.o[0] - S.o[0], js = j + R.o[1] - S.o[1];
            if (is < NG || is >= NG + S.bs[0] || js < NG || js >= NG + S.bs[1]) return -1;
            const int t2n = (S.bs[1] + kPushTile - 1) / kPushTile;
            if (g.push3tab[S.push3[side] + ((is - NG) / kPushTile) * t2n + (js - NG) / kPushTile] != r) return -1;
            const int zs = side ? z + S.bs[2] : z - R.bs[2];
            return (is * S.D[1] + js) * S.D[2] + zs;
        };
        auto pushed = [&](const AmrEntry &e) { return pushed_from(e.dblk, e.dcell, e.sblk) >= 0; };
        auto mid = std::stable_partition(g.ex[2].begin(), g.ex[2].end(), [&](const AmrEntry &e) { return !pushed(e); });
        g.npush3 = (int64_t)(g.ex[2].end() - mid);
        int64_t sent = 0;
        for (int n = 0; n < nblk; n++)
            for (int side = 0; side < 2; side++) {
                const BlockDev &B = b.blk[n];
                if (B.push3[side] < 0) continue;
                const int t1n = (B.bs[0] + kPushTile - 1) / kPushTile, t2n = (B.bs[1] + kPushTile - 1) / kPushTile;
                for (int t = 0; t < t1n * t2n; t++)
                    if (g.push3tab[B.push3[side] + t] >= 0)
                        sent += (int64_t)NG * (std::min(kPushTile, B.bs[0] - (t / t2n) * kPushTile)) *
                                std::min(kPushTile, B.bs[1] - (t % t2n) * kPushTile);
            }
        if (sent != g.npush3) return "amr: the push table writes ghosts that the exchange fills otherwise";
        for (auto it = mid; it != g.ex[2].end(); ++it)
            if (it->scode != ((AS_CELL << 28) | pushed_from(it->dblk, it->dcell, it->sblk)))
                return "amr: an x3 ghost row of the push table has another source";
    }
    // flux correction at level jumps (flux_rec1-3 with calc_corr = 1,
    // boundcomF.c:583-1315): every interior face cell of a block face whose
    // neighbour across it is one level finer, with the finer block's face that
    // tiles its lowest transverse quarter
    for (int d = 0; d < 3; d++) {
        g.rfx[d].clear();
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;  // transverse, slower first
        for (int n = 0; n < nblk; n++) {
            const BlockDev &C = b.blk[n];
            const Level &L = b.lev[C.level];
            for (int side = 0; side < 2; side++) {
                int Y[3];
                Y[d] = side ? C.o[d] + C.bs[d] : C.o[d] - 1;
                if (d < 2 && (Y[d] < 0 || Y[d] >= L.N[d])) continue;  // radial edge or axis
                if (d == 2) Y[2] = ((Y[2] % L.N[2]) + L.N[2]) % L.N[2];
                for (int ya = C.o[ta]; ya < C.o[ta] + C.bs[ta]; ya++)
                    for (int yb = C.o[tb]; yb < C.o[tb] + C.bs[tb]; yb++) {
                        Y[ta] = ya;
                        Y[tb] = yb;
                        int lc = -1;
                        const int f = b.find(C.level, Y, &lc);
                        if (f < 0 || lc != C.level + 1) continue;
                        const BlockDev &Fb = b.blk[f];
                        int Z[3];
                        Z[d] = side ? 2 * Y[d] : 2 * Y[d] + 1;  // the finer cell beside the face
                        Z[ta] = 2 * ya;
                        Z[tb] = 2 * yb;
                        if (!interior_of(Fb, Z)) return "amr: flux correction source outside its block";
                        const int sd = d == 0 ? Fb.D[1] * Fb.D[2] : d == 1 ? Fb.D[2] : 1;
                        int K[3] = {Y[0], Y[1], Y[2]};
                        K[d] = side ? C.o[d] + C.bs[d] : C.o[d];
                        g.rfx[d].push_back({n, cell_in(C, K), f, cell_in(Fb, Z) + (side ? 0 : sd)});
                    }
            }
        }
    }
    // EMF correction at level jumps (E_send1-3 / E_rec1-3 with calc_corr = 1 and
    // the edge exchanges E*_send_corn / E*_receive_corn, boundcomE.c): every
    // edge EMF on a block's boundary that one of the (up to four) leaves around
    // the edge refines by one level takes the mean of that leaf's two edges
    // along it.  The leaf is the first finer one in a fixed order of the
    // quadrants around the edge, so every block holding the edge gets the same
    // value.
    {
        g.emf.clear();
        std::vector<char> dest;  // (block, comp, cell) marks, to check that no source is rewritten
        std::vector<int64_t> base(nblk + 1, 0);
        for (int n = 0; n < nblk; n++)
            base[n + 1] = base[n] + 3 * (int64_t)b.blk[n].D[0] * b.blk[n].D[1] * b.blk[n].D[2];
        dest.assign((size_t)base[nblk], 0);
        for (int n = 0; n < nblk; n++) {
            const BlockDev &C = b.blk[n];
            const int l = C.level;
            if (l + 1 >= a.nlev) continue;
            const Level &L = b.lev[l];
            const Level &Lf = b.lev[l + 1];
            for (int t = 0; t < 3; t++) {
                const int p = (t + 1) % 3, q = (t + 2) % 3;
                for (int kt = C.o[t]; kt < C.o[t] + C.bs[t]; kt++)
                    for (int kp = C.o[p]; kp <= C.o[p] + C.bs[p]; kp++) {
                        const bool pb = (kp == C.o[p] || kp == C.o[p] + C.bs[p]);
                        for (int kq = C.o[q]; kq <= C.o[q] + C.bs[q]; kq += (pb ? 1 : C.bs[q])) {
                            int f = -1;
                            for (int qd = 0; qd < 4 && f < 0; qd++) {
                                int Y[3];
                                Y[t] = kt;
                                Y[p] = kp - 1 + (qd >> 1);
                                Y[q] = kq - 1 + (qd & 1);
                                bool in = true;
                                for (int x : {p, q}) {
                                    if (x < 2 && (Y[x] < 0 || Y[x] >= L.N[x])) in = false;
                                    if (x == 2) Y[2] = ((Y[2] % L.N[2]) + L.N[2]) % L.N[2];
                                }
                                if (!in) continue;
                                int lc = -1;
                                const int cand = b.find(l, Y, &lc);
                                if (cand >= 0 && lc == l + 1) f = cand;
                            }
                            if (f < 0) continue;
                            const BlockDev &Fb = b.blk[f];
                            int Z[3];
                            Z[t] = 2 * kt;
                            Z[p] = 2 * kp;
                            Z[q] = 2 * kq;
                            if (Z[2] < Fb.o[2] || Z[2] > Fb.o[2] + Fb.bs[2])  // across the periodic x3 seam
                                Z[2] += (Z[2] < Fb.o[2]) ? Lf.N[2] : -Lf.N[2];
                            for (int x = 0; x < 3; x++) {
                                const int hi = Fb.o[x] + Fb.bs[x] - (x == t ? 2 : 0);
                                if (Z[x] < Fb.o[x] || Z[x] > hi) return "amr: EMF correction source outside its block";
                            }
                            int K[3];
                            K[t] = kt;
                            K[p] = kp;
                            K[q] = kq;
                            const int cc = cell_in(C, K), fc = cell_in(Fb, Z);
                            dest[base[n] + (int64_t)t * C.D[0] * C.D[1] * C.D[2] + cc] = 1;
                            g.emf.push_back({n, cc | (t << 28), f, fc});
                        }
                    }
            }
        }
        for (const AmrEmf &e : g.emf) {
            const int t = e.cc >> 28;
            const BlockDev &Fb = b.blk[e.fb];
            const int64_t nc = (int64_t)Fb.D[0] * Fb.D[1] * Fb.D[2];
            const int st = t == 0 ? Fb.D[1] * Fb.D[2] : t == 1 ? Fb.D[2] : 1;
            if (dest[base[e.fb] + t * nc + e.fc] || dest[base[e.fb] + t * nc + e.fc + st])
                return "amr: an EMF correction source is itself corrected";
        }
    }
    g.blocks = b.blk;
    g.nlev = a.nlev;
    for (int l = 0; l < a.nlev; l++)
        for (int k = 0; k < 3; k++) g.N[l][k] = b.lev[l].N[k];
    return validate_amr(g);
}

// Every index a table entry names lies inside its block's slab: checked on the
// host before the tables reach the device.
static bool src_ok(const std::vector<BlockDev> &blk, int b, int code)
{
    if (b < 0 || b >= (int)blk.size()) return false;
    const BlockDev &S = blk[b];
    const int kind = code >> 28, c = code & 0x0FFFFFFF;
    const int64_t ncell = (int64_t)S.D[0] * S.D[1] * S.D[2];
    if (kind == AS_NONE) return true;
    if (kind == AS_CELL) return c >= 0 && c < ncell;
    if (kind != AS_AVG8) return false;
    const int64_t last = (int64_t)c + (int64_t)S.D[1] * S.D[2] + S.D[2] + 1;
    return c >= 0 && last < ncell;
}

// ---- local adaptive time stepping ----

namespace {

// Leaf lookup over a finished block set: owner of each 16^3 unit per level.
struct LeafMap {
    int nlev = 0;
    int N[kAmrMaxLevels][3] = {};
    std::vector<int> own[kAmrMaxLevels];
    int U(int l, int k) const { return N[l][k] / UNIT; }
    int &at(int l, int a, int b, int c) { return own[l][((size_t)a * U(l, 1) + b) * U(l, 2) + c]; }
    int get(int l, int a, int b, int c) const { return own[l][((size_t)a * U(l, 1) + b) * U(l, 2) + c]; }
    explicit LeafMap(const AmrGrid &g)
    {
        nlev = g.nlev;
        for (int l = 0; l < nlev; l++) {
            for (int k = 0; k < 3; k++) N[l][k] = g.N[l][k];
            own[l].assign((size_t)U(l, 0) * U(l, 1) * U(l, 2), -1);
        }
        for (int n = 0; n < (int)g.blocks.size(); n++) {
            const BlockDev &B = g.blocks[n];
            for (int a = B.o[0] / UNIT; a < (B.o[0] + B.bs[0]) / UNIT; a++)
                for (int b = B.o[1] / UNIT; b < (B.o[1] + B.bs[1]) / UNIT; b++)
                    for (int c = B.o[2] / UNIT; c < (B.o[2] + B.bs[2]) / UNIT; c++) at(B.level, a, b, c) = n;
        }
    }
    // leaf holding cell Y of level l (inside the grid), *lc its level; -1 if none
    int find(int l, const int Y[3], int *lc) const
    {
        int o = get(l, Y[0] / UNIT, Y[1] / UNIT, Y[2] / UNIT);
        if (o >= 0) { *lc = l; return o; }
        if (l + 1 < nlev) {
            o = get(l + 1, 2 * Y[0] / UNIT, 2 * Y[1] / UNIT, 2 * Y[2] / UNIT);
            if (o >= 0) { *lc = l + 1; return o; }
        }
        for (int c = l - 1; c >= 0; c--) {
            const int sh = l - c;
            o = get(c, (Y[0] >> sh) / UNIT, (Y[1] >> sh) / UNIT, (Y[2] >> sh) / UNIT);
            if (o >= 0) { *lc = c; return o; }
        }
        return -1;
    }
};

}  // namespace

std::vector<std::pair<int, int>> jump_pairs(const AmrGrid &g)
{
    std::vector<std::pair<int, int>> v;
    for (int d = 0; d < 3; d++)
        for (const AmrReflux &e : g.rfx[d]) v.emplace_back(e.cb, e.fb);
    for (const AmrEmf &e : g.emf) v.emplace_back(e.cb, e.fb);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

void las_rule(int nblk, const double *bdt, double dt, int maxtl, const std::vector<int> &ring, int ring_len,
              const std::vector<std::pair<int, int>> &pairs, int *tl)
{
    int imax = 0;
    while ((2 << imax) <= maxtl) imax++;
    for (int n = 0; n < nblk; n++) {
        tl[n] = 1;
        for (int i = imax; i >= 0; i--)
            if (bdt[n] / dt >= 1.05 * std::pow(2., i)) {
                tl[n] = 1 << i;
                break;
            }
    }
    const int nring = ring_len > 0 ? (int)ring.size() / ring_len : 0;
    for (bool changed = true; changed;) {
        changed = false;
        for (int r = 0; r < nring; r++) {
            int m = kMaxTimeLevel;
            for (int k = 0; k < ring_len; k++) m = std::min(m, tl[ring[(size_t)r * ring_len + k]]);
            for (int k = 0; k < ring_len; k++) {
                int &x = tl[ring[(size_t)r * ring_len + k]];
                if (x != m) { x = m; changed = true; }
            }
        }
        for (const auto &pr : pairs)
            if (tl[pr.second] > tl[pr.first]) {
                tl[pr.second] = tl[pr.first];
                changed = true;
            }
    }
}

std::string build_las_tables(const AmrGrid &g, const std::vector<int> &tl, LasTables &t)
{
    const auto &blk = g.blocks;
    const int nblk = (int)blk.size();
    if ((int)tl.size() != nblk) return "las: one time level per block";
    for (int n = 0; n < nblk; n++)
        if (tl[n] < 1 || tl[n] > kMaxTimeLevel || (tl[n] & (tl[n] - 1)))
            return "las: time levels are powers of 2 from 1 to " + std::to_string(kMaxTimeLevel);
    if (g.ring_len > 0)
        for (size_t r = 0; r < g.ring.size(); r += g.ring_len)
            for (int k = 1; k < g.ring_len; k++)
                if (tl[g.ring[r + k]] != tl[g.ring[r]]) return "las: the blocks of a polar ring share one time level";
    for (const auto &pr : jump_pairs(g))
        if (tl[pr.second] > tl[pr.first])
            return "las: a finer block at a level jump has a larger time level than its coarser neighbour";
    const LeafMap M(g);
    auto cell = [](const BlockDev &B, const int K[3]) {
        return ((K[0] - B.o[0] + NG) * B.D[1] + (K[1] - B.o[1] + NG)) * B.D[2] + (K[2] - B.o[2] + NG);
    };
    // fluxes: level jumps, then same-level faces whose neighbour is faster
    for (int d = 0; d < 3; d++) {
        auto &v = t.flux[d];
        v.clear();
        for (const AmrReflux &e : g.rfx[d]) v.push_back({e.cb, e.cf, e.fb, e.ff | LAS_MEAN});
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        for (int n = 0; n < nblk; n++) {
            const BlockDev &B = blk[n];
            const int l = B.level;
            for (int side = 0; side < 2; side++) {
                int Y[3];
                Y[d] = side ? B.o[d] + B.bs[d] : B.o[d] - 1;
                if (d < 2 && (Y[d] < 0 || Y[d] >= M.N[l][d])) continue;
                if (d == 2) Y[2] = ((Y[2] % M.N[l][2]) + M.N[l][2]) % M.N[l][2];
                for (int ya = B.o[ta]; ya < B.o[ta] + B.bs[ta]; ya++)
                    for (int yb = B.o[tb]; yb < B.o[tb] + B.bs[tb]; yb++) {
                        Y[ta] = ya;
                        Y[tb] = yb;
                        int lc = -1;
                        const int m = M.find(l, Y, &lc);
                        if (m < 0 || lc != l || m == n || tl[m] >= tl[n]) continue;
                        const BlockDev &S = blk[m];
                        int K[3] = {Y[0], Y[1], Y[2]}, Ks[3] = {Y[0], Y[1], Y[2]};
                        K[d] = side ? B.o[d] + B.bs[d] : B.o[d];
                        Ks[d] = side ? S.o[d] : S.o[d] + S.bs[d];
                        v.push_back({n, cell(B, K), m, cell(S, Ks)});
                    }
            }
        }
    }
    // EMFs: level-jump edges as g.emf; other boundary edges take the EMF of
    // the fastest same-level leaf around them (first in the quadrant order
    // (-,-), (-,+), (+,-), (+,+) on ties) when it is faster than the block
    t.emf_jump.clear();
    t.emf_same.clear();
    for (const AmrEmf &e : g.emf) t.emf_jump.push_back({e.cb, e.cc, e.fb, e.fc | LAS_MEAN});
    for (int n = 0; n < nblk; n++) {
        const BlockDev &B = blk[n];
        const int l = B.level;
        for (int tt = 0; tt < 3; tt++) {
            const int p = (tt + 1) % 3, q = (tt + 2) % 3;
            int K[3];
            for (K[tt] = B.o[tt]; K[tt] < B.o[tt] + B.bs[tt]; K[tt]++)
                for (K[p] = B.o[p]; K[p] <= B.o[p] + B.bs[p]; K[p]++) {
                    const bool pb = (K[p] == B.o[p] || K[p] == B.o[p] + B.bs[p]);
                    for (K[q] = B.o[q]; K[q] <= B.o[q] + B.bs[q]; K[q] += (pb ? 1 : B.bs[q])) {
                        int best = -1, finer = 0;
                        for (int qd = 0; qd < 4; qd++) {
                            int Y[3] = {K[0], K[1], K[2]}, in = 1;
                            Y[p] += (qd >> 1) - 1;
                            Y[q] += (qd & 1) - 1;
                            for (int x = 0; x < 2; x++)
                                if (Y[x] < 0 || Y[x] >= M.N[l][x]) in = 0;
                            Y[2] = ((Y[2] % M.N[l][2]) + M.N[l][2]) % M.N[l][2];
                            if (!in) continue;
                            int lc = -1;
                            const int m = M.find(l, Y, &lc);
                            if (m < 0) continue;
                            if (lc == l + 1) finer = 1;
                            if (lc == l && (best < 0 || tl[m] < tl[best])) best = m;
                        }
                        if (finer || best < 0 || tl[best] >= tl[n]) continue;
                        const BlockDev &S = blk[best];
                        int Ks[3] = {K[0], K[1], K[2]};
                        if (Ks[2] < S.o[2]) Ks[2] += M.N[l][2];
                        if (Ks[2] > S.o[2] + S.bs[2]) Ks[2] -= M.N[l][2];
                        for (int x = 0; x < 3; x++) {
                            const int hi = S.o[x] + S.bs[x] - (x == tt ? 1 : 0);
                            if (Ks[x] < S.o[x] || Ks[x] > hi) return "las: EMF correction source outside its block";
                        }
                        t.emf_same.push_back({n, cell(B, K) | (tt << 28), best, cell(S, Ks)});
                    }
                }
        }
    }
    return "";
}

std::string validate_amr(const AmrGrid &g)
{
    const auto &blk = g.blocks;
    auto dst_ok = [&](int b, int c) {
        if (b < 0 || b >= (int)blk.size()) return false;
        const BlockDev &B = blk[b];
        return c >= 0 && (int64_t)c < (int64_t)B.D[0] * B.D[1] * B.D[2];
    };
    for (int s = 0; s < 3; s++) {
        for (const AmrEntry &e : g.ex[s])
            if (!dst_ok(e.dblk, e.dcell) || !src_ok(blk, e.sblk, e.scode) || (e.scode >> 28) == AS_NONE)
                return "amr: exchange entry out of range";
        for (const AmrProlong &e : g.pro[s]) {
            if (!dst_ok(e.dblk, e.dcell) || !src_ok(blk, e.cblk, (AS_CELL << 28) | e.ccell))
                return "amr: prolongation entry out of range";
            for (int k = 0; k < 6; k++)
                if (!src_ok(blk, e.nb[k][0], e.nb[k][1])) return "amr: prolongation neighbour out of range";
        }
    }
    for (int d = 0; d < 3; d++)
        for (const AmrReflux &e : g.rfx[d]) {
            if (!dst_ok(e.cb, e.cf)) return "amr: flux correction entry out of range";
            if (e.fb < 0 || e.fb >= (int)blk.size()) return "amr: flux correction source out of range";
            const BlockDev &F = blk[e.fb];
            const int sa = d == 0 ? F.D[2] : F.D[1] * F.D[2], sb = d == 2 ? F.D[2] : 1;
            if (!dst_ok(e.fb, e.ff) || !dst_ok(e.fb, e.ff + sa + sb))
                return "amr: flux correction source out of range";
        }
    for (const AmrEmf &e : g.emf) {
        const int t = e.cc >> 28;
        if (t < 0 || t > 2 || !dst_ok(e.cb, e.cc & 0x0FFFFFFF)) return "amr: EMF correction entry out of range";
        if (e.fb < 0 || e.fb >= (int)blk.size()) return "amr: EMF correction source out of range";
        const BlockDev &F = blk[e.fb];
        const int st = t == 0 ? F.D[1] * F.D[2] : t == 1 ? F.D[2] : 1;
        if (!dst_ok(e.fb, e.fc) || !dst_ok(e.fb, e.fc + st)) return "amr: EMF correction source out of range";
    }
    for (size_t n = 0; n < blk.size(); n++) {
        const BlockDev &B = blk[n];
        for (int f = 0; f < 6; f++)
            if (B.kind[f] == FK_POLE && (B.nbr[f] < 0 || B.nbr[f] >= (int)blk.size())) return "amr: bad pole partner";
        // every receiver holds the tile's ghost rows (the kernels index its slab with them)
        for (int side = 0; side < 2; side++) {
            if (B.push3[side] < 0) continue;
            const int t1n = (B.bs[0] + kPushTile - 1) / kPushTile, t2n = (B.bs[1] + kPushTile - 1) / kPushTile;
            if (B.bs[2] < 2 * NG || B.push3[side] + t1n * t2n > (int)g.push3tab.size()) return "amr: bad push table";
            for (int t = 0; t < t1n * t2n; t++) {
                const int m = g.push3tab[B.push3[side] + t];
                if (m < 0) continue;
                if (m >= (int)blk.size()) return "amr: bad push table";
                const BlockDev &M = blk[m];
                const int lo[2] = {B.o[0] + (t / t2n) * kPushTile, B.o[1] + (t % t2n) * kPushTile};
                const int hi[2] = {std::min(lo[0] + kPushTile, B.o[0] + B.bs[0]), std::min(lo[1] + kPushTile, B.o[1] + B.bs[1])};
                for (int k = 0; k < 2; k++)
                    if (lo[k] < M.o[k] || hi[k] > M.o[k] + M.bs[k]) return "amr: bad push table";
            }
        }
    }
    if (g.npush3 < 0 || g.npush3 > (int64_t)g.ex[2].size()) return "amr: bad push3 count";
    for (int r : g.ring)
        if (r < 0 || r >= (int)blk.size()) return "amr: bad ring entry";
    return "";
}

}  // namespace hamr
