// hamr_kernels_dev.h -- device helpers shared by the kernel translation units:
// work-list walking, metric addressing and loads, wavefront reductions.
#pragma once

#include "hamr_kernels.h"

namespace hamr {


__device__ __forceinline__ void region_index(int t, int n2, int n3, int &a, int &b, int &c)
{
    c = t % n3;
    int q = t / n3;
    b = q % n2;
    a = q / n2;
}

// Item w of the list in visiting order (see WorkList).
__device__ __forceinline__ int2 wl_get(const WorkList &wl, int w)
{
    if (wl.swz) {
        const int q = wl.n >> 3, r = wl.n & 7, x = w & 7, k = w >> 3;
        w = x * q + (x < r ? x : r) + k;
    }
    return wl.items[w];
}

#define FOR_ITEMS(wl)                                                                  \
    for (int w_ = blockIdx.x; w_ < (wl).n; w_ += gridDim.x)

// Metric addressing of a block's geometry slab (axisymmetric: one column per
// (i, j) of the block, padded to kcol; moff columns of earlier slabs).
__device__ __forceinline__ int64_t mg10(const BlockDev &B, int pos, int s) { return B.moff * (NPOS * 10) + ((int64_t)pos * 10 + s) * B.kcol; }
__device__ __forceinline__ int64_t mg1(const BlockDev &B, int pos) { return B.moff * NPOS + (int64_t)pos * B.kcol; }

// Metric at one position of one column in the arithmetic R: the FP64 metric,
// or for R = float its FP32 copy (the FP64 values rounded, and 1 - alpha
// formed in FP64 and rounded, by k_geom32 when the metric was uploaded).
template <class R = double>
__device__ __forceinline__ Geom<R> load_geom(const KParams &P, const BlockDev &B, int pos, int64_t col)
{
    Geom<R> G;
    if constexpr (sizeof(R) == 4) {
        const float *g = P.g32 + B.moff * (NPOS * G32N) + (int64_t)pos * G32N * B.kcol + col;
#pragma unroll
        for (int s = 0; s < 10; s++) {
            G.gcov[s] = g[s * B.kcol];
            G.gcon[s] = g[(10 + s) * B.kcol];
        }
        G.g = g[20 * B.kcol];
        G.oma = g[21 * B.kcol];
    } else {
        const double *gc = P.gcov + mg10(B, pos, 0) + col;
        const double *gi = P.gcon + mg10(B, pos, 0) + col;
#pragma unroll
        for (int s = 0; s < 10; s++) {
            G.gcov[s] = gc[s * B.kcol];
            G.gcon[s] = gi[s * B.kcol];
        }
        G.g = P.gdet[mg1(B, pos) + col];
    }
    return G;
}

// The step Dt of block B: the launch's Dt (dt_ptr[0] dt_mult when the launch
// reads dt from the device, else Dt_val), scaled by the block's time level in
// a local-adaptive-time-stepping cycle (a block of time level tl advances by
// tl dt per full step, advance_GPU, step_ch.c:1184-1262).
__device__ __forceinline__ double block_Dt(const KParams &P, const BlockDev &B, double Dt_val,
                                           const double *__restrict__ dt_ptr, double dt_mult)
{
    const double Dt = dt_ptr ? dt_ptr[0] * dt_mult : Dt_val;
    return P.las_s >= 0 ? Dt * (double)(B.tl > 0 ? B.tl : 1) : Dt;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Where the update and face-field kernels also write cell (i, j, z) (index g)
// of block B, one of the NG lowest (side 0) or highest (side 1) cells of its
// z-row: the x3 ghost cell of the same-level block across that face that the
// push table names for the cell's tile, as slab offset, component stride and
// cell index there.  false: nothing (the exchange fills that ghost), and in
// a time-level cycle when the receiver does not step in this sub-iteration
// (its exchange fills a ghost only when both blocks step; when they do, the
// caller writes both of the receiver's array pairs).  When one entry serves
// each face (kPushWhole: every face of a uniform grid), both faces' entries
// are read at wave-uniform addresses and each lane selects its side's.
struct PushTarget {
    int64_t off, ks, cell;
};
// TILED = false: the kernels' form for uniform grids, whose faces all have
// one entry with a cell offset (no per-tile entries, no general index).
template <bool TILED = true>
__device__ __forceinline__ bool x3_push_target(const KParams &P, const BlockDev &B, int i, int j, int z, int64_t g,
                                               int side, PushTarget &T)
{
    if (!TILED && P.las_s >= 0) return false;
    const int r0 = B.push3[0], r1 = B.push3[1];
    int m, delta, ks;
    int64_t off;
    if (!TILED || ((r0 < 0 || (r0 & kPushWhole)) && (r1 < 0 || (r1 & kPushWhole)))) {
        const PushEntry none{-1, 0, 0, 0, 0};
        const PushEntry E0 = r0 >= 0 ? P.push3[r0 & ~kPushWhole] : none;
        const PushEntry E1 = r1 >= 0 ? P.push3[r1 & ~kPushWhole] : none;
        m = side ? E1.m : E0.m;
        delta = side ? E1.delta : E0.delta;
        ks = side ? E1.ks : E0.ks;
        off = side ? E1.off : E0.off;
    } else {
        const int row = side ? r1 : r0;
        if (row < 0) return false;
        const int t2n = (B.bs[1] + kPushTile - 1) / kPushTile;
        const PushEntry &E = P.push3[(row & kPushWhole) ? row & ~kPushWhole
                                                        : row + ((i - NG) / kPushTile) * t2n + (j - NG) / kPushTile];
        m = E.m;
        delta = E.delta;
        ks = E.ks;
        off = E.off;
    }
    if (m < 0) return false;
    if (TILED && P.las_s >= 0 && las_phase(P.blk[m].tl, P.las_s) == LAS_IDLE) return false;
    T.off = off;
    T.ks = ks;
    if (!TILED || delta != kPushGeneral) {
        T.cell = g + delta;
        return true;
    }
    const BlockDev &M = P.blk[m];
    const int im = i + B.o[0] - M.o[0], jm = j + B.o[1] - M.o[1];
    const int zm = side ? z - B.bs[2] : z + M.bs[2];
    T.cell = ((int64_t)im * M.D[1] + jm) * M.D[2] + zm;
    return true;
}

// Whether the x3 ghosts of row (i, j) of block n (inside its x1 and x2 range)
// beyond face f (4: lower, 5: upper) are written by the update of the copy
// neighbour across f (its push table names n for that row's tile).
__device__ __forceinline__ bool x3_row_pushed(const KParams &P, const BlockDev &B, int n, int f, int i, int j)
{
    const int s = B.nbr[f];
    if (P.las_s >= 0 || s < 0) return false;
    const BlockDev &S = P.blk[s];
    const int row = S.push3[f == 4 ? 1 : 0];
    if (row < 0) return false;
    const int is = i + B.o[0] - S.o[0], js = j + B.o[1] - S.o[1];
    if (is < NG || is >= NG + S.bs[0] || js < NG || js >= NG + S.bs[1]) return false;
    if (row & kPushWhole) return P.push3[row & ~kPushWhole].m == n;
    const int t2n = (S.bs[1] + kPushTile - 1) / kPushTile;
    return P.push3[row + ((is - NG) / kPushTile) * t2n + (js - NG) / kPushTile].m == n;
}

// launch geometry of a work-list launch
static inline dim3 gdim(const Launch &L) { return dim3((unsigned)L.grid); }
static inline dim3 bdim(const Launch &L) { return dim3((unsigned)L.threads); }

}  // namespace hamr
