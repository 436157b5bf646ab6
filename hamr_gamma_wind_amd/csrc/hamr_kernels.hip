// hamr_kernels.hip -- gfx950 kernels of the GRMHD fluid step.
//
// Every kernel walks a work list of (block, tile) items, so that one launch
// serves a single block (the reference-shaped per-block calls), the batched step
// over the whole resident block set, and block sets of uneven sizes on several
// refinement levels.  Block size, slab offsets, metric slab and spacing come
// from the block table.  Threads run along z (the fastest index of the block
// layout) so that every global access of a wavefront is one or two contiguous
// 512-byte segments.  Built without FMA contraction so that the FP64 arithmetic
// matches the reference's expression by expression.
#include "hamr_recon.h"

#include <algorithm>
#include <cstdio>

namespace hamr {

int64_t class_work(const BlockDev &B, int cls)
{
    const int64_t b1 = B.bs[0], b2 = B.bs[1], b3 = B.bs[2], d1 = B.D[0], d2 = B.D[1], d3 = B.D[2];
    switch (cls) {
    case KC_FLUX1: return (b1 + 1) * (b2 + 2) * (b3 + 2);
    case KC_FLUX2: return (b1 + 2) * (b2 + 1) * (b3 + 2);
    case KC_FLUX3: return (b1 + 2) * (b2 + 2) * (b3 + 1);
    case KC_ECENT: return (b1 + 2) * (b2 + 2) * (b3 + 2);
    case KC_CORNER: return (b1 + 1) * (b2 + 1) * (b3 + 1);
    case KC_INTERIOR: return b1 * b2 * b3;
    case KC_BOUND1: return 2 * NG * d2 * d3;
    case KC_BOUND2: return 2 * NG * d1 * d3;
    default: return 2 * NG * (d1 * d2 + 1);
    }
}

// Polar E1: mean over each pole block's interior corners in phi (weighted by
// dt as packsend2E does), mean over the ring of phi-blocks, divided back by dt
// (E_average / read_E_avg / write_E_avg, const_trans.c:268-543).  One thread per
// (ring, radial corner); the blocks of a ring share level, radial range and size.
template <class R>
__global__ void k_pole_average(KParams P, const int *__restrict__ ring, const int *__restrict__ rside, int nring,
                               int len, int ni_max, double dt_val, const double *__restrict__ dt_ptr, R *__restrict__ E)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nring * ni_max) return;
    const int il = t % ni_max, rg = t / ni_max;
    const int *rb = ring + (int64_t)rg * len;
    const BlockDev &B0 = P.blk[rb[0]];
    if (il > B0.bs[0]) return;
    // in a local-time-stepping cycle: the rings of this launch's phase, each
    // weighted by its time level's step (a ring's blocks share one time level)
    if (P.las_s >= 0 && las_phase(B0.tl, P.las_s) != P.las_want) return;
    const int side = rside[rg];
    const int i = il + NG, j = side ? NG + B0.bs[1] : NG;
    const double dt = block_Dt(P, B0, dt_val, dt_ptr, 1.0);
    double avg = 0.;
    for (int b = 0; b < len; b++) {
        const BlockDev &B = P.blk[rb[b]];
        const R *e1 = E + 3 * B.off + ((int64_t)i * B.D[1] + j) * B.D[2] + NG;
        double s = 0.;
        // the sum runs in z order; the loads of 16 terms are issued before
        // their additions, so that one thread does not wait on one load at a time
        int zz = 0;
        for (; zz + 16 <= B.bs[2]; zz += 16) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; q++) v[q] = (double)e1[zz + q];
#pragma unroll
            for (int q = 0; q < 16; q++) s += dt * v[q];
        }
        for (; zz < B.bs[2]; zz++) s += dt * (double)e1[zz];
        s /= (double)(B.bs[2]);
        if (b == 0) avg = s / ((double)len);
        else avg += s / ((double)len);
    }
    const double v = avg / dt;
    for (int b = 0; b < len; b++) {
        const BlockDev &B = P.blk[rb[b]];
        R *e = E + 3 * B.off + ((int64_t)i * B.D[1] + j) * B.D[2];
        for (int z = NG; z <= NG + B.bs[2]; z++) {
            e[z] = (R)v;
            e[2 * (int64_t)B.ks + z] = R(0.);
        }
    }
}

// Face-B update from the curl of the corner EMF, consttransport3
// (GPU_const_trans.h:426-530); faces [NG, NG+BS] along their normal.  psi and
// psf may alias (each thread reads only the faces it writes).
template <class R, bool TILED>
__global__ void __launch_bounds__(256) k_ct_update(KParams P, WorkList wl, double Dt_val,
                                                   const double *__restrict__ dt_ptr, double dt_mult, const double *psi,
                                                   double *psf, const R *__restrict__ E)
{
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const BlockDev &B = P.blk[it.x];
        const int e1 = B.bs[0] + 1, e2 = B.bs[1] + 1, e3 = B.bs[2] + 1;
        const int t = it.y * blockDim.x + threadIdx.x;
        if (t >= e1 * e2 * e3) continue;
        int a, b, c;
        region_index(t, e2, e3, a, b, c);
        const int i = a + NG, j = b + NG, z = c + NG;
        const double Dt = block_Dt(P, B, Dt_val, dt_ptr, dt_mult);
        const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2], sj = B.D[2];
        const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z, col = (int64_t)i * B.D[1] + j;
        const double gF1 = P.gdet[mg1(B, FACE1) + col], gF2 = P.gdet[mg1(B, FACE2) + col],
                     gF3 = P.gdet[mg1(B, FACE3) + col];
        const R *Ebr = E + 3 * B.off + g;
        auto Eb = [&](int64_t o) { return (double)Ebr[o]; };
        const double *si_ = psi + 3 * B.off + g;
        double *sf = psf + 3 * B.off + g;
        const double dx1 = B.dx[0], dx2 = B.dx[1], dx3 = B.dx[2];
        double temp;
        // the x3 ghosts of the push table's receiver, as in finish_cell: B1 and
        // B2 of the NG lowest / highest cells of the z-row, B3 of the faces
        // that are not the block's own x3 boundary faces (each block forms those)
        const int bs3 = B.bs[2];
        const bool inner = i < NG + B.bs[0] && j < NG + B.bs[1];
        PushTarget T;
        double *sm = nullptr, *sm2 = nullptr;  // sm2: a time-level cycle's second array pair
        if (inner && (z < 2 * NG || (z >= bs3 && z < NG + bs3)) &&
            x3_push_target<TILED>(P, B, i, j, z, g, z < 2 * NG ? 0 : 1, T)) {
            if (TILED && P.las_s >= 0) {
                sm = P.xps + 3 * T.off + T.cell;
                sm2 = P.xpsh + 3 * T.off + T.cell;
            } else {
                sm = psf + 3 * T.off + T.cell;
            }
        }
        const int64_t KM = sm ? T.ks : 0;
        if (j < NG + B.bs[1] && z < NG + B.bs[2]) {
            temp = 0.;
            temp += 1.0 / 1.0 * si_[0] * gF1;
            temp += -Dt / (1.0 * dx2) * (Eb(2 * K + sj) - Eb(2 * K));
            temp += Dt / (1.0 * dx3) * (Eb(K + 1) - Eb(K));
            sf[0] = temp / gF1;
            if (sm) sm[0] = sf[0];
            if (sm2) sm2[0] = sf[0];
        }
        if (i < NG + B.bs[0] && j < NG + B.bs[1]) {
            temp = si_[2 * K] * gF3;
            temp += -Dt / dx1 * (Eb(K + si) - Eb(K));
            temp += Dt / dx2 * (Eb(sj) - Eb(0));
            sf[2 * K] = temp / gF3;
            if (sm && z != NG) sm[2 * KM] = sf[2 * K];
            if (sm2 && z != NG) sm2[2 * KM] = sf[2 * K];
        }
        if (i < NG + B.bs[0] && z < NG + B.bs[2]) {
            temp = 0.;
            temp += 1.0 / 1.0 * si_[K] * gF2;
            temp += Dt / (1.0 * dx1) * (Eb(2 * K + si) - Eb(2 * K));
            temp += -Dt / (1.0 * dx3) * (Eb(1) - Eb(0));
            sf[K] = temp / gF2;
            if (sm) sm[KM] = sf[K];
            if (sm2) sm2[KM] = sf[K];
        }
    }
}

// Conserved update, source, face-B average, Noble cascade, floors and flags:
// the fixup kernel of GPU_fixup.h:5-425.  The step runs it as k_update (the
// updated conserved variables and the 2D inversion of each cell in one pass,
// U' kept in registers) and k_fallback (the 1D inversions of the few cells
// whose 2D inversion failed, so that they do not set k_update's register
// budget).  k_cons + k_invert are the same update in two kernels with U'
// passed through memory (Unew): k_cons forms the updated conserved variables
// (full = 0, predictor: U^n from pi is stored in U0, source from pi; full = 1,
// corrector: U from U0, source from pb) and k_invert inverts them with the
// same initial guess (pi or pb), applies the floors and sets the flags.  All
// forms give the same result bit for bit.

// Per-column metric values k_cons stages in LDS (every z-row of a block shares
// one column): gcov, gcon, gdet at CENT, the connection, and the face gdets the
// face-field average uses.
enum { MV_GCOV = 0, MV_GCON = 10, MV_G = 20, MV_CONN = 21, MV_GF1 = 61, MV_GF1P = 62, MV_GF2 = 63, MV_GF2P = 64,
       MV_GF3 = 65, MV_N = 66 };

__device__ __forceinline__ double metric_value(const KParams &P, const BlockDev &B, int v, int64_t col)
{
    if (v < MV_GCON) return P.gcov[mg10(B, CENT, v) + col];
    if (v < MV_G) return P.gcon[mg10(B, CENT, v - MV_GCON) + col];
    if (v == MV_G) return P.gdet[mg1(B, CENT) + col];
    if (v < MV_GF1) return P.conn[B.moff * 40 + (int64_t)(v - MV_CONN) * B.kcol + col];
    if (v == MV_GF1) return P.gdet[mg1(B, FACE1) + col];
    if (v == MV_GF1P) return P.gdet[mg1(B, FACE1) + col + B.D[1]];
    if (v == MV_GF2) return P.gdet[mg1(B, FACE2) + col];
    if (v == MV_GF2P) return P.gdet[mg1(B, FACE2) + col + 1];
    return P.gdet[mg1(B, FACE3) + col];
}

// Neumaier's compensated sum: s + c carries the running total without the
// rounding of each addition (used by the mixed step, whose FP32 flux
// differences are exact in FP64 and are added to U^n without loss).
__device__ __forceinline__ void comp_add(double &s, double &c, double t)
{
    const double u = s + t;
    if (fabs(s) >= fabs(t)) c += (s - u) + t;
    else c += (t - u) + s;
    s = u;
}

// One cell of the conserved update (cell g of block B, metric column values
// through mv(v), v one of MV_*): U^n (predictor: from pi, stored in U0;
// corrector: read from U0), minus Dt times the flux divergence, plus Dt times
// the source from the state (pi or pb), the face-field average into the B
// slots; returned in U.  pf returns the state the source was formed from (pi
// or pb at the cell), the inversion's initial guess.  f1l / f1h are the x1
// fluxes at the cell's lower and upper faces (the caller loads or carries
// them).  R is the storage type of the fluxes: double (FP64 step, the
// reference's arithmetic) or float (mixed step, compensated FP64 accumulation).
template <class R, class MV>
__device__ __forceinline__ void cons_cell(const KParams &P, const BlockDev &B, int full, double Dt, int64_t g,
                                          const double *__restrict__ pi, const double *__restrict__ pb,
                                          double *__restrict__ U0, const double *__restrict__ psf, const R *f1l,
                                          const R *f1h, const R *__restrict__ F2, const R *__restrict__ F3, MV mv,
                                          double (&U)[NPR], double (&pf)[NPR])
{
    const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2], sj = B.D[2];
    const int64_t base = NPR * B.off + g;
    Geom<double> G;
#pragma unroll
    for (int s = 0; s < 10; s++) {
        G.gcov[s] = mv(MV_GCOV + s);
        G.gcon[s] = mv(MV_GCON + s);
    }
    G.g = mv(MV_G);
    double dU[NPR];
    State<double> q;
    if (!full) {
        double d1, d2;
#pragma unroll
        for (int k = 0; k < NPR; k++) pf[k] = pi[base + k * K];
        get_state(pf, G, q);
        primtoflux(pf, q, 0, G, P.gam, U, &d1, &d2);
#pragma unroll
        for (int k = 0; k < NPR; k++) U0[base + k * K] = U[k];
    } else {
#pragma unroll
        for (int k = 0; k < NPR; k++) U[k] = U0[base + k * K];
#pragma unroll
        for (int k = 0; k < NPR; k++) pf[k] = pb[base + k * K];
        get_state(pf, G, q);
    }
    // fluxes kept in their storage type until used (half the registers in the mixed step)
    R f2l[NPR], f2h[NPR], f3l[NPR], f3h[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        const R *f2 = F2 + base + k * K, *f3 = F3 + base + k * K;
        f2l[k] = f2[0];
        f2h[k] = f2[sj];
        f3l[k] = f3[0];
        f3h[k] = f3[1];
    }
    source(pf, q, [&](int lam, int s) { return mv(MV_CONN + lam * 10 + s); }, G.g, P.gam, dU);
    if (sizeof(R) == sizeof(double)) {
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            U[k] -= Dt * (f1h[k] - f1l[k]) / (B.dx[0] * 1.0);
            U[k] -= Dt * (f2h[k] - f2l[k]) / (B.dx[1] * 1.0);
            U[k] -= Dt * (f3h[k] - f3l[k]) / (B.dx[2] * 1.0);
        }
#pragma unroll
        for (int k = 0; k < NPR; k++) U[k] += Dt * (dU[k]);
    } else {
        // Dt/dx once per direction (27 FP64 divisions per cell fewer):
        // the FP32 flux differences are exact in FP64, so each term
        // differs from the reference's Dt*dF/dx by FP64 roundings only
        const double c1 = Dt / B.dx[0], c2 = Dt / B.dx[1], c3 = Dt / B.dx[2];
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            double sum = U[k], cmp = 0.;
            comp_add(sum, cmp, -(c1 * ((double)f1h[k] - (double)f1l[k])));
            comp_add(sum, cmp, -(c2 * ((double)f2h[k] - (double)f2l[k])));
            comp_add(sum, cmp, -(c3 * ((double)f3h[k] - (double)f3l[k])));
            comp_add(sum, cmp, Dt * dU[k]);
            U[k] = sum + cmp;
        }
    }
    {
        const double *sp = psf + 3 * B.off + g;
        U[B1] = 0.0;
        U[B2] = 0.0;
        U[B1] += (sp[0] * mv(MV_GF1) + sp[si] * mv(MV_GF1P)) / (2.0 * 1.0);
        U[B2] += (sp[K] * mv(MV_GF2) + sp[K + sj] * mv(MV_GF2P)) / (2.0 * 1.0);
        U[B3] = (sp[2 * K] * mv(MV_GF3) + sp[2 * K + 1] * mv(MV_GF3)) / 2.0;
    }
}

// One thread per interior cell, tiles of a block's cells z fastest; the
// metric columns a tile touches are staged in LDS.
template <class R>
__global__ void __launch_bounds__(256) k_cons(KParams P, WorkList wl, int full, double Dt_val,
                                              const double *__restrict__ dt_ptr, double dt_mult,
                                              const double *__restrict__ pi, const double *__restrict__ pb,
                                              double *__restrict__ U0, double *__restrict__ Unew,
                                              const double *__restrict__ psf, const R *__restrict__ F1,
                                              const R *__restrict__ F2, const R *__restrict__ F3)
{
    extern __shared__ double smem[];
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const BlockDev &B = P.blk[it.x];
        const int bs2 = B.bs[1], bs3 = B.bs[2];
        const int nint = B.bs[0] * bs2 * bs3;
        const int t0 = it.y * blockDim.x;
        const int t = t0 + threadIdx.x;
        // stage the tile's columns
        const int q0 = t0 / bs3, q1 = mymin(t0 + (int)blockDim.x - 1, nint - 1) / bs3;
        const int ncw = q1 - q0 + 1;
        for (int e = threadIdx.x; e < MV_N * ncw; e += blockDim.x) {
            const int v = e / ncw, qq = e - v * ncw, q = q0 + qq;
            const int64_t col = (int64_t)(q / bs2 + NG) * B.D[1] + (q % bs2 + NG);
            smem[e] = metric_value(P, B, v, col);
        }
        __syncthreads();
        if (t < nint) {
            int a, b, c;
            region_index(t, bs2, bs3, a, b, c);
            const int i = a + NG, j = b + NG, z = c + NG;
            const double *mvp = smem + (t / bs3 - q0);
            const double Dt = block_Dt(P, B, Dt_val, dt_ptr, dt_mult);
            const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2];
            const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
            const int64_t base = NPR * B.off + g;
            R f1l[NPR], f1h[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                const R *f1 = F1 + base + k * K;
                f1l[k] = f1[0];
                f1h[k] = f1[si];
            }
            double U[NPR], pf[NPR];
            cons_cell<R>(P, B, full, Dt, g, pi, pb, U0, psf, f1l, f1h, F2, F3, [&](int v) { return mvp[v * ncw]; },
                         U, pf);
#pragma unroll
            for (int k = 0; k < NPR; k++) Unew[base + k * K] = U[k];
        }
        __syncthreads();
    }
}

// The same update marching along x1: one thread per (j, z) of a block's
// interior x2-x3 plane walks i upward, so each x1 face flux is read once and
// carried from one cell to the next in registers (the per-cell form reads
// every interior x1 face twice, the second time from HBM unless the tile of
// the next x1 plane ran recently on the same XCD).  At each i the workgroup
// stages the metric of the columns its threads touch in LDS, as k_cons does
// per tile.  Grid: x over the plane in RPLANE tiles, y over the blocks.
constexpr int RPLANE = 128;
template <class R>
__global__ void __launch_bounds__(RPLANE) __attribute__((amdgpu_waves_per_eu(3)))
k_cons_march(KParams P, int full, double Dt_val, const double *__restrict__ dt_ptr, double dt_mult,
             const double *__restrict__ pi, const double *__restrict__ pb, double *__restrict__ U0,
             double *__restrict__ Unew, const double *__restrict__ psf, const R *__restrict__ F1,
             const R *__restrict__ F2, const R *__restrict__ F3)
{
    extern __shared__ double smem[];
    const BlockDev &B = P.blk[blockIdx.y];
    const int bs2 = B.bs[1], bs3 = B.bs[2], nplane = bs2 * bs3;
    const int t0 = blockIdx.x * RPLANE, t = t0 + threadIdx.x;
    if (t0 >= nplane) return;  // the whole workgroup leaves together
    const bool act = t < nplane;
    const int q0 = t0 / bs3, q1 = mymin(t0 + RPLANE - 1, nplane - 1) / bs3, ncw = q1 - q0 + 1;
    const int j = (act ? t : t0) / bs3 + NG, z = (act ? t : t0) % bs3 + NG;
    const double *mvp = smem + (j - NG - q0);
    const double Dt = block_Dt(P, B, Dt_val, dt_ptr, dt_mult);
    const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2];
    R fl[NPR], fh[NPR];
    if (act) {
        const int64_t base = NPR * B.off + ((int64_t)NG * B.D[1] + j) * B.D[2] + z;
#pragma unroll
        for (int k = 0; k < NPR; k++) fl[k] = F1[base + k * K];
    }
#pragma unroll 1
    for (int i = NG; i < NG + B.bs[0]; i++) {
        __syncthreads();  // the previous plane's metric is no longer read
        for (int e = threadIdx.x; e < MV_N * ncw; e += RPLANE) {
            const int v = e / ncw, qq = e - v * ncw;
            smem[e] = metric_value(P, B, v, (int64_t)i * B.D[1] + (q0 + qq + NG));
        }
        __syncthreads();
        if (act) {
            const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
            const int64_t base = NPR * B.off + g;
#pragma unroll
            for (int k = 0; k < NPR; k++) fh[k] = F1[base + k * K + si];
            double U[NPR], pf[NPR];
            cons_cell<R>(P, B, full, Dt, g, pi, pb, U0, psf, fl, fh, F2, F3, [&](int v) { return mvp[v * ncw]; },
                         U, pf);
#pragma unroll
            for (int k = 0; k < NPR; k++) Unew[base + k * K] = U[k];
#pragma unroll
            for (int k = 0; k < NPR; k++) fl[k] = fh[k];
        }
    }
}

void launch_cons_march(const KParams &P, int nblk, int max_plane, int min_bs3, int full, double Dt_val,
                       const double *dt_ptr, double dt_mult, const double *pi, const double *pb, double *U0,
                       double *Unew, const double *psf, const void *F1, const void *F2, const void *F3, int mixed,
                       hipStream_t s)
{
    if (nblk <= 0 || min_bs3 <= 0) return;
    const dim3 grid((unsigned)((max_plane + RPLANE - 1) / RPLANE), (unsigned)nblk);
    const size_t lds = sizeof(double) * MV_N * (size_t)((RPLANE - 1) / min_bs3 + 2);
    if (mixed)
        hipLaunchKernelGGL((k_cons_march<float>), grid, dim3(RPLANE), lds, s, P, full, Dt_val, dt_ptr, dt_mult, pi,
                           pb, U0, Unew, psf, (const float *)F1, (const float *)F2, (const float *)F3);
    else
        hipLaunchKernelGGL((k_cons_march<double>), grid, dim3(RPLANE), lds, s, P, full, Dt_val, dt_ptr, dt_mult, pi,
                           pb, U0, Unew, psf, (const double *)F1, (const double *)F2, (const double *)F3);
}

// Adds the number of active lanes with `pred` set to *cnt with one atomic per
// wavefront.  All lanes of a wavefront work on the same block (a tile never
// spans two blocks), so they share the counter.  Most atmosphere cells sit on
// the floor, so a per-cell atomic would issue millions of device-scope atomics
// per half step.
__device__ __forceinline__ void wave_count(int *cnt, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const unsigned long long act = __ballot(1);
    if (__lane_id() == __ffsll((unsigned long long)act) - 1) atomicAdd(cnt, __popcll(m));
}

// Floors (this library's own specification: rho >= rho_floor r^-1.5, u >= u_floor
// r^-2.5, flag -333 as GPU_fixup.h:399-415) and the result of one cell.  A
// cell among the NG lowest or highest interior cells of its z-row is also
// written into the x3 ghost cell of the periodic neighbour that the ghost fill
// would copy it to (the neighbour's upper ghosts from this block's lowest
// cells, its lower ghosts from the highest), so that the ghost fill after the
// update only has the rows beyond the block's x1 and x2 edges left along x3.
template <bool TILED = true>
__device__ __forceinline__ void finish_cell(const KParams &P, const BlockDev &B, int n, int i, int j, int64_t col,
                                            int64_t g, int z, int64_t base, int flag, double *pf,
                                            double *__restrict__ pf_out, int *__restrict__ pflag,
                                            int *__restrict__ failcnt)
{
    const int64_t K = B.ks;
    const double r = P.rcen[B.moff + col], sr = sqrt(r);
    int hit = 0;
    if (P.rho_floor > 0.) {
        const double rhoflr = P.rho_floor / (r * sr);
        if (pf[RHO] < rhoflr) { pf[RHO] = rhoflr; hit = 1; }
    }
    if (P.u_floor > 0.) {
        const double uflr = P.u_floor / (r * r * sr);
        if (pf[UU] < uflr) { pf[UU] = uflr; hit = 1; }
    }
    if (hit) flag = -333;
    wave_count(&failcnt[n * 4 + 3], hit != 0);
    pflag[B.off + g] = flag;
#pragma unroll
    for (int k = 0; k < NPR; k++) pf_out[base + k * K] = pf[k];
    const int bs3 = B.bs[2];
    PushTarget T;
    if ((z < 2 * NG || z >= bs3) && x3_push_target<TILED>(P, B, i, j, z, g, z < 2 * NG ? 0 : 1, T)) {
        pflag[T.off + T.cell] = flag;
        // a time-level cycle writes the receiver's ghost into both of its arrays
        const bool both = TILED && P.las_s >= 0;
        for (int pass = 0; pass < (both ? 2 : 1); pass++) {
            double *dst = both ? (pass ? P.xph : P.xp) : pf_out;
#pragma unroll
            for (int k = 0; k < NPR; k++) dst[NPR * T.off + T.cell + k * T.ks] = pf[k];
        }
    }
}

// Marks a cell whose 2D inversion failed until k_fallback has run the 1D
// inversions on it (never a return code of the inversions).
constexpr int kFallbackPending = 0x40000000;

// 2D inversion of every interior cell.  A failed cell keeps the primitives the
// 2D solver left in pf_out and is marked for k_fallback, so that the rare
// 1D inversions do not set this kernel's register budget.
__global__ void __launch_bounds__(256) k_invert(KParams P, WorkList wl, int full, const double *__restrict__ pi,
                                                const double *__restrict__ pb, const double *__restrict__ Unew,
                                                double *__restrict__ pf_out, int *__restrict__ pflag,
                                                int *__restrict__ failcnt)
{
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const int n = it.x;
        const BlockDev &B = P.blk[n];
        const int t = it.y * blockDim.x + threadIdx.x;
        if (t >= B.bs[0] * B.bs[1] * B.bs[2]) continue;
        int a, b, c;
        region_index(t, B.bs[1], B.bs[2], a, b, c);
        const int i = a + NG, j = b + NG, z = c + NG;
        const int64_t K = B.ks;
        const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z, col = (int64_t)i * B.D[1] + j;
        const int64_t base = NPR * B.off + g;
        const Geom<double> G = load_geom(P, B, CENT, col);
        const double *guess = full ? pb : pi;
        double pf[NPR], U[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            U[k] = Unew[base + k * K];
            pf[k] = guess[base + k * K];
        }
        const int flag = utoprim(0, U, G.gcov, G.gcon, G.g, P.gam, pf);
        if (flag == 0) {
            finish_cell(P, B, n, i, j, col, g, z, base, 0, pf, pf_out, pflag, failcnt);
            continue;
        }
        atomicAdd(&failcnt[n * 4 + 0], 1);
        pflag[B.off + g] = kFallbackPending;
#pragma unroll
        for (int k = 0; k < NPR; k++) pf_out[base + k * K] = pf[k];
    }
}

// The conserved update and the 2D inversion of each interior cell in one
// pass (fixup, GPU_fixup.h:5-425, without the 1D fallbacks): k_cons's update
// with the column metric staged in LDS, then k_invert's inversion on the
// updated U in registers, from the state the source was formed from (pi or pb)
// as the guess.  U' and the guess never go through memory, except for a cell
// whose 2D inversion fails: its U' goes to Unew, its flag is kFallbackPending,
// and k_fallback finishes it.  Same arithmetic as k_cons + k_invert, bit for
// bit.  Two waves per SIMD (<= 256 VGPRs): held at three (168 VGPRs) it
// spills 43-65 registers and runs 1.2x slower.
template <class R, bool TILED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_update(KParams P, WorkList wl, int full, double Dt_val, const double *__restrict__ dt_ptr, double dt_mult,
         const double *__restrict__ pi, const double *__restrict__ pb, double *__restrict__ pf_out,
         double *__restrict__ U0, double *__restrict__ Unew, const double *__restrict__ psf, const R *__restrict__ F1,
         const R *__restrict__ F2, const R *__restrict__ F3, int *__restrict__ pflag, int *__restrict__ failcnt)
{
    extern __shared__ double smem[];
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const int n = it.x;
        const BlockDev &B = P.blk[n];
        const int bs2 = B.bs[1], bs3 = B.bs[2];
        const int nint = B.bs[0] * bs2 * bs3;
        const int t0 = it.y * blockDim.x;
        const int t = t0 + threadIdx.x;
        const int q0 = t0 / bs3, q1 = mymin(t0 + (int)blockDim.x - 1, nint - 1) / bs3;
        const int ncw = q1 - q0 + 1;
        for (int e = threadIdx.x; e < MV_N * ncw; e += blockDim.x) {
            const int v = e / ncw, qq = e - v * ncw, q = q0 + qq;
            const int64_t col = (int64_t)(q / bs2 + NG) * B.D[1] + (q % bs2 + NG);
            smem[e] = metric_value(P, B, v, col);
        }
        __syncthreads();
        if (t < nint) {
            int a, b, c;
            region_index(t, bs2, bs3, a, b, c);
            const int i = a + NG, j = b + NG, z = c + NG;
            const double *mvp = smem + (t / bs3 - q0);
            const double Dt = block_Dt(P, B, Dt_val, dt_ptr, dt_mult);
            const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2];
            const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z, col = (int64_t)i * B.D[1] + j;
            const int64_t base = NPR * B.off + g;
            double U[NPR], pf[NPR];
            {
                R f1l[NPR], f1h[NPR];
#pragma unroll
                for (int k = 0; k < NPR; k++) {
                    const R *f1 = F1 + base + k * K;
                    f1l[k] = f1[0];
                    f1h[k] = f1[si];
                }
                cons_cell<R>(P, B, full, Dt, g, pi, pb, U0, psf, f1l, f1h, F2, F3,
                             [&](int v) { return mvp[v * ncw]; }, U, pf);
            }
            Geom<double> G;
#pragma unroll
            for (int s = 0; s < 10; s++) {
                G.gcov[s] = mvp[(MV_GCOV + s) * ncw];
                G.gcon[s] = mvp[(MV_GCON + s) * ncw];
            }
            G.g = mvp[MV_G * ncw];
            const int flag = utoprim(0, U, G.gcov, G.gcon, G.g, P.gam, pf);
            if (flag == 0) {
                finish_cell<TILED>(P, B, n, i, j, col, g, z, base, 0, pf, pf_out, pflag, failcnt);
            } else {
                atomicAdd(&failcnt[n * 4 + 0], 1);
                pflag[B.off + g] = kFallbackPending;
#pragma unroll
                for (int k = 0; k < NPR; k++) {
                    pf_out[base + k * K] = pf[k];
                    Unew[base + k * K] = U[k];
                }
            }
        }
        __syncthreads();
    }
}

// The rest of the cascade for the cells k_invert marked: entropy-based 1D
// (GPU_utoprim_1dvsq2fix1.h), then 1D (GPU_utoprim_1dfix1.h), then floors.
__global__ void __launch_bounds__(256) k_fallback(KParams P, WorkList wl, const double *__restrict__ Unew,
                                                  double *__restrict__ pf_out, int *__restrict__ pflag,
                                                  int *__restrict__ failcnt)
{
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const int n = it.x;
        const BlockDev &B = P.blk[n];
        const int t = it.y * blockDim.x + threadIdx.x;
        if (t >= B.bs[0] * B.bs[1] * B.bs[2]) continue;
        int a, b, c;
        region_index(t, B.bs[1], B.bs[2], a, b, c);
        const int i = a + NG, j = b + NG, z = c + NG;
        const int64_t K = B.ks;
        const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z, col = (int64_t)i * B.D[1] + j;
        if (pflag[B.off + g] != kFallbackPending) continue;
        const int64_t base = NPR * B.off + g;
        const Geom<double> G = load_geom(P, B, CENT, col);
        double pf[NPR], U[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            U[k] = Unew[base + k * K];
            pf[k] = pf_out[base + k * K];
        }
        int flag = utoprim(1, U, G.gcov, G.gcon, G.g, P.gam, pf);
        if (flag) {
            atomicAdd(&failcnt[n * 4 + 1], 1);
            flag = utoprim(2, U, G.gcov, G.gcon, G.g, P.gam, pf);
            if (flag) atomicAdd(&failcnt[n * 4 + 2], 1);
        }
        finish_cell(P, B, n, i, j, col, g, z, base, flag, pf, pf_out, pflag, failcnt);
    }
}

// Ghost fill of one direction for the listed blocks: same-level copy, radial
// outflow with inflow check, or the polar mirror from the pi-shifted partner.
// Stage order x1, x2, x3, each over the full extent of the others.  Faces of
// kind FK_AMR are left to the multi-level exchange.
//
// Thread mapping: for x1 and x2 the fastest thread index runs along z, so
// every access of a wavefront is contiguous.  For x3 the ghosts are the first
// and last NG cells of each z-row; a group of 2*NG consecutive threads writes
// the upper ghosts of row r-1 and the lower ghosts of row r, which are adjacent
// in memory, so each group stores one contiguous 48-byte segment.
template <int DIR>
__global__ void __launch_bounds__(256) k_bound(KParams P, WorkList wl, double *p, double *ps, int *pflag,
                                               int outflow_only, int x3_rows_done)
{
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const int n = it.x;
        const BlockDev &B = P.blk[n];
        const int d1 = B.D[0], d2 = B.D[1], d3 = B.D[2];
        const int t = it.y * blockDim.x + threadIdx.x;
        int aa, bb, gg, side;
        if (DIR == 3) {
            const int nrow = d1 * d2;
            if (t >= 2 * NG * (nrow + 1)) continue;
            const int g6 = t % (2 * NG), r = t / (2 * NG);
            side = (g6 < NG) ? 1 : 0;
            gg = side ? g6 : g6 - NG;
            const int row = side ? r - 1 : r;
            if (row < 0 || row >= nrow) continue;
            aa = row / d2;
            bb = row % d2;
        } else {
            const int ta = (DIR == 1) ? d2 : d1, tb = d3;
            if (t >= 2 * NG * ta * tb) continue;
            bb = t % tb;
            aa = (t / tb) % ta;
            gg = (t / (tb * ta)) % NG;
            side = t / (tb * ta * NG);
        }
        const int face = 2 * (DIR - 1) + side;
        const int kind = B.kind[face];
        if (kind == FK_AMR) continue;
        // rows inside the block's x1 and x2 range whose x3 ghosts the update
        // and face-field kernels already wrote from the periodic neighbour
        if (DIR == 3 && x3_rows_done && aa >= NG && aa < NG + B.bs[0] && bb >= NG && bb < NG + B.bs[1] &&
            x3_row_pushed(P, B, n, face, aa, bb))
            continue;
        if (outflow_only) {
            if (kind != FK_OUTFLOW) continue;
            // rows beyond a polar face keep the mirror image of the partner's ghosts
            if (((B.pole & 1) && aa < NG) || ((B.pole & 2) && aa >= NG + B.bs[1])) continue;
        }
        const int bsd = B.bs[DIR - 1];
        // ghost and source index along DIR
        const int gidx = side ? NG + bsd + gg : gg;
        int sidx, src = B.nbr[face];
        if (kind == FK_COPY) sidx = side ? NG + gg : bsd + gg;
        else if (kind == FK_OUTFLOW) { sidx = side ? NG + bsd - 1 : NG; src = n; }
        else sidx = side ? NG + bsd - 1 - gg : 2 * NG - 1 - gg;
        int i, j, z, is, js, zs;
        if (DIR == 1) { i = gidx; j = aa; z = bb; is = sidx; js = j; zs = z; }
        else if (DIR == 2) { i = aa; j = gidx; z = bb; is = i; js = sidx; zs = z; }
        else { i = aa; j = bb; z = gidx; is = i; js = j; zs = sidx; }
        // a block that is its own pole partner (one block around the axis): shift by half a turn
        if (kind == FK_POLE && src == n) zs = ((z - NG + B.bs[2] / 2) % B.bs[2] + B.bs[2]) % B.bs[2] + NG;
        const BlockDev &S = P.blk[src];
        const int64_t K = B.ks, KS = S.ks;
        const int64_t cg = ((int64_t)i * d2 + j) * d3 + z, cs = ((int64_t)is * S.D[1] + js) * S.D[2] + zs;
        double *pd = p + NPR * B.off;
        const double *psrc = p + NPR * S.off;
        double *sd = ps + 3 * B.off;
        const double *ssrc = ps + 3 * S.off;
        double pr[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) pr[k] = psrc[k * KS + cs];
        if (kind == FK_COPY) {
#pragma unroll
            for (int k = 0; k < NPR; k++) pd[k * K + cg] = pr[k];
            sd[cg] = ssrc[cs];
            sd[K + cg] = ssrc[KS + cs];
            sd[2 * K + cg] = ssrc[2 * KS + cs];
            if (pflag) pflag[B.off + cg] = pflag[S.off + cs];
        } else if (kind == FK_OUTFLOW) {
            const Geom<double> G = load_geom(P, B, CENT, (int64_t)i * d2 + j);
            inflow_check(pr, G, side ? 1 : 0);
#pragma unroll
            for (int k = 0; k < NPR; k++) pd[k * K + cg] = pr[k];
            sd[K + cg] = ssrc[KS + cs];
            sd[2 * K + cg] = ssrc[2 * KS + cs];
            if (pflag) pflag[B.off + cg] = pflag[B.off + cs];
        } else {
            pr[U2] *= -1.0;
            pr[U3] *= -1.0;
            pr[B2] *= -1.0;
            pr[B3] *= -1.0;
#pragma unroll
            for (int k = 0; k < NPR; k++) pd[k * K + cg] = pr[k];
            sd[cg] = ssrc[cs];
            sd[2 * K + cg] = -ssrc[2 * KS + cs];
            if (pflag) pflag[B.off + cg] = pflag[S.off + cs];
        }
    }
}

// ---------------------------------------------------------------------------
// Multi-level ghost exchange (level ratio 2).  Each entry names one ghost cell
// and its source: a same-level cell (copy) or the 2x2x2 finer cells that tile
// it (restriction, the mean of packsendaverage1, GPU_program2.cu:154-197); the
// face fields of a restriction are the means of the four finer faces that tile
// each coarse face, the component normal to the exchange direction weighted by
// the finer faces' gdet and divided by the coarse face's (the flux mean that
// packsendaverage1-3 pack, GPU_program2.cu:176-185, :231-236, :281-285, over
// the receiver's gdet, as the disabled store of unpackreceive1 does,
// :376-380).  A prolongation entry names the coarse cell that contains the
// ghost and its six neighbours; the ghost gets the coarse value plus a quarter
// of each direction's MC slope (slope_lim_2, GPU_program2.cu:10-45), signed by
// the side of the coarse cell it lies on (unpackreceivecoarse1,
// GPU_program2.cu:615-803).  A face field along direction a carries no slope
// along a; a fine face that halves a coarse cell starts from the coarse
// cell-centred field, one on a coarse face from that face's value.
//
// No entry of the exchange along d writes the d-component of the face field
// at the block's own upper boundary face (ghost index NG+bs along d, interior
// across): the block's constrained-transport update owns that face, as the
// reference's receive never stores the normal component (unpackreceive1,
// GPU_program2.cu:376-380; unpackreceivecoarse1 stores only the tangential
// ones, :779-796), so the divergence of every interior cell's field stays at
// its roundoff.  The same component in ghost rows beyond another direction's
// edge is written (the fluxes of the transverse ghost faces read it).

// Whether ghost cell (i, j, z) of block B holds B's own upper boundary face
// along direction d0 (0..2): index NG+bs along d0, interior across.
__device__ __forceinline__ bool own_upper_face(const BlockDev &B, int d0, const int loc[3])
{
    bool own = loc[d0] == NG + B.bs[d0];
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (k != d0) own = own && loc[k] >= NG && loc[k] < NG + B.bs[k];
    return own;
}

// value of source s (cell or 2x2x2 mean) for the 9 primitives and 3 face fields
__device__ __forceinline__ void amr_source(const KParams &P, const double *__restrict__ p,
                                           const double *__restrict__ ps, int blk, int code, double v[NPR + 3])
{
    const BlockDev &S = P.blk[blk];
    const int kind = code >> 28;
    const int64_t c = code & 0x0FFFFFFF;
    const int64_t K = S.ks;
    const double *pp = p + NPR * S.off + c, *sp = ps + 3 * S.off + c;
    if (kind == AS_CELL) {
#pragma unroll
        for (int k = 0; k < NPR; k++) v[k] = pp[k * K];
#pragma unroll
        for (int a = 0; a < 3; a++) v[NPR + a] = sp[a * K];
        return;
    }
    const int64_t si = (int64_t)S.D[1] * S.D[2], sj = S.D[2];
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        const double *q = pp + k * K;
        v[k] = 0.125 * (q[0] + q[1] + q[sj] + q[sj + 1] + q[si] + q[si + 1] + q[si + sj] + q[si + sj + 1]);
    }
    v[NPR + 0] = 0.25 * (sp[0] + sp[1] + sp[sj] + sp[sj + 1]);
    v[NPR + 1] = 0.25 * (sp[K] + sp[K + 1] + sp[K + si] + sp[K + si + 1]);
    v[NPR + 2] = 0.25 * (sp[2 * K] + sp[2 * K + sj] + sp[2 * K + si] + sp[2 * K + si + sj]);
}

// gdet-weighted mean of the four finer d0-faces of the 2x2x2 cells from cell c
// of block S (term order of packsendaverage1-3), over the coarse face's gdet gc
__device__ __forceinline__ double amr_flux_mean(const KParams &P, const BlockDev &S, const double *__restrict__ ps,
                                                int d0, int64_t c, double gc)
{
    const int64_t K = S.ks, si = (int64_t)S.D[1] * S.D[2], sj = S.D[2];
    const int64_t col = c / S.D[2];
    const double *sp = ps + 3 * S.off + (int64_t)d0 * K + c;
    const double *gf = P.gdet + mg1(S, d0 == 0 ? FACE1 : d0 == 1 ? FACE2 : FACE3) + col;
    double m;
    if (d0 == 0) m = 0.25 * (sp[0] * gf[0] + sp[1] * gf[0] + sp[sj] * gf[1] + sp[sj + 1] * gf[1]);
    else if (d0 == 1) m = 0.25 * (sp[0] * gf[0] + sp[1] * gf[0] + sp[si] * gf[S.D[1]] + sp[si + 1] * gf[S.D[1]]);
    else m = 0.25 * (sp[0] * gf[0] + sp[sj] * gf[1] + sp[si] * gf[S.D[1]] + sp[si + sj] * gf[S.D[1] + 1]);
    return m / gc;
}

// dir: the exchange direction of the entries (1..3)
// Exchange arrays of a local-time-stepping cycle (P.las_s >= 0; ph non-null):
// a ghost is filled only when its block and its source block both step in
// this sub-iteration; the source is read from the arrays of its own phase
// (its predictor's ph/psh, or p/ps after its corrector; a block that does not
// step here, a slope neighbour of a prolongation, from p/ps), and the ghost is
// written into both of its block's array pairs, as the reference's receive
// sets p and ph (unpackreceive1, GPU_program2.cu:345-388).
__device__ __forceinline__ void las_arrays(const KParams &P, int blk, double *p, double *ps, double *ph,
                                           double *psh, const double *&sp, const double *&sps)
{
    const bool half = ph && las_phase(P.blk[blk].tl, P.las_s) == LAS_HALF;
    sp = half ? ph : p;
    sps = half ? psh : ps;
}
__device__ __forceinline__ bool las_steps(const KParams &P, int blk)
{
    return las_phase(P.blk[blk].tl, P.las_s) != LAS_IDLE;
}

__global__ void __launch_bounds__(256) k_amr_copy(KParams P, const AmrEntry *__restrict__ e, int64_t n, int dir,
                                                  double *p, double *ps, double *ph, double *psh, int *pflag)
{
    const int d0 = dir - 1;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const AmrEntry E = e[t];
        if (ph && (!las_steps(P, E.dblk) || !las_steps(P, E.sblk))) continue;
        const double *sp, *sps;
        las_arrays(P, E.sblk, p, ps, ph, psh, sp, sps);
        double v[NPR + 3];
        amr_source(P, sp, sps, E.sblk, E.scode, v);
        const BlockDev &B = P.blk[E.dblk];
        const BlockDev &S = P.blk[E.sblk];
        int loc[3];
        region_index(E.dcell, B.D[1], B.D[2], loc[0], loc[1], loc[2]);
        if ((E.scode >> 28) == AS_AVG8) {
            const double gc = P.gdet[mg1(B, d0 == 0 ? FACE1 : d0 == 1 ? FACE2 : FACE3) + (int64_t)loc[0] * B.D[1] + loc[1]];
            v[NPR + d0] = amr_flux_mean(P, S, sps, d0, E.scode & 0x0FFFFFFF, gc);
        }
        const bool own = own_upper_face(B, d0, loc);
        const int64_t K = B.ks;
        for (int pass = 0; pass < (ph ? 2 : 1); pass++) {
            double *pd = (pass ? ph : p) + NPR * B.off + E.dcell, *sd = (pass ? psh : ps) + 3 * B.off + E.dcell;
#pragma unroll
            for (int k = 0; k < NPR; k++) pd[k * K] = v[k];
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (!(a == d0 && own)) sd[a * K] = v[NPR + a];
        }
        if (pflag) pflag[B.off + E.dcell] = pflag[S.off + (E.scode & 0x0FFFFFFF)];
    }
}

__global__ void __launch_bounds__(256) k_amr_prolong(KParams P, const AmrProlong *__restrict__ e, int64_t n,
                                                     int dir, double *p, double *ps, double *ph, double *psh,
                                                     int *pflag)
{
    const int d0 = dir - 1;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const AmrProlong E = e[t];
        if (ph && (!las_steps(P, E.dblk) || !las_steps(P, E.cblk))) continue;
        const BlockDev &B = P.blk[E.dblk];
        const BlockDev &C = P.blk[E.cblk];
        // side of the coarse cell along each direction (global fine index parity)
        int loc[3];
        region_index(E.dcell, B.D[1], B.D[2], loc[0], loc[1], loc[2]);
        const bool own = own_upper_face(B, d0, loc);
        double sgn[3];
#pragma unroll
        for (int d = 0; d < 3; d++) sgn[d] = ((B.o[d] + loc[d] - NG) & 1) ? 1.0 : -1.0;
        double c0[NPR + 3], dq[3][NPR + 3];
        const double *sp, *sps;
        las_arrays(P, E.cblk, p, ps, ph, psh, sp, sps);
        amr_source(P, sp, sps, E.cblk, (AS_CELL << 28) | E.ccell, c0);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int cm = E.nb[2 * d][1], cp = E.nb[2 * d + 1][1];
            if ((cm >> 28) == AS_NONE || (cp >> 28) == AS_NONE) {
#pragma unroll
                for (int k = 0; k < NPR + 3; k++) dq[d][k] = 0.;
                continue;
            }
            double vm[NPR + 3], vp[NPR + 3];
            las_arrays(P, E.nb[2 * d][0], p, ps, ph, psh, sp, sps);
            amr_source(P, sp, sps, E.nb[2 * d][0], cm, vm);
            las_arrays(P, E.nb[2 * d + 1][0], p, ps, ph, psh, sp, sps);
            amr_source(P, sp, sps, E.nb[2 * d + 1][0], cp, vp);
#pragma unroll
            for (int k = 0; k < NPR + 3; k++) dq[d][k] = mc_slope(vm[k], c0[k], vp[k]);
        }
        const int64_t K = B.ks;
        double pv[NPR], fv[3];
#pragma unroll
        for (int k = 0; k < NPR; k++)
            pv[k] = c0[k] + 0.25 * sgn[0] * dq[0][k] + 0.25 * sgn[1] * dq[1][k] + 0.25 * sgn[2] * dq[2][k];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // odd fine face along a: halves the coarse cell -> cell-centred field
            const int kk = (sgn[a] > 0.) ? B1 + a : NPR + a;
            double v = c0[kk];
#pragma unroll
            for (int d = 0; d < 3; d++)
                if (d != a) v = v + 0.25 * sgn[d] * dq[d][kk];
            fv[a] = v;
        }
        for (int pass = 0; pass < (ph ? 2 : 1); pass++) {
            double *pd = (pass ? ph : p) + NPR * B.off + E.dcell, *sd = (pass ? psh : ps) + 3 * B.off + E.dcell;
#pragma unroll
            for (int k = 0; k < NPR; k++) pd[k * K] = pv[k];
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (!(a == d0 && own)) sd[a * K] = fv[a];
        }
        if (pflag) pflag[B.off + E.dcell] = pflag[C.off + E.ccell];
    }
}
// Flux correction at level jumps (unpackreceive1flux-3flux with calc_corr = 1,
// GPU_program2.cu:1258-1617, fed by packsendfluxaverage1-3, :1618-1771): the
// coarse face takes ((0.25 dt) (f00 + f01 + f10 + f11)) / dt of the four finer
// faces that tile it, f_ab at transverse offsets a (slower) and b.  One thread
// per coarse face; the corrector's fluxes only (the reference sends them from
// blocks doing their full step).  Coarse faces written and finer faces read
// are distinct faces, so the entries of one direction need no ordering.
template <class R>
__global__ void __launch_bounds__(256) k_amr_reflux(KParams P, const AmrReflux *__restrict__ e, int64_t n, int dir,
                                                    R *__restrict__ F, double dt_val, const double *__restrict__ dt_ptr)
{
    const double dt = dt_ptr ? dt_ptr[0] : dt_val;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const AmrReflux x = e[t];
        const BlockDev &C = P.blk[x.cb], &Fb = P.blk[x.fb];
        const int64_t sa = dir == 1 ? Fb.D[2] : (int64_t)Fb.D[1] * Fb.D[2], sb = dir == 3 ? Fb.D[2] : 1;
        const R *f = F + NPR * Fb.off + x.ff;
        R *c = F + NPR * C.off + x.cf;
#pragma unroll
        for (int k = 0; k < NPR; k++) {
            const R *q = f + (int64_t)k * Fb.ks;
            const double sum = (double)q[0] + (double)q[sb] + (double)q[sa] + (double)q[sa + sb];
            c[(int64_t)k * C.ks] = (R)(((0.25 * dt) * sum) / dt);
        }
    }
}


// EMF correction at level jumps (the receiving side of E_rec1-3 with calc_corr
// = 1, unpackreceive1E, GPU_program2.cu:2106-2230, fed by packsendEaverage1-3,
// :1875-2105): the coarse edge takes ((0.5 dt) (e0 + e1)) / dt of its two finer
// edges, in the corrector after the polar average.  Coarse edges written and
// finer edges read are distinct (a source is never corrected itself).
template <class R>
__global__ void __launch_bounds__(256) k_amr_emf(KParams P, const AmrEmf *__restrict__ e, int64_t n,
                                                 R *__restrict__ E, double dt_val, const double *__restrict__ dt_ptr)
{
    const double dt = dt_ptr ? dt_ptr[0] : dt_val;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const AmrEmf x = e[t];
        const int comp = x.cc >> 28;
        const BlockDev &C = P.blk[x.cb], &Fb = P.blk[x.fb];
        const int64_t st = comp == 0 ? (int64_t)Fb.D[1] * Fb.D[2] : comp == 1 ? (int64_t)Fb.D[2] : 1;
        const R *f = E + 3 * Fb.off + (int64_t)comp * Fb.ks + x.fc;
        const double sum = (double)f[0] + (double)f[st];
        E[3 * C.off + (int64_t)comp * C.ks + (x.cc & 0x0FFFFFFF)] = (R)(((0.5 * dt) * sum) / dt);
    }
}

// Corrections of a local-adaptive-time-stepping cycle (flux_send / flux_rec
// and E_send / E_rec with calc_corr = 1 and PRESTEP = PRESTEP2 = 0,
// boundcomF.c:5-1315, boundcomE.c; packsendfluxaverage1-3 and
// unpackreceive1flux-3flux, GPU_program2.cu:1258-1771, the factor and
// first-step rule of GPU_boundcomF.cu:7-8 and :109-110).  One thread per entry
// and, for fluxes, all NPR components.  Stage 0 (senders at their corrector):
// acc = c sum on the first corrector after the receiver's last (s % 2tl_r ==
// 2tl_s - 1), acc += c sum after, with c = w (dt tl_s), w = 1/4 (four finer
// faces), 1/2 (two finer edges) or 1, and sum the sender's values in the term
// order of k_amr_reflux / k_amr_emf.  Stage 1 (receivers at their corrector):
// value = acc / (dt tl_r).
template <class R>
__global__ void __launch_bounds__(256) k_las_corr(KParams P, const LasCorr *__restrict__ e, int64_t n, int dir,
                                                  int stage, R *__restrict__ FE, double *__restrict__ acc,
                                                  const double *__restrict__ dt_ptr)
{
    const double dt = dt_ptr[0];
    const int s = P.las_s;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const LasCorr x = e[t];
        const BlockDev &Rb = P.blk[x.rb], &Sb = P.blk[x.sb];
        const int tr = Rb.tl > 0 ? Rb.tl : 1, ts = Sb.tl > 0 ? Sb.tl : 1;
        const bool mean = (x.scell & LAS_MEAN) != 0;
        const int sc = x.scell & 0x3FFFFFFF;
        if (dir == 0) {  // EMF: component in bits 28-29 of rcell
            const int comp = x.rcell >> 28, rc = x.rcell & 0x0FFFFFFF;
            if (stage == 0) {
                if (las_phase(ts, s) != LAS_FULL) continue;
                const R *f = FE + 3 * Sb.off + (int64_t)comp * Sb.ks + sc;
                double sum, c;
                if (mean) {
                    const int64_t st = comp == 0 ? (int64_t)Sb.D[1] * Sb.D[2] : comp == 1 ? (int64_t)Sb.D[2] : 1;
                    sum = (double)f[0] + (double)f[st];
                    c = 0.5 * (dt * (double)ts);
                } else {
                    sum = (double)f[0];
                    c = 1.0 * (dt * (double)ts);
                }
                const bool first = s % (2 * tr) == 2 * ts - 1;
                acc[t] = first ? c * sum : acc[t] + c * sum;
            } else {
                if (las_phase(tr, s) != LAS_FULL) continue;
                FE[3 * Rb.off + (int64_t)comp * Rb.ks + rc] = (R)(acc[t] / (dt * (double)tr));
            }
            continue;
        }
        if (stage == 0) {
            if (las_phase(ts, s) != LAS_FULL) continue;
            const bool first = s % (2 * tr) == 2 * ts - 1;
            const int64_t sa = dir == 1 ? Sb.D[2] : (int64_t)Sb.D[1] * Sb.D[2], sb = dir == 3 ? Sb.D[2] : 1;
            const double c = (mean ? 0.25 : 1.0) * (dt * (double)ts);
            const R *f = FE + NPR * Sb.off + sc;
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                const R *q = f + (int64_t)k * Sb.ks;
                const double sum = mean ? (double)q[0] + (double)q[sb] + (double)q[sa] + (double)q[sa + sb] : (double)q[0];
                double &a = acc[t * NPR + k];
                a = first ? c * sum : a + c * sum;
            }
        } else {
            if (las_phase(tr, s) != LAS_FULL) continue;
            R *o = FE + NPR * Rb.off + x.rcell;
#pragma unroll
            for (int k = 0; k < NPR; k++) o[(int64_t)k * Rb.ks] = (R)(acc[t * NPR + k] / (dt * (double)tr));
        }
    }
}

// FP32 copy of one geometry slab's metric (KParams::g32): each value rounded
// once, and 1 - alpha = 1 - 1/sqrt(-g^00) formed in FP64 before it is rounded
// (the mixed step's FP32 energy flux needs 1 - alpha to its own precision).
__global__ void __launch_bounds__(256) k_geom32(const double *__restrict__ gcov, const double *__restrict__ gcon,
                                                const double *__restrict__ gdet, float *__restrict__ g32, int64_t kcol)
{
    const int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (col >= kcol) return;
    for (int pos = 0; pos < NPOS; pos++) {
        float *o = g32 + (int64_t)pos * G32N * kcol + col;
        const double *gc = gcov + (int64_t)pos * 10 * kcol + col, *gi = gcon + (int64_t)pos * 10 * kcol + col;
        for (int s = 0; s < 10; s++) {
            o[s * kcol] = (float)gc[s * kcol];
            o[(10 + s) * kcol] = (float)gi[s * kcol];
        }
        o[20 * kcol] = (float)gdet[(int64_t)pos * kcol + col];
        o[21 * kcol] = (float)(1. - 1. / sqrt(-gi[0]));
    }
}

void launch_geom32(const double *gcov, const double *gcon, const double *gdet, float *g32, int64_t kcol,
                   hipStream_t s)
{
    if (kcol <= 0) return;
    hipLaunchKernelGGL(k_geom32, dim3((unsigned)((kcol + 255) / 256)), dim3(256), 0, s, gcov, gcon, gdet, g32, kcol);
}

// Courant dt of the step: per-block harmonic of the direction minima, minimum
// over blocks (step_ch.c:1313-1322), clamped to 1.3 dt (step_ch.c:974).  Resets
// the minima for the next corrector.  dt[0] <- new dt, dt[1] <- previous dt;
// bdt[n] <- block n's harmonic (bdt[][0] of the reference) when bdt is set.
__global__ void k_dt_finalize(int nblk, unsigned long long *__restrict__ dtmin, double *__restrict__ dt,
                              double *__restrict__ bdt)
{
    __shared__ double red[256];
    double m = 1.e9;
    for (int n = threadIdx.x; n < nblk; n += blockDim.x) {
        const double d1 = __longlong_as_double((long long)dtmin[n * 3 + 0]);
        const double d2 = __longlong_as_double((long long)dtmin[n * 3 + 1]);
        const double d3 = __longlong_as_double((long long)dtmin[n * 3 + 2]);
        const double b0 = 1. / (1. / d1 + 1. / d2 + 1. / d3);
        m = mymin(m, b0);
        if (bdt) bdt[n] = b0;
        const unsigned long long init = (unsigned long long)__double_as_longlong(1.e9);
        dtmin[n * 3 + 0] = init;
        dtmin[n * 3 + 1] = init;
        dtmin[n * 3 + 2] = init;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = mymin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double old = dt[0];
        double ndt = red[0];
        if (ndt > 1.3 * old) ndt = 1.3 * old;
        dt[1] = old;
        dt[0] = ndt;
    }
}

// Plain streaming copy (16 B per thread per iteration): the known-size kernel
// that calibrates the HBM byte counters.
__global__ void __launch_bounds__(256) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

void launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t s)
{
    const int64_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, (const double2 *)src, (double2 *)dst, n);
}

__global__ void k_fill_u64(unsigned long long *a, int n, unsigned long long v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// ---------------------------------------------------------------------------
// launchers.  The FP32 reconstruction, flux and EMF kernels are instantiated in
// hamr_kernels_f32.hip, the FP64 ones here.

void launch_flux(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, void *F,
                 unsigned long long *dtmin, int mixed, hipStream_t s)
{
    if (mixed) launch_flux_f32(P, dir, L, p, ps, (float *)F, dtmin, s);
    else recon_flux(P, dir, L, p, ps, (double *)F, dtmin, s);
}

void launch_emf_centre(const KParams &P, const Launch &L, const double *p, void *Ec, int mixed, hipStream_t s)
{
    if (mixed) launch_emf_centre_f32(P, L, p, (float *)Ec, s);
    else recon_emf_centre(P, L, p, (double *)Ec, s);
}

void launch_emf_corner(const KParams &P, const Launch &L, const void *Ec, const void *F1, const void *F2,
                       const void *F3, void *E, int mixed, hipStream_t s)
{
    if (mixed)
        launch_emf_corner_f32(P, L, (const float *)Ec, (const float *)F1, (const float *)F2, (const float *)F3,
                              (float *)E, s);
    else
        recon_emf_corner(P, L, (const double *)Ec, (const double *)F1, (const double *)F2, (const double *)F3,
                         (double *)E, s);
}

void launch_emf(const KParams &P, const Launch &centre, const Launch &corner, const double *p, const void *F1,
                const void *F2, const void *F3, void *Ec, void *E, int mixed, hipStream_t s)
{
    if (centre.wl.n <= 0) return;
    launch_emf_centre(P, centre, p, Ec, mixed, s);
    launch_emf_corner(P, corner, Ec, F1, F2, F3, E, mixed, s);
}

void launch_pole(const KParams &P, const int *ring, const int *rside, int nring, int len, int ni_max, double dt_val,
                 const double *dt_ptr, void *E, int mixed, hipStream_t s)
{
    if (nring <= 0) return;
    const int64_t w = (int64_t)nring * ni_max;
    const dim3 g((unsigned)((w + 63) / 64));
    if (mixed)
        hipLaunchKernelGGL((k_pole_average<float>), g, dim3(64), 0, s, P, ring, rside, nring, len, ni_max, dt_val, dt_ptr,
                           (float *)E);
    else
        hipLaunchKernelGGL((k_pole_average<double>), g, dim3(64), 0, s, P, ring, rside, nring, len, ni_max, dt_val,
                           dt_ptr, (double *)E);
}

void launch_ct_update(const KParams &P, const Launch &L, double Dt_val, const double *dt_ptr, double dt_mult,
                      const double *psi, double *psf, const void *E, int mixed, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (mixed) {
        if (P.tiled_push)
            hipLaunchKernelGGL((k_ct_update<float, true>), gdim(L), bdim(L), 0, s, P, L.wl, Dt_val, dt_ptr, dt_mult,
                               psi, psf, (const float *)E);
        else
            hipLaunchKernelGGL((k_ct_update<float, false>), gdim(L), bdim(L), 0, s, P, L.wl, Dt_val, dt_ptr, dt_mult,
                               psi, psf, (const float *)E);
    } else {
        if (P.tiled_push)
            hipLaunchKernelGGL((k_ct_update<double, true>), gdim(L), bdim(L), 0, s, P, L.wl, Dt_val, dt_ptr, dt_mult,
                               psi, psf, (const double *)E);
        else
            hipLaunchKernelGGL((k_ct_update<double, false>), gdim(L), bdim(L), 0, s, P, L.wl, Dt_val, dt_ptr, dt_mult,
                               psi, psf, (const double *)E);
    }
}

void launch_update_part(const KParams &P, int part, const Launch &L, int full, double Dt_val, const double *dt_ptr,
                        double dt_mult, const double *pi, const double *pb, double *pf, double *U0, double *Unew,
                        const double *psf, const void *F1, const void *F2, const void *F3, int *pflag, int *failcnt,
                        int mixed, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (part == 3) {  // k_update: parts 0 and 1 in one kernel
        const size_t lds = sizeof(double) * MV_N * (size_t)L.lds_cols;
        if (mixed) {
            auto k = P.tiled_push ? k_update<float, true> : k_update<float, false>;
            hipLaunchKernelGGL(k, gdim(L), bdim(L), lds, s, P, L.wl, full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, U0,
                               Unew, psf, (const float *)F1, (const float *)F2, (const float *)F3, pflag, failcnt);
        } else {
            auto k = P.tiled_push ? k_update<double, true> : k_update<double, false>;
            hipLaunchKernelGGL(k, gdim(L), bdim(L), lds, s, P, L.wl, full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, U0,
                               Unew, psf, (const double *)F1, (const double *)F2, (const double *)F3, pflag, failcnt);
        }
        return;
    }
    if (part == 0) {
        const size_t lds = sizeof(double) * MV_N * (size_t)L.lds_cols;
        if (mixed)
            hipLaunchKernelGGL((k_cons<float>), gdim(L), bdim(L), lds, s, P, L.wl, full, Dt_val, dt_ptr, dt_mult, pi, pb,
                               U0, Unew, psf, (const float *)F1, (const float *)F2, (const float *)F3);
        else
            hipLaunchKernelGGL((k_cons<double>), gdim(L), bdim(L), lds, s, P, L.wl, full, Dt_val, dt_ptr, dt_mult, pi,
                               pb, U0, Unew, psf, (const double *)F1, (const double *)F2, (const double *)F3);
    } else if (part == 1) {
        hipLaunchKernelGGL(k_invert, gdim(L), bdim(L), 0, s, P, L.wl, full, pi, pb, Unew, pf, pflag, failcnt);
    } else {
        hipLaunchKernelGGL(k_fallback, gdim(L), bdim(L), 0, s, P, L.wl, Unew, pf, pflag, failcnt);
    }
}

void launch_update(const KParams &P, const Launch &L, int full, double Dt_val, const double *dt_ptr, double dt_mult,
                   const double *pi, const double *pb, double *pf, double *U0, double *Unew, const double *psf,
                   const void *F1, const void *F2, const void *F3, int *pflag, int *failcnt, int mixed, hipStream_t s)
{
    for (int part : {3, 2})
        launch_update_part(P, part, L, full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, U0, Unew, psf, F1, F2, F3, pflag,
                           failcnt, mixed, s);
}

void launch_cons(const KParams &P, const Launch &L, double Dt, const double *p, double *U0, double *Unew,
                 const double *psf, const void *F1, const void *F2, const void *F3, int mixed, hipStream_t s)
{
    launch_update_part(P, 0, L, 0, Dt, nullptr, 1.0, p, p, nullptr, U0, Unew, psf, F1, F2, F3, nullptr, nullptr, mixed,
                       s);
}

void launch_bound(const KParams &P, int dir, const Launch &L, double *p, double *ps, int *pflag, int outflow_only,
                  hipStream_t s, int x3_rows_done)
{
    if (L.wl.n <= 0) return;
    if (dir == 1) hipLaunchKernelGGL(k_bound<1>, gdim(L), bdim(L), 0, s, P, L.wl, p, ps, pflag, outflow_only, 0);
    else if (dir == 2) hipLaunchKernelGGL(k_bound<2>, gdim(L), bdim(L), 0, s, P, L.wl, p, ps, pflag, outflow_only, 0);
    else hipLaunchKernelGGL(k_bound<3>, gdim(L), bdim(L), 0, s, P, L.wl, p, ps, pflag, outflow_only, x3_rows_done);
}

void launch_amr_exchange(const KParams &P, int dir, const AmrEntry *e, int64_t n, const AmrProlong *pe, int64_t npro,
                         double *p, double *ps, int *pflag, int threads, int grid, hipStream_t s, double *ph,
                         double *psh)
{
    if (n > 0) {
        const int64_t need = (n + threads - 1) / threads;
        const unsigned g = (unsigned)(grid > 0 && grid < need ? grid : need);
        hipLaunchKernelGGL(k_amr_copy, dim3(g), dim3(threads), 0, s, P, e, n, dir, p, ps, ph, psh, pflag);
    }
    if (npro > 0) {
        const int64_t need = (npro + threads - 1) / threads;
        const unsigned g = (unsigned)(grid > 0 && grid < need ? grid : need);
        hipLaunchKernelGGL(k_amr_prolong, dim3(g), dim3(threads), 0, s, P, pe, npro, dir, p, ps, ph, psh, pflag);
    }
}

void launch_amr_reflux(const KParams &P, const AmrReflux *e, int64_t n, int dir, void *F, double dt_val,
                       const double *dt_ptr, int mixed, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    if (mixed)
        hipLaunchKernelGGL(k_amr_reflux<float>, dim3(g), dim3(256), 0, s, P, e, n, dir, (float *)F, dt_val, dt_ptr);
    else
        hipLaunchKernelGGL(k_amr_reflux<double>, dim3(g), dim3(256), 0, s, P, e, n, dir, (double *)F, dt_val, dt_ptr);
}

void launch_amr_emf(const KParams &P, const AmrEmf *e, int64_t n, void *E, double dt_val, const double *dt_ptr,
                    int mixed, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    if (mixed)
        hipLaunchKernelGGL(k_amr_emf<float>, dim3(g), dim3(256), 0, s, P, e, n, (float *)E, dt_val, dt_ptr);
    else
        hipLaunchKernelGGL(k_amr_emf<double>, dim3(g), dim3(256), 0, s, P, e, n, (double *)E, dt_val, dt_ptr);
}

void launch_dt_finalize(int nblk, unsigned long long *dtmin, double *dt, hipStream_t s, double *bdt)
{
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, s, nblk, dtmin, dt, bdt);
}

void launch_las_corr(const KParams &P, const LasCorr *e, int64_t n, int dir, int stage, void *FE, double *acc,
                     const double *dt_ptr, int mixed, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
    if (mixed)
        hipLaunchKernelGGL(k_las_corr<float>, dim3(g), dim3(256), 0, s, P, e, n, dir, stage, (float *)FE, acc, dt_ptr);
    else
        hipLaunchKernelGGL(k_las_corr<double>, dim3(g), dim3(256), 0, s, P, e, n, dir, stage, (double *)FE, acc,
                           dt_ptr);
}

void launch_dt_reset(int nblk, unsigned long long *dtmin, hipStream_t s)
{
    const int n = nblk * 3;
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, dtmin, n,
                       (unsigned long long)0x41CDCD6500000000ULL /* 1e9 */);
}

}  // namespace hamr
