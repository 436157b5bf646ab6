// hamr_recon.h -- reconstruction, HLL flux and EMF kernels, templated on the
// arithmetic R of the face states (double: the FP64 step; float: the mixed
// step).  hamr_kernels.hip instantiates R = double, hamr_kernels_f32.hip
// R = float.  Both units are built without FMA contraction: the FP64 step
// follows the reference expression by expression, and in the FP32 fluxes
// contraction made the face-field error of one mixed step against the FP64
// step about 50 times larger (the corner EMFs difference fluxes of nearly
// equal size).
#pragma once

#include "hamr_hllc.h"
#include "hamr_kernels_dev.h"

namespace hamr {

// ---------------------------------------------------------------------------
// Fused PPM reconstruction + HLL flux through the lower dir-faces of a block:
// fluxcalcprep (GPU_program1.cu:78-190) and fluxcalc2D2 (:791-1068) in one pass,
// without the dq1/dq2 round trip through HBM.  Face range: [NG, NG+BS] along
// dir, [NG-1, NG+BS] across (the transverse ghost faces feed the corner EMFs).
//
// R is the arithmetic of reconstruction, state, fluxes and speeds: double for
// the FP64 step, float for the mixed step (primitives, face fields and metric
// are read in FP64 and rounded once; fluxes are stored as R).

// HLL flux through face `cell` (cell index of the face's upper cell; (i, j) its
// metric column) from the reconstructed states pl (lower side) and pr (upper
// side), fluxcalc2D2's flux and speed part (GPU_program1.cu:811-1045).  Stores
// the 9 fluxes and returns the face's Courant dt, or 1e9 where it is not a
// valid bound (no signal speed).
template <int DIR, class R>
__device__ __forceinline__ double hll_face(const KParams &P, const BlockDev &B, const double *__restrict__ ps, R *pl,
                                           R *pr, int i, int j, int64_t cell, R *__restrict__ F)
{
    const int64_t K = B.ks;
    const R bn = (R)ps[3 * B.off + (DIR - 1) * K + cell];
    pl[B1 + DIR - 1] = bn;
    pr[B1 + DIR - 1] = bn;
    if (DIR == 2 && (((B.pole & 1) && j == NG) || ((B.pole & 2) && j == NG + B.bs[1]))) {
        pl[B1] = R(0.);
        pr[B1] = R(0.);
    }
    const int64_t col = (int64_t)i * B.D[1] + j;
    const Geom<R> G = load_geom<R>(P, B, DIR == 1 ? FACE1 : DIR == 2 ? FACE2 : FACE3, col);
    const R gam = (R)P.gam;
    R t1[NPR], t2[NPR], t3[NPR], t4[NPR], cmax_l = R(0.), cmin_l = R(0.), cmax_r = R(0.), cmin_r = R(0.), dum;
    State<R> q;
    get_state(pl, G, q);
    primtoflux(pl, q, DIR, G, gam, t1, &cmax_l, &cmin_l);
    primtoflux(pl, q, 0, G, gam, t2, &dum, &dum);
    get_state(pr, G, q);
    primtoflux(pr, q, DIR, G, gam, t3, &cmax_r, &cmin_r);
    primtoflux(pr, q, 0, G, gam, t4, &dum, &dum);
    const R cmax = fabs(mymax(mymax(R(0.), cmax_l), cmax_r));
    const R cmin = fabs(mymax(mymax(R(0.), -cmin_l), -cmin_r));
    const R ctop = mymax(cmax, cmin);
    R *Fb = F + NPR * B.off + cell;
#pragma unroll
    for (int k = 0; k < NPR; k++)
        Fb[(int64_t)k * K] = dv(cmax * t1[k] + cmin * t3[k] - cmax * cmin * (t4[k] - t2[k]), cmax + cmin + R(kSMALL));
    const double v = P.cour * B.dx[DIR - 1] / (double)ctop;
    return (v < 1.e9 && v > 1.e-6) ? v : 1.e9;
}

// HLLC flux through the same face (fluxcalc2D_FT, GPU_hllc.h:3-247): the
// Courant dt from the coordinate speeds as in hll_face, the flux from the
// orthonormal-frame solver of hamr_hllc.h with the face's tetrad.
template <int DIR, class R>
__device__ __forceinline__ double hllc_face(const KParams &P, const BlockDev &B, const double *__restrict__ ps, R *pl,
                                            R *pr, int i, int j, int64_t cell, R *__restrict__ F)
{
    const int64_t K = B.ks;
    const R bn = (R)ps[3 * B.off + (DIR - 1) * K + cell];
    pl[B1 + DIR - 1] = bn;
    pr[B1 + DIR - 1] = bn;
    if (DIR == 2 && (((B.pole & 1) && j == NG) || ((B.pole & 2) && j == NG + B.bs[1]))) {
        pl[B1] = R(0.);
        pr[B1] = R(0.);
    }
    const int64_t col = (int64_t)i * B.D[1] + j;
    const Geom<R> G = load_geom<R>(P, B, DIR == 1 ? FACE1 : DIR == 2 ? FACE2 : FACE3, col);
    const R gam = (R)P.gam;
    State<R> ql, qr;
    get_state(pl, G, ql);
    get_state(pr, G, qr);
    R cmax_l, cmin_l, cmax_r, cmin_r;
    vchar_state(pl, ql, DIR, G, gam, &cmax_l, &cmin_l);
    vchar_state(pr, qr, DIR, G, gam, &cmax_r, &cmin_r);
    const R cmax = fabs(mymax(mymax(R(0.), cmax_l), cmax_r));
    const R cmin = fabs(mymax(mymax(R(0.), -cmin_l), -cmin_r));
    const R ctop = mymax(cmax, cmin);
    R mud[16], mudi[16];
    const double *T = P.tet + B.moff * 96 + (int64_t)(DIR - 1) * 32 * B.kcol + col;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        mud[e] = (R)T[(int64_t)e * B.kcol];
        mudi[e] = (R)T[(int64_t)(16 + e) * B.kcol];
    }
    R F1[NPR];
    hllc_frame_flux<DIR>(pl, pr, ql.ucon, ql.bcon, qr.ucon, qr.bcon, mud, mudi, face_velocity<DIR>(G.gcon), gam,
                         sizeof(R) == 4 ? G.oma : R(0.), F1);
    R *Fb = F + NPR * B.off + cell;
#pragma unroll
    for (int k = 0; k < NPR; k++) Fb[(int64_t)k * K] = G.g * F1[k];
    const double v = P.cour * B.dx[DIR - 1] / (double)ctop;
    return (v < 1.e9 && v > 1.e-6) ? v : 1.e9;
}

__device__ __forceinline__ void dt_reduce(double vdt, unsigned long long *__restrict__ dtmin, int slot)
{
    vdt = wave_min(vdt);
    if ((threadIdx.x & 63) == 0 && vdt < 1.e9) atomicMin(&dtmin[slot], (unsigned long long)__double_as_longlong(vdt));
}

// One thread per face: both parabolas of the face's two cells.  At least 3
// waves per SIMD (<= 168 VGPRs): the FP64 flux needs 169 registers
// unconstrained (the FP32 flux needs 96, five waves).  A variant that marches
// each thread through a segment of faces along x1 or x2 and forms each cell's
// parabola once issues 29% fewer FP32 VALU instructions per face, but its
// carried state takes 157 registers (three waves) and it ran 7-35% slower.
// RS selects the Riemann solver: 0 HLL, 1 HLLC.  The FP64 HLLC kernel is held
// at two waves per SIMD: 256 VGPRs with about 20 spilled to scratch ran 1.4x
// faster than one wave with none spilled.  The FP32 HLLC kernel at four waves
// (128 VGPRs, 9 spilled) ran 4% faster than at three (139, none spilled).
template <int DIR, class R, int RS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RS ? (sizeof(R) == 4 ? 4 : 2) : 3)))
k_flux(KParams P, WorkList wl, const double *__restrict__ p, const double *__restrict__ ps, R *__restrict__ F,
       unsigned long long *__restrict__ dtmin)
{
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const int n = it.x;
        const BlockDev &B = P.blk[n];
        const int e1 = B.bs[0] + (DIR == 1 ? 1 : 2), e2 = B.bs[1] + (DIR == 2 ? 1 : 2), e3 = B.bs[2] + (DIR == 3 ? 1 : 2);
        const int t = it.y * blockDim.x + threadIdx.x;
        double vdt = 1.e9;
        if (t < e1 * e2 * e3) {
            int a, b, c;
            region_index(t, e2, e3, a, b, c);
            const int i = a + (DIR == 1 ? NG : NG - 1);
            const int j = b + (DIR == 2 ? NG : NG - 1);
            const int z = c + (DIR == 3 ? NG : NG - 1);
            const int64_t K = B.ks;
            const int64_t cell = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
            const int64_t sd = (DIR == 1) ? (int64_t)B.D[1] * B.D[2] : (DIR == 2) ? (int64_t)B.D[2] : 1;
            const double *pb = p + NPR * B.off;
            R pl[NPR], pr[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                if (k == B1 + DIR - 1) continue;  // the face-normal field comes from ps
                const double *q = pb + (int64_t)k * K + cell;
                const R ym3 = (R)q[-3 * sd], ym2 = (R)q[-2 * sd], ym1 = (R)q[-sd], y0 = (R)q[0], yp1 = (R)q[sd],
                        yp2 = (R)q[2 * sd];
                R l, r;
                para(ym3, ym2, ym1, y0, yp1, l, r);
                pl[k] = r;
                para(ym2, ym1, y0, yp1, yp2, l, r);
                pr[k] = l;
            }
            if constexpr (RS == 1) vdt = hllc_face<DIR>(P, B, ps, pl, pr, i, j, cell, F);
            else vdt = hll_face<DIR>(P, B, ps, pl, pr, i, j, cell, F);
        }
        if (dtmin) dt_reduce(vdt, dtmin, n * 3 + DIR - 1);
    }
}

// ---------------------------------------------------------------------------
// Line-staged reconstruction + Riemann flux: the faces, states and fluxes of
// k_flux, with each cell's MC slope and limited parabola formed once instead
// of once per face that reaches it.
//
// A work-list item is a tile (line_tiles): one segment of S consecutive faces
// along DIR on L consecutive lines of one row, a line being one transverse
// position of the face range, with one thread per cell of each line's S + 3
// cells f0 - 2 .. f0 + S (f0: the segment's first face, as the index of its
// upper cell).  In three steps, with a barrier between them:
//   1. every cell's MC slope from its own three-point stencil (read from
//      memory, rounded to R) into LDS;
//   2. cells f0 - 1 .. f0 + S - 1: the interface values of both boundaries
//      (the neighbours' slopes from LDS) and the limited parabola; the upper
//      edge goes to LDS, the lower edge stays in registers;
//   3. one thread per face (the face's upper cell): the lower cell's upper
//      edge from LDS and its own lower edge are the face states.
// Every value is para's expression on para's operands, so the states and
// fluxes are k_flux's bit for bit.  Thread t of a tile: line q and cell c
// (c = 0: cell f0 - 2) with t = c L + q for DIR = 1, 2 (consecutive threads
// on consecutive lines, i.e. consecutive z) and t = q (S + 3) + c for DIR = 3
// (consecutive threads along the z-row).  SC is the segment length when it is
// known at compile time (33: blocks of 32 or 64 cells along DIR; 17: 16
// cells), 0 otherwise.  A shorter last segment masks the cells past its faces.
template <int DIR, class R, int SC>
__device__ __forceinline__ void line_tile_states(const BlockDev &B, int item, const double *__restrict__ p, R *sm,
                                                 R *pl, R *pr, bool &fact, int64_t &cell, int &i, int &j)
{
    constexpr int T = kLineT, NV = NPR - 1;
    const int t = threadIdx.x;
    const LineTiles lt = line_tiles(B.bs[0], B.bs[1], B.bs[2], DIR);
    const int S = SC ? SC : lt.S, W = S + 3, L = SC ? T / (SC + 3) : lt.L;
    const int seg = item % lt.nseg, rest = item / lt.nseg, row = rest / lt.tpr, tb0 = (rest - row * lt.tpr) * L;
    const int nl = mymin(L, lt.n2 - tb0);
    const int nf = mymin(S, B.bs[DIR - 1] + 1 - seg * S);  // faces of this segment
    const int f0 = NG + seg * S;
    int q, c;
    if (DIR == 3) { q = t / W; c = t - q * W; }
    else { c = t / L; q = t - c * L; }
    const int sn = DIR == 3 ? 1 : L;  // LDS distance to the next cell along DIR
    const bool ok = q < nl && c < W;
    const bool slope = ok && c <= nf + 2, edge = ok && c >= 1 && c < nf + 2;
    fact = ok && c >= 2 && c < nf + 2;
    const int64_t D12 = (int64_t)B.D[1] * B.D[2];
    const int64_t sdir = DIR == 1 ? D12 : DIR == 2 ? (int64_t)B.D[2] : 1;
    const int64_t sq = DIR == 3 ? (int64_t)B.D[2] : 1;  // between consecutive lines of a row
    const int ta = row + NG - 1, tb = tb0 + NG - 1;
    const int64_t base = DIR == 1 ? (int64_t)ta * B.D[2] + tb : DIR == 2 ? (int64_t)ta * D12 + tb
                                                                         : (int64_t)ta * D12 + (int64_t)tb * B.D[2];
    const int pos = f0 - 2 + c;  // this thread's cell along DIR
    cell = base + q * sq + (int64_t)pos * sdir;
    i = DIR == 1 ? pos : ta;  // the face's metric column (i, j)
    j = DIR == 1 ? ta : DIR == 2 ? pos : tb + q;
    R *const DQ = sm, *const RE = sm + NV * T;
    const int64_t K = B.ks;
    const double *pc = p + NPR * B.off + cell;
    R ym[NV], y0[NV], yp[NV], dq[NV];
#pragma unroll
    for (int k = 0, v = 0; k < NPR; k++) {
        if (k == B1 + DIR - 1) continue;  // the face-normal field comes from ps
        if (slope) {
            const double *y = pc + (int64_t)k * K;
            ym[v] = (R)y[-sdir];
            y0[v] = (R)y[0];
            yp[v] = (R)y[sdir];
            dq[v] = mc_slope(ym[v], y0[v], yp[v]);
            DQ[v * T + t] = dq[v];
        }
        v++;
    }
    __syncthreads();
    R le[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        if (edge) {
            R l, r;
            ppm_limit(y0[v], ppm_face(ym[v], y0[v], DQ[v * T + t - sn], dq[v]),
                      ppm_face(y0[v], yp[v], dq[v], DQ[v * T + t + sn]), l, r);
            le[v] = l;
            RE[v * T + t] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0, v = 0; k < NPR; k++) {
        if (k == B1 + DIR - 1) continue;
        if (fact) {
            pl[k] = RE[v * T + t - sn];
            pr[k] = le[v];
        }
        v++;
    }
}

// Occupancy: four waves per SIMD in FP32 (<= 128 VGPRs), two in FP64 (the
// FP64 states, speeds and fluxes need ~220 VGPRs).  The LDS of one tile:
// 2 (NPR - 1) kLineT values of R.
template <int DIR, class R, int RS>
__global__ void __launch_bounds__(kLineT) __attribute__((amdgpu_waves_per_eu(RS || sizeof(R) == 8 ? 2 : 4)))
k_flux_lines(KParams P, WorkList wl, const double *__restrict__ p, const double *__restrict__ ps, R *__restrict__ F,
             unsigned long long *__restrict__ dtmin)
{
    extern __shared__ double smem_d[];
    R *const sm = reinterpret_cast<R *>(smem_d);
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const int n = it.x;
        const BlockDev &B = P.blk[n];
        const int nfd = B.bs[DIR - 1] + 1;
        R pl[NPR], pr[NPR];
        bool fact;
        int64_t cell;
        int i, j;
        if (nfd == 33 || nfd == 65) line_tile_states<DIR, R, 33>(B, it.y, p, sm, pl, pr, fact, cell, i, j);
        else if (nfd == 17) line_tile_states<DIR, R, 17>(B, it.y, p, sm, pl, pr, fact, cell, i, j);
        else line_tile_states<DIR, R, 0>(B, it.y, p, sm, pl, pr, fact, cell, i, j);
        double vdt = 1.e9;
        if (fact) {
            if constexpr (RS == 1) vdt = hllc_face<DIR>(P, B, ps, pl, pr, i, j, cell, F);
            else vdt = hll_face<DIR>(P, B, ps, pl, pr, i, j, cell, F);
        }
        if (dtmin) dt_reduce(vdt, dtmin, n * 3 + DIR - 1);
        // the next item's first barrier orders its LDS writes after these reads
    }
}

// ---------------------------------------------------------------------------
// Cell-centre EMF, consttransport1 (GPU_const_trans.h:63-119), cells [NG-1, NG+BS].
template <class R>
__global__ void __launch_bounds__(256) k_emf_centre(KParams P, WorkList wl, const double *__restrict__ p,
                                                    R *__restrict__ Ec)
{
    FOR_ITEMS(wl)
    {
        const int2 it = wl_get(wl, w_);
        const BlockDev &B = P.blk[it.x];
        const int e1 = B.bs[0] + 2, e2 = B.bs[1] + 2, e3 = B.bs[2] + 2;
        const int t = it.y * blockDim.x + threadIdx.x;
        if (t >= e1 * e2 * e3) continue;
        int a, b, c;
        region_index(t, e2, e3, a, b, c);
        const int i = a + NG - 1, j = b + NG - 1, z = c + NG - 1;
        const int64_t K = B.ks;
        const int64_t cell = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
        const double *pb = p + NPR * B.off + cell;
        R pr[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) pr[k] = (R)pb[(int64_t)k * K];
        const Geom<R> G = load_geom<R>(P, B, CENT, (int64_t)i * B.D[1] + j);
        R ucon[4], ucov[4], bcon[4];
        ucon_calc(pr, G, ucon);
        lower10(ucon, G.gcov, ucov);
        bcon_calc(pr, ucon, ucov, bcon);
        R *E = Ec + 3 * B.off + cell;
        E[0] = -G.g * (ucon[2] * bcon[3] - ucon[3] * bcon[2]);
        E[K] = -G.g * (ucon[3] * bcon[1] - ucon[1] * bcon[3]);
        E[2 * K] = -G.g * (ucon[1] * bcon[2] - ucon[2] * bcon[1]);
    }
}

// Corner EMF upwinded on the mass-flux sign, consttransport2
// (GPU_const_trans.h:121-339) with the polar rule (:233-236); corners [NG, NG+BS].
template <class R>
__global__ void __launch_bounds__(256) k_emf_corner(KParams P, WorkList wl, const R *__restrict__ Ec,
                                                    const R *__restrict__ F1, const R *__restrict__ F2,
                                                    const R *__restrict__ F3, R *__restrict__ E)
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
        const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
        const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2], sj = B.D[2];
        const R *E1c = Ec + 3 * B.off, *E2c = E1c + K, *E3c = E2c + K;
        const R *f1 = F1 + NPR * B.off, *f2 = F2 + NPR * B.off, *f3 = F3 + NPR * B.off;
        const R *F1R = f1 + RHO * K, *F2R = f2 + RHO * K, *F3R = f3 + RHO * K;
        const R *F1B2 = f1 + B2 * K, *F1B3 = f1 + B3 * K;
        const R *F2B1 = f2 + B1 * K, *F2B3 = f2 + B3 * K;
        const R *F3B1 = f3 + B1 * K, *F3B2 = f3 + B2 * K;
        const R dEL13_1 = E1c[g] - F3B2[g];
        const R dEL13_2 = E1c[g - sj] - F3B2[g - sj];
        const R dER13_1 = F3B2[g] - E1c[g - 1];
        const R dER13_2 = F3B2[g - sj] - E1c[g - sj - 1];
        const R dEL12_1 = E1c[g] + F2B3[g];
        const R dEL12_2 = E1c[g - 1] + F2B3[g - 1];
        const R dER12_1 = -F2B3[g] - E1c[g - sj];
        const R dER12_2 = -F2B3[g - 1] - E1c[g - sj - 1];
        const R dEL21_1 = E2c[g] - F1B3[g];
        const R dEL21_2 = E2c[g - 1] - F1B3[g - 1];
        const R dER21_1 = F1B3[g] - E2c[g - si];
        const R dER21_2 = F1B3[g - 1] - E2c[g - si - 1];
        const R dEL23_1 = E2c[g] + F3B1[g];
        const R dEL23_2 = E2c[g - si] + F3B1[g - si];
        const R dER23_1 = -F3B1[g] - E2c[g - 1];
        const R dER23_2 = -F3B1[g - si] - E2c[g - si - 1];
        const R dEL31_1 = E3c[g] + F1B2[g];
        const R dEL31_2 = E3c[g - sj] + F1B2[g - sj];
        const R dER31_1 = -F1B2[g] - E3c[g - si];
        const R dER31_2 = -F1B2[g - sj] - E3c[g - si - sj];
        const R dEL32_1 = E3c[g] - F2B1[g];
        const R dEL32_2 = E3c[g - si] - F2B1[g - si];
        const R dER32_1 = F2B1[g] - E3c[g - sj];
        const R dER32_2 = F2B1[g - si] - E3c[g - si - sj];
#define LE(x) ((R)((x) <= R(0.)))
#define GT(x) ((R)((x) > R(0.)))
        R e1v = R(0.25) * ((-F2B3[g] - (dEL13_1 * LE(F2R[g]) + dEL13_2 * GT(F2R[g])))
            + (-F2B3[g - 1] + (dER13_1 * LE(F2R[g - 1]) + dER13_2 * GT(F2R[g - 1])))
            + (F3B2[g] - (dEL12_1 * LE(F3R[g]) + dEL12_2 * GT(F3R[g])))
            + (F3B2[g - sj] + (dER12_1 * LE(F3R[g - sj]) + dER12_2 * GT(F3R[g - sj]))));
        const R e2v = R(0.25) * ((-F3B1[g] - (dEL21_1 * LE(F3R[g]) + dEL21_2 * GT(F3R[g])))
            + (-F3B1[g - si] + (dER21_1 * LE(F3R[g - si]) + dER21_2 * GT(F3R[g - si])))
            + (F1B3[g] - (dEL23_1 * LE(F1R[g]) + dEL23_2 * GT(F1R[g])))
            + (F1B3[g - 1] + (dER23_1 * LE(F1R[g - 1]) + dER23_2 * GT(F1R[g - 1]))));
        R e3v = R(0.25) * ((F2B1[g] - (dEL31_1 * LE(F2R[g]) + dEL31_2 * GT(F2R[g])))
            + (F2B1[g - si] + (dER31_1 * LE(F2R[g - si]) + dER31_2 * GT(F2R[g - si])))
            + (-F1B2[g] - (dEL32_1 * LE(F1R[g]) + dEL32_2 * GT(F1R[g])))
            + (-F1B2[g - sj] + (dER32_1 * LE(F1R[g - sj]) + dER32_2 * GT(F1R[g - sj]))));
#undef LE
#undef GT
        if (((B.pole & 1) && j == NG) || ((B.pole & 2) && j == NG + B.bs[1])) {
            e3v = R(0.);
            e1v = R(-0.5) * (F2B3[g] + F2B3[g - 1]);
        }
        R *Eb = E + 3 * B.off + g;
        Eb[0] = e1v;
        Eb[K] = e2v;
        Eb[2 * K] = e3v;
    }
}

// Launchers of one arithmetic (R = double here and in hamr_kernels.hip, R =
// float in hamr_kernels_f32.hip).
template <int DIR, class R, int RS>
static void launch_flux_lines(const KParams &P, const Launch &L, const double *p, const double *ps, R *F,
                              unsigned long long *dtmin, hipStream_t s)
{
    const size_t lds = (size_t)2 * (NPR - 1) * kLineT * sizeof(R);
    hipLaunchKernelGGL((k_flux_lines<DIR, R, RS>), gdim(L), bdim(L), lds, s, P, L.wl, p, ps, F, dtmin);
}

template <class R>
void recon_flux(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, R *F,
                unsigned long long *dtmin, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (L.line) {
        if (P.riemann == 1) {
            if (dir == 1) launch_flux_lines<1, R, 1>(P, L, p, ps, F, dtmin, s);
            else if (dir == 2) launch_flux_lines<2, R, 1>(P, L, p, ps, F, dtmin, s);
            else launch_flux_lines<3, R, 1>(P, L, p, ps, F, dtmin, s);
            return;
        }
        if (dir == 1) launch_flux_lines<1, R, 0>(P, L, p, ps, F, dtmin, s);
        else if (dir == 2) launch_flux_lines<2, R, 0>(P, L, p, ps, F, dtmin, s);
        else launch_flux_lines<3, R, 0>(P, L, p, ps, F, dtmin, s);
        return;
    }
    if (P.riemann == 1) {
        if (dir == 1) hipLaunchKernelGGL((k_flux<1, R, 1>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
        else if (dir == 2) hipLaunchKernelGGL((k_flux<2, R, 1>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
        else hipLaunchKernelGGL((k_flux<3, R, 1>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
        return;
    }
    if (dir == 1) hipLaunchKernelGGL((k_flux<1, R, 0>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
    else if (dir == 2) hipLaunchKernelGGL((k_flux<2, R, 0>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
    else hipLaunchKernelGGL((k_flux<3, R, 0>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
}

template <class R>
void recon_emf_centre(const KParams &P, const Launch &L, const double *p, R *Ec, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    hipLaunchKernelGGL((k_emf_centre<R>), gdim(L), bdim(L), 0, s, P, L.wl, p, Ec);
}

template <class R>
void recon_emf_corner(const KParams &P, const Launch &L, const R *Ec, const R *F1, const R *F2, const R *F3, R *E,
                      hipStream_t s)
{
    if (L.wl.n <= 0) return;
    hipLaunchKernelGGL((k_emf_corner<R>), gdim(L), bdim(L), 0, s, P, L.wl, Ec, F1, F2, F3, E);
}

// The FP32 instantiations (hamr_kernels_f32.hip).
void launch_flux_f32(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, float *F,
                     unsigned long long *dtmin, hipStream_t s);
void launch_emf_centre_f32(const KParams &P, const Launch &L, const double *p, float *Ec, hipStream_t s);
void launch_emf_corner_f32(const KParams &P, const Launch &L, const float *Ec, const float *F1, const float *F2,
                           const float *F3, float *E, hipStream_t s);

}  // namespace hamr
