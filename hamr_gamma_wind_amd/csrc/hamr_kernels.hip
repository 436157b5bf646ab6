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
#include "hamr_kernels.h"

#include <cstdio>

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

// Metric at one position of one column, rounded to R (the metric is stored in FP64).
template <class R = double>
__device__ __forceinline__ Geom<R> load_geom(const KParams &P, const BlockDev &B, int pos, int64_t col)
{
    Geom<R> G;
    const double *gc = P.gcov + mg10(B, pos, 0) + col;
    const double *gi = P.gcon + mg10(B, pos, 0) + col;
#pragma unroll
    for (int s = 0; s < 10; s++) {
        G.gcov[s] = (R)gc[s * B.kcol];
        G.gcon[s] = (R)gi[s * B.kcol];
    }
    G.g = (R)P.gdet[mg1(B, pos) + col];
    if constexpr (sizeof(R) == 4) G.oma = (R)(1. - 1. / sqrt(-gi[0]));
    return G;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

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

// ---------------------------------------------------------------------------
// Fused PPM reconstruction + HLL flux through the lower dir-faces of a block:
// fluxcalcprep (GPU_program1.cu:78-190) and fluxcalc2D2 (:791-1068) in one pass,
// without the dq1/dq2 round trip through HBM.  Face range: [NG, NG+BS] along
// dir, [NG-1, NG+BS] across (the transverse ghost faces feed the corner EMFs).
//
// R is the arithmetic of reconstruction, state, fluxes and speeds: double for
// the FP64 step, float for the mixed step (primitives, face fields and metric
// are read in FP64 and rounded once; fluxes are stored as R).
// At least 3 waves per SIMD (<= 168 VGPRs): the FP64 flux needs 169 registers
// unconstrained, one too many for the third wave.
template <int DIR, class R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_flux(KParams P, WorkList wl, const double *__restrict__ p,
                                              const double *__restrict__ ps, R *__restrict__ F,
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
                const double *q = pb + (int64_t)k * K + cell;
                const R ym3 = (R)q[-3 * sd], ym2 = (R)q[-2 * sd], ym1 = (R)q[-sd], y0 = (R)q[0], yp1 = (R)q[sd],
                        yp2 = (R)q[2 * sd];
                R l, r;
                para(ym3, ym2, ym1, y0, yp1, l, r);
                pl[k] = r;
                para(ym2, ym1, y0, yp1, yp2, l, r);
                pr[k] = l;
            }
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
            if (v < 1.e9 && v > 1.e-6) vdt = v;
        }
        if (dtmin) {
            vdt = wave_min(vdt);
            if ((threadIdx.x & 63) == 0 && vdt < 1.e9)
                atomicMin(&dtmin[n * 3 + DIR - 1], (unsigned long long)__double_as_longlong(vdt));
        }
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
    const int side = rside[rg];
    const int i = il + NG, j = side ? NG + B0.bs[1] : NG;
    const double dt = dt_ptr ? dt_ptr[0] : dt_val;
    double avg = 0.;
    for (int b = 0; b < len; b++) {
        const BlockDev &B = P.blk[rb[b]];
        const R *e1 = E + 3 * B.off + ((int64_t)i * B.D[1] + j) * B.D[2];
        double s = 0.;
        for (int zz = 0; zz < B.bs[2]; zz++) s += dt * (double)e1[NG + zz];
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
template <class R>
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
        const double Dt = dt_ptr ? dt_ptr[0] * dt_mult : Dt_val;
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
        if (j < NG + B.bs[1] && z < NG + B.bs[2]) {
            temp = 0.;
            temp += 1.0 / 1.0 * si_[0] * gF1;
            temp += -Dt / (1.0 * dx2) * (Eb(2 * K + sj) - Eb(2 * K));
            temp += Dt / (1.0 * dx3) * (Eb(K + 1) - Eb(K));
            sf[0] = temp / gF1;
        }
        if (i < NG + B.bs[0] && j < NG + B.bs[1]) {
            temp = si_[2 * K] * gF3;
            temp += -Dt / dx1 * (Eb(K + si) - Eb(K));
            temp += Dt / dx2 * (Eb(sj) - Eb(0));
            sf[2 * K] = temp / gF3;
        }
        if (i < NG + B.bs[0] && z < NG + B.bs[2]) {
            temp = 0.;
            temp += 1.0 / 1.0 * si_[K] * gF2;
            temp += Dt / (1.0 * dx1) * (Eb(2 * K + si) - Eb(2 * K));
            temp += -Dt / (1.0 * dx3) * (Eb(1) - Eb(0));
            sf[K] = temp / gF2;
        }
    }
}

// Conserved update, source, face-B average, Noble cascade, floors and flags:
// the fixup kernel of GPU_fixup.h:5-425, as three kernels.  k_cons forms the
// updated conserved variables (full = 0, predictor: U^n from pi is stored in
// U0, source from pi; full = 1, corrector: U from U0, source from pb) and
// writes them to Unew; k_invert runs the 2D inversion on Unew with the same
// initial guess (pi or pb), applies the floors and sets the flags; k_fallback
// finishes the cascade for the cells whose 2D inversion failed.  The split
// keeps each kernel's registers low enough for several waves per SIMD, which
// hides the latency of its loads; the arithmetic is that of one fused kernel.

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

// R is the storage type of the fluxes: double (FP64 step, the reference's
// arithmetic) or float (mixed step, compensated FP64 accumulation).
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
            const int qq = t / bs3 - q0;
            const double *mv = smem + qq;
            const double Dt = dt_ptr ? dt_ptr[0] * dt_mult : Dt_val;
            const int64_t K = B.ks, si = (int64_t)B.D[1] * B.D[2], sj = B.D[2];
            const int64_t g = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
            const int64_t base = NPR * B.off + g;
            Geom<double> G;
#pragma unroll
            for (int s = 0; s < 10; s++) {
                G.gcov[s] = mv[(MV_GCOV + s) * ncw];
                G.gcon[s] = mv[(MV_GCON + s) * ncw];
            }
            G.g = mv[MV_G * ncw];
            double pf[NPR], U[NPR], dU[NPR];
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
            R f1l[NPR], f1h[NPR], f2l[NPR], f2h[NPR], f3l[NPR], f3h[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                const R *f1 = F1 + base + k * K, *f2 = F2 + base + k * K, *f3 = F3 + base + k * K;
                f1l[k] = f1[0];
                f1h[k] = f1[si];
                f2l[k] = f2[0];
                f2h[k] = f2[sj];
                f3l[k] = f3[0];
                f3h[k] = f3[1];
            }
            source(pf, q, [&](int lam, int s) { return mv[(MV_CONN + lam * 10 + s) * ncw]; }, G.g, P.gam, dU);
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
#pragma unroll
                for (int k = 0; k < NPR; k++) {
                    double sum = U[k], cmp = 0.;
                    comp_add(sum, cmp, -(Dt * ((double)f1h[k] - (double)f1l[k]) / B.dx[0]));
                    comp_add(sum, cmp, -(Dt * ((double)f2h[k] - (double)f2l[k]) / B.dx[1]));
                    comp_add(sum, cmp, -(Dt * ((double)f3h[k] - (double)f3l[k]) / B.dx[2]));
                    comp_add(sum, cmp, Dt * dU[k]);
                    U[k] = sum + cmp;
                }
            }
            {
                const double *sp = psf + 3 * B.off + g;
                U[B1] = 0.0;
                U[B2] = 0.0;
                U[B1] += (sp[0] * mv[MV_GF1 * ncw] + sp[si] * mv[MV_GF1P * ncw]) / (2.0 * 1.0);
                U[B2] += (sp[K] * mv[MV_GF2 * ncw] + sp[K + sj] * mv[MV_GF2P * ncw]) / (2.0 * 1.0);
                U[B3] = (sp[2 * K] * mv[MV_GF3 * ncw] + sp[2 * K + 1] * mv[MV_GF3 * ncw]) / 2.0;
            }
#pragma unroll
            for (int k = 0; k < NPR; k++) Unew[base + k * K] = U[k];
        }
        __syncthreads();
    }
}

// Floors (the build's specification: rho >= rho_floor r^-1.5, u >= u_floor
// r^-2.5, flag -333 as GPU_fixup.h:399-415) and the result of one cell.
__device__ __forceinline__ void finish_cell(const KParams &P, const BlockDev &B, int n, int64_t col, int64_t g,
                                            int64_t base, int flag, double *pf, double *__restrict__ pf_out,
                                            int *__restrict__ pflag, int *__restrict__ failcnt)
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
    if (hit) {
        flag = -333;
        atomicAdd(&failcnt[n * 4 + 3], 1);
    }
    pflag[B.off + g] = flag;
#pragma unroll
    for (int k = 0; k < NPR; k++) pf_out[base + k * K] = pf[k];
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
            finish_cell(P, B, n, col, g, base, 0, pf, pf_out, pflag, failcnt);
            continue;
        }
        atomicAdd(&failcnt[n * 4 + 0], 1);
        pflag[B.off + g] = kFallbackPending;
#pragma unroll
        for (int k = 0; k < NPR; k++) pf_out[base + k * K] = pf[k];
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
        finish_cell(P, B, n, col, g, base, flag, pf, pf_out, pflag, failcnt);
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
                                               int outflow_only)
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
// each coarse face.  A prolongation entry names the coarse cell that contains
// the ghost and its six neighbours; the ghost gets the coarse value plus
// a quarter of each direction's MC slope (slope_lim_2, GPU_program2.cu:10-45),
// signed by the side of the coarse cell it lies on (unpackreceivecoarse1,
// GPU_program2.cu:615-803).  A face field along direction a carries no slope
// along a; a fine face that halves a coarse cell starts from the coarse
// cell-centred field, one on a coarse face from that face's value.

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

__global__ void __launch_bounds__(256) k_amr_copy(KParams P, const AmrEntry *__restrict__ e, int64_t n, double *p,
                                                  double *ps, int *pflag)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const AmrEntry E = e[t];
        double v[NPR + 3];
        amr_source(P, p, ps, E.sblk, E.scode, v);
        const BlockDev &B = P.blk[E.dblk];
        const BlockDev &S = P.blk[E.sblk];
        const int64_t K = B.ks;
        double *pd = p + NPR * B.off + E.dcell, *sd = ps + 3 * B.off + E.dcell;
#pragma unroll
        for (int k = 0; k < NPR; k++) pd[k * K] = v[k];
#pragma unroll
        for (int a = 0; a < 3; a++) sd[a * K] = v[NPR + a];
        if (pflag) pflag[B.off + E.dcell] = pflag[S.off + (E.scode & 0x0FFFFFFF)];
    }
}

__global__ void __launch_bounds__(256) k_amr_prolong(KParams P, const AmrProlong *__restrict__ e, int64_t n,
                                                     double *p, double *ps, int *pflag)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const AmrProlong E = e[t];
        const BlockDev &B = P.blk[E.dblk];
        const BlockDev &C = P.blk[E.cblk];
        // side of the coarse cell along each direction (global fine index parity)
        int loc[3];
        region_index(E.dcell, B.D[1], B.D[2], loc[0], loc[1], loc[2]);
        double sgn[3];
#pragma unroll
        for (int d = 0; d < 3; d++) sgn[d] = ((B.o[d] + loc[d] - NG) & 1) ? 1.0 : -1.0;
        double c0[NPR + 3], dq[3][NPR + 3];
        amr_source(P, p, ps, E.cblk, (AS_CELL << 28) | E.ccell, c0);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int cm = E.nb[2 * d][1], cp = E.nb[2 * d + 1][1];
            if ((cm >> 28) == AS_NONE || (cp >> 28) == AS_NONE) {
#pragma unroll
                for (int k = 0; k < NPR + 3; k++) dq[d][k] = 0.;
                continue;
            }
            double vm[NPR + 3], vp[NPR + 3];
            amr_source(P, p, ps, E.nb[2 * d][0], cm, vm);
            amr_source(P, p, ps, E.nb[2 * d + 1][0], cp, vp);
#pragma unroll
            for (int k = 0; k < NPR + 3; k++) dq[d][k] = mc_slope(vm[k], c0[k], vp[k]);
        }
        const int64_t K = B.ks;
        double *pd = p + NPR * B.off + E.dcell, *sd = ps + 3 * B.off + E.dcell;
#pragma unroll
        for (int k = 0; k < NPR; k++)
            pd[k * K] = c0[k] + 0.25 * sgn[0] * dq[0][k] + 0.25 * sgn[1] * dq[1][k] + 0.25 * sgn[2] * dq[2][k];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // odd fine face along a: halves the coarse cell -> cell-centred field
            const int kk = (sgn[a] > 0.) ? B1 + a : NPR + a;
            double v = c0[kk];
#pragma unroll
            for (int d = 0; d < 3; d++)
                if (d != a) v = v + 0.25 * sgn[d] * dq[d][kk];
            sd[a * K] = v;
        }
        if (pflag) pflag[B.off + E.dcell] = pflag[C.off + E.ccell];
    }
}

// Courant dt of the step: per-block harmonic of the direction minima, minimum
// over blocks (step_ch.c:1313-1322), clamped to 1.3 dt (step_ch.c:974).  Resets
// the minima for the next corrector.  dt[0] <- new dt, dt[1] <- previous dt.
__global__ void k_dt_finalize(int nblk, unsigned long long *__restrict__ dtmin, double *__restrict__ dt)
{
    __shared__ double red[256];
    double m = 1.e9;
    for (int n = threadIdx.x; n < nblk; n += blockDim.x) {
        const double d1 = __longlong_as_double((long long)dtmin[n * 3 + 0]);
        const double d2 = __longlong_as_double((long long)dtmin[n * 3 + 1]);
        const double d3 = __longlong_as_double((long long)dtmin[n * 3 + 2]);
        const double b0 = 1. / (1. / d1 + 1. / d2 + 1. / d3);
        m = mymin(m, b0);
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
// launchers
static inline dim3 gdim(const Launch &L) { return dim3((unsigned)L.grid); }
static inline dim3 bdim(const Launch &L) { return dim3((unsigned)L.threads); }

template <class R>
static void flux_t(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, R *F,
                   unsigned long long *dtmin, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (dir == 1) hipLaunchKernelGGL((k_flux<1, R>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
    else if (dir == 2) hipLaunchKernelGGL((k_flux<2, R>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
    else hipLaunchKernelGGL((k_flux<3, R>), gdim(L), bdim(L), 0, s, P, L.wl, p, ps, F, dtmin);
}

void launch_flux(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, void *F,
                 unsigned long long *dtmin, int mixed, hipStream_t s)
{
    if (mixed) flux_t(P, dir, L, p, ps, (float *)F, dtmin, s);
    else flux_t(P, dir, L, p, ps, (double *)F, dtmin, s);
}

template <class R>
static void emf_t(const KParams &P, const Launch &Lc, const Launch &Lk, const double *p, const R *F1, const R *F2,
                  const R *F3, R *Ec, R *E, hipStream_t s)
{
    if (Lc.wl.n <= 0) return;
    hipLaunchKernelGGL((k_emf_centre<R>), gdim(Lc), bdim(Lc), 0, s, P, Lc.wl, p, Ec);
    hipLaunchKernelGGL((k_emf_corner<R>), gdim(Lk), bdim(Lk), 0, s, P, Lk.wl, Ec, F1, F2, F3, E);
}

void launch_emf(const KParams &P, const Launch &centre, const Launch &corner, const double *p, const void *F1,
                const void *F2, const void *F3, void *Ec, void *E, int mixed, hipStream_t s)
{
    if (mixed)
        emf_t(P, centre, corner, p, (const float *)F1, (const float *)F2, (const float *)F3, (float *)Ec, (float *)E, s);
    else
        emf_t(P, centre, corner, p, (const double *)F1, (const double *)F2, (const double *)F3, (double *)Ec, (double *)E,
              s);
}

void launch_emf_centre(const KParams &P, const Launch &L, const double *p, void *Ec, int mixed, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (mixed) hipLaunchKernelGGL((k_emf_centre<float>), gdim(L), bdim(L), 0, s, P, L.wl, p, (float *)Ec);
    else hipLaunchKernelGGL((k_emf_centre<double>), gdim(L), bdim(L), 0, s, P, L.wl, p, (double *)Ec);
}

void launch_emf_corner(const KParams &P, const Launch &L, const void *Ec, const void *F1, const void *F2,
                       const void *F3, void *E, int mixed, hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (mixed)
        hipLaunchKernelGGL((k_emf_corner<float>), gdim(L), bdim(L), 0, s, P, L.wl, (const float *)Ec, (const float *)F1,
                           (const float *)F2, (const float *)F3, (float *)E);
    else
        hipLaunchKernelGGL((k_emf_corner<double>), gdim(L), bdim(L), 0, s, P, L.wl, (const double *)Ec,
                           (const double *)F1, (const double *)F2, (const double *)F3, (double *)E);
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
    if (mixed)
        hipLaunchKernelGGL((k_ct_update<float>), gdim(L), bdim(L), 0, s, P, L.wl, Dt_val, dt_ptr, dt_mult, psi, psf,
                           (const float *)E);
    else
        hipLaunchKernelGGL((k_ct_update<double>), gdim(L), bdim(L), 0, s, P, L.wl, Dt_val, dt_ptr, dt_mult, psi, psf,
                           (const double *)E);
}

void launch_update_part(const KParams &P, int part, const Launch &L, int full, double Dt_val, const double *dt_ptr,
                        double dt_mult, const double *pi, const double *pb, double *pf, double *U0, double *Unew,
                        const double *psf, const void *F1, const void *F2, const void *F3, int *pflag, int *failcnt,
                        int mixed, hipStream_t s)
{
    if (L.wl.n <= 0) return;
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
    for (int part = 0; part < 3; part++)
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
                  hipStream_t s)
{
    if (L.wl.n <= 0) return;
    if (dir == 1) hipLaunchKernelGGL(k_bound<1>, gdim(L), bdim(L), 0, s, P, L.wl, p, ps, pflag, outflow_only);
    else if (dir == 2) hipLaunchKernelGGL(k_bound<2>, gdim(L), bdim(L), 0, s, P, L.wl, p, ps, pflag, outflow_only);
    else hipLaunchKernelGGL(k_bound<3>, gdim(L), bdim(L), 0, s, P, L.wl, p, ps, pflag, outflow_only);
}

void launch_amr_exchange(const KParams &P, const AmrEntry *e, int64_t n, const AmrProlong *pe, int64_t npro,
                         double *p, double *ps, int *pflag, int threads, int grid, hipStream_t s)
{
    if (n > 0) {
        const int64_t need = (n + threads - 1) / threads;
        const unsigned g = (unsigned)(grid > 0 && grid < need ? grid : need);
        hipLaunchKernelGGL(k_amr_copy, dim3(g), dim3(threads), 0, s, P, e, n, p, ps, pflag);
    }
    if (npro > 0) {
        const int64_t need = (npro + threads - 1) / threads;
        const unsigned g = (unsigned)(grid > 0 && grid < need ? grid : need);
        hipLaunchKernelGGL(k_amr_prolong, dim3(g), dim3(threads), 0, s, P, pe, npro, p, ps, pflag);
    }
}

void launch_dt_finalize(int nblk, unsigned long long *dtmin, double *dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, s, nblk, dtmin, dt);
}

void launch_dt_reset(int nblk, unsigned long long *dtmin, hipStream_t s)
{
    const int n = nblk * 3;
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, dtmin, n,
                       (unsigned long long)0x41CDCD6500000000ULL /* 1e9 */);
}

}  // namespace hamr
