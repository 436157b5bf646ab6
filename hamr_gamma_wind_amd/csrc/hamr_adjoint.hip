// hamr_adjoint.hip -- adjoint of one forward stage of the step (FP64).
//
// Forward stage (the predictor of the step with the full Dt):
//   fluxes F_d = HLL(PPM(P), face B) through every computed face,
//   corner EMF E(P, F) upwinded on the mass-flux signs,
//   face field  B_f' = B_f - Dt curl(E) / gdet_f,
//   U' = U(P) - Dt sum_d (F_d(c + e_d) - F_d(c)) / dx_d + Dt S(P), U'[B] = face average of B_f'.
// For a loss L = sum w . U' the kernels below form dL/dP and dL/dB_f (the face
// field at the start of the stage) in reverse order, recomputing the
// reconstructed states, fluxes and EMF selections from P instead of storing
// them.  The source term is zero in the flat metric the adjoint serves (a
// periodic box of one block); ghost cells are folded back onto the interior
// cells they copy.
#include "hamr_adjoint.h"
#include "hamr_kernels.h"

namespace hamr {

namespace {

struct AdjArgs {
    int D[3], bs[3];
    int64_t K;           // slab length
    double dx[3], Dt, gam;
    const double *gcov, *gcon, *gdet;  // geometry slab 0
    int64_t kcol;
};

__device__ __forceinline__ int64_t cell_of(const AdjArgs &A, int i, int j, int z)
{
    return ((int64_t)i * A.D[1] + j) * A.D[2] + z;
}

// periodic image of a (possibly ghost) index inside the interior
__device__ __forceinline__ int wrap(int i, int bs) { return ((i - NG) % bs + bs) % bs + NG; }

__device__ __forceinline__ int64_t fold(const AdjArgs &A, int i, int j, int z)
{
    return cell_of(A, wrap(i, A.bs[0]), wrap(j, A.bs[1]), wrap(z, A.bs[2]));
}

__device__ __forceinline__ void decode(int64_t t, int e1, int e2, int e3, int &a, int &b, int &c)
{
    c = (int)(t % e3);
    const int64_t q = t / e3;
    b = (int)(q % e2);
    a = (int)(q / e2);
}

__device__ __forceinline__ Geom<double> geom_at(const AdjArgs &A, int pos, int64_t col)
{
    Geom<double> G;
    for (int s = 0; s < 10; s++) {
        G.gcov[s] = A.gcov[((int64_t)pos * 10 + s) * A.kcol + col];
        G.gcon[s] = A.gcon[((int64_t)pos * 10 + s) * A.kcol + col];
    }
    G.g = A.gdet[(int64_t)pos * A.kcol + col];
    return G;
}

__device__ __forceinline__ bool interior(const AdjArgs &A, int i, int j, int z)
{
    return i >= NG && i < NG + A.bs[0] && j >= NG && j < NG + A.bs[1] && z >= NG && z < NG + A.bs[2];
}

}  // namespace

// lam_psf: adjoint of the updated face fields from w[B] (the face average of U'[B]).
__global__ void __launch_bounds__(256) k_adj_seed(AdjArgs A, const double *__restrict__ w, double *__restrict__ lpsf)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)A.D[0] * A.D[1] * A.D[2];
    if (t >= N) return;
    int i, j, z;
    decode(t, A.D[0], A.D[1], A.D[2], i, j, z);
    const int64_t col = (int64_t)i * A.D[1] + j;
    const int64_t K = A.K, si = (int64_t)A.D[1] * A.D[2], sj = A.D[2];
    for (int d = 0; d < 3; d++) {
        const int pos = d == 0 ? FACE1 : d == 1 ? FACE2 : FACE3;
        const int64_t sd = d == 0 ? si : d == 1 ? sj : 1;
        const double gf = A.gdet[(int64_t)pos * A.kcol + col];
        double v = 0.;
        if (interior(A, i, j, z)) v += w[(B1 + d) * K + t];
        const int im = i - (d == 0), jm = j - (d == 1), zm = z - (d == 2);
        if (interior(A, im, jm, zm)) v += w[(B1 + d) * K + t - sd];
        lpsf[d * K + t] = gf * v / 2.;
    }
}

// lam_E from lam_psf (transpose of the face-field update), and the direct part
// of lam_ps (the updated face field carries the old one with weight 1).
__global__ void __launch_bounds__(256) k_adj_ct(AdjArgs A, const double *__restrict__ lpsf, double *__restrict__ lE,
                                                double *__restrict__ lps)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)A.D[0] * A.D[1] * A.D[2];
    if (t >= N) return;
    int i, j, z;
    decode(t, A.D[0], A.D[1], A.D[2], i, j, z);
    const int64_t K = A.K, si = (int64_t)A.D[1] * A.D[2], sj = A.D[2];
    const int64_t col = (int64_t)i * A.D[1] + j;
    const double Dt = A.Dt, dx1 = A.dx[0], dx2 = A.dx[1], dx3 = A.dx[2];
    auto gF = [&](int pos, int64_t c) { return A.gdet[(int64_t)pos * A.kcol + c]; };
    auto L0 = [&](int64_t g) { return lpsf[g]; };
    auto L1 = [&](int64_t g) { return lpsf[K + g]; };
    auto L2 = [&](int64_t g) { return lpsf[2 * K + g]; };
    // faces the forward updates (psf0: i in [NG, NG+bs1], j, z interior; cyclic for 1, 2)
    auto ok0 = [&](int a, int b, int c) { return a >= NG && a <= NG + A.bs[0] && b >= NG && b < NG + A.bs[1] && c >= NG && c < NG + A.bs[2]; };
    auto ok1 = [&](int a, int b, int c) { return a >= NG && a < NG + A.bs[0] && b >= NG && b <= NG + A.bs[1] && c >= NG && c < NG + A.bs[2]; };
    auto ok2 = [&](int a, int b, int c) { return a >= NG && a < NG + A.bs[0] && b >= NG && b < NG + A.bs[1] && c >= NG && c <= NG + A.bs[2]; };
    // direct part (faces on the upper block boundary are ghost copies: folded)
    if (ok0(i, j, z)) atomicAdd(&lps[fold(A, i, j, z)], L0(t));
    if (ok1(i, j, z)) atomicAdd(&lps[K + fold(A, i, j, z)], L1(t));
    if (ok2(i, j, z)) atomicAdd(&lps[2 * K + fold(A, i, j, z)], L2(t));
    if (i < NG || i > NG + A.bs[0] || j < NG || j > NG + A.bs[1] || z < NG || z > NG + A.bs[2]) return;
    double e1 = 0., e2 = 0., e3 = 0.;
    // psf0[g] = (ps0 gF1 - Dt/dx2 (E3[g+sj] - E3[g]) + Dt/dx3 (E2[g+1] - E2[g])) / gF1
    if (ok0(i, j, z)) {
        const double l = L0(t) / gF(FACE1, col);
        e3 += l * Dt / dx2;
        e2 -= l * Dt / dx3;
    }
    if (ok0(i, j - 1, z)) e3 -= L0(t - sj) / gF(FACE1, col - 1) * Dt / dx2;
    if (ok0(i, j, z - 1)) e2 += L0(t - 1) / gF(FACE1, col) * Dt / dx3;
    // psf2[g] = (ps2 gF3 - Dt/dx1 (E2[g+si] - E2[g]) + Dt/dx2 (E1[g+sj] - E1[g])) / gF3
    if (ok2(i, j, z)) {
        const double l = L2(t) / gF(FACE3, col);
        e2 += l * Dt / dx1;
        e1 -= l * Dt / dx2;
    }
    if (ok2(i - 1, j, z)) e2 -= L2(t - si) / gF(FACE3, col - A.D[1]) * Dt / dx1;
    if (ok2(i, j - 1, z)) e1 += L2(t - sj) / gF(FACE3, col - 1) * Dt / dx2;
    // psf1[g] = (ps1 gF2 + Dt/dx1 (E3[g+si] - E3[g]) - Dt/dx3 (E1[g+1] - E1[g])) / gF2
    if (ok1(i, j, z)) {
        const double l = L1(t) / gF(FACE2, col);
        e3 -= l * Dt / dx1;
        e1 += l * Dt / dx3;
    }
    if (ok1(i - 1, j, z)) e3 += L1(t - si) / gF(FACE2, col - A.D[1]) * Dt / dx1;
    if (ok1(i, j, z - 1)) e1 -= L1(t - 1) / gF(FACE2, col) * Dt / dx3;
    lE[t] = e1;
    lE[K + t] = e2;
    lE[2 * K + t] = e3;
}

// lam_F (all components) from the flux divergence of U' (the field components
// of U' are the face averages, not the divergence).
__global__ void __launch_bounds__(256) k_adj_div(AdjArgs A, const double *__restrict__ w, double *__restrict__ lF1,
                                                 double *__restrict__ lF2, double *__restrict__ lF3)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)A.D[0] * A.D[1] * A.D[2];
    if (t >= N) return;
    int i, j, z;
    decode(t, A.D[0], A.D[1], A.D[2], i, j, z);
    const int64_t K = A.K, si = (int64_t)A.D[1] * A.D[2], sj = A.D[2];
    const bool in0 = interior(A, i, j, z);
    const bool in1 = interior(A, i - 1, j, z), in2 = interior(A, i, j - 1, z), in3 = interior(A, i, j, z - 1);
    for (int k = 0; k < NPR; k++) {
        const bool fld = (k == B1 || k == B2 || k == B3);
        const double w0 = (in0 && !fld) ? w[k * K + t] : 0.;
        const double wi = (in1 && !fld) ? w[k * K + t - si] : 0.;
        const double wj = (in2 && !fld) ? w[k * K + t - sj] : 0.;
        const double wz = (in3 && !fld) ? w[k * K + t - 1] : 0.;
        lF1[k * K + t] = A.Dt / (A.dx[0] * 1.0) * (w0 - wi);
        lF2[k * K + t] = A.Dt / (A.dx[1] * 1.0) * (w0 - wj);
        lF3[k * K + t] = A.Dt / (A.dx[2] * 1.0) * (w0 - wz);
    }
}

// lam_P (initialised here) from U(P) of the interior cells.
__global__ void __launch_bounds__(256) k_adj_u0(AdjArgs A, const double *__restrict__ p, const double *__restrict__ w,
                                                double *__restrict__ lP)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)A.D[0] * A.D[1] * A.D[2];
    if (t >= N) return;
    int i, j, z;
    decode(t, A.D[0], A.D[1], A.D[2], i, j, z);
    const int64_t K = A.K;
    double lpr[NPR];
    for (int k = 0; k < NPR; k++) lpr[k] = 0.;
    if (interior(A, i, j, z)) {
        double pr[NPR], lU[NPR];
        for (int k = 0; k < NPR; k++) {
            pr[k] = p[k * K + t];
            lU[k] = (k == B1 || k == B2 || k == B3) ? 0. : w[k * K + t];
        }
        const Geom<double> G = geom_at(A, CENT, (int64_t)i * A.D[1] + j);
        conserved_vjp(pr, G, A.gam, lU, lpr);
    }
    for (int k = 0; k < NPR; k++) lP[k * K + t] = lpr[k];
}

// Corner EMF (upwind selections from the mass fluxes), transposed: scatter
// lam_E into the flux field components and the centre EMFs.
__global__ void __launch_bounds__(256) k_adj_emf(AdjArgs A, const double *__restrict__ F1, const double *__restrict__ F2,
                                                 const double *__restrict__ F3, const double *__restrict__ lE,
                                                 double *lF1, double *lF2, double *lF3, double *lEc)
{
    const int e = A.bs[0] + 1, f = A.bs[1] + 1, h = A.bs[2] + 1;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)e * f * h) return;
    int a, b, c;
    decode(t, e, f, h, a, b, c);
    const int i = a + NG, j = b + NG, z = c + NG;
    const int64_t K = A.K, si = (int64_t)A.D[1] * A.D[2], sj = A.D[2];
    const int64_t g = cell_of(A, i, j, z);
    auto L = [](double x) { return (double)(x <= 0.0); };
    const double q1 = 0.25 * lE[g], q2 = 0.25 * lE[K + g], q3 = 0.25 * lE[2 * K + g];
    const double *F1R = F1 + RHO * K, *F2R = F2 + RHO * K, *F3R = F3 + RHO * K;
    double *lF1B2 = lF1 + B2 * K, *lF1B3 = lF1 + B3 * K, *lF2B1 = lF2 + B1 * K, *lF2B3 = lF2 + B3 * K;
    double *lF3B1 = lF3 + B1 * K, *lF3B2 = lF3 + B2 * K;
    double *lE1c = lEc, *lE2c = lEc + K, *lE3c = lEc + 2 * K;
    {   // E1
        const double aa = L(F2R[g]), bb = L(F2R[g - 1]), cc = L(F3R[g]), dd = L(F3R[g - sj]);
        atomicAdd(&lF2B3[g], q1 * (-1. - cc - dd));
        atomicAdd(&lF2B3[g - 1], q1 * (-1. - (1. - cc) - (1. - dd)));
        atomicAdd(&lF3B2[g], q1 * (aa + bb + 1.));
        atomicAdd(&lF3B2[g - sj], q1 * ((1. - aa) + (1. - bb) + 1.));
        atomicAdd(&lE1c[g], q1 * (-aa - cc));
        atomicAdd(&lE1c[g - sj], q1 * (-(1. - aa) - dd));
        atomicAdd(&lE1c[g - 1], q1 * (-bb - (1. - cc)));
        atomicAdd(&lE1c[g - sj - 1], q1 * (-(1. - bb) - (1. - dd)));
    }
    {   // E2
        const double aa = L(F3R[g]), bb = L(F3R[g - si]), cc = L(F1R[g]), dd = L(F1R[g - 1]);
        atomicAdd(&lF3B1[g], q2 * (-1. - cc - dd));
        atomicAdd(&lF3B1[g - si], q2 * (-1. - (1. - cc) - (1. - dd)));
        atomicAdd(&lF1B3[g], q2 * (aa + bb + 1.));
        atomicAdd(&lF1B3[g - 1], q2 * ((1. - aa) + (1. - bb) + 1.));
        atomicAdd(&lE2c[g], q2 * (-aa - cc));
        atomicAdd(&lE2c[g - 1], q2 * (-(1. - aa) - dd));
        atomicAdd(&lE2c[g - si], q2 * (-bb - (1. - cc)));
        atomicAdd(&lE2c[g - si - 1], q2 * (-(1. - bb) - (1. - dd)));
    }
    {   // E3
        const double aa = L(F2R[g]), bb = L(F2R[g - si]), cc = L(F1R[g]), dd = L(F1R[g - sj]);
        atomicAdd(&lF2B1[g], q3 * (1. + cc + dd));
        atomicAdd(&lF2B1[g - si], q3 * (1. + (1. - cc) + (1. - dd)));
        atomicAdd(&lF1B2[g], q3 * (-aa - bb - 1.));
        atomicAdd(&lF1B2[g - sj], q3 * (-(1. - aa) - (1. - bb) - 1.));
        atomicAdd(&lE3c[g], q3 * (-aa - cc));
        atomicAdd(&lE3c[g - sj], q3 * (-(1. - aa) - dd));
        atomicAdd(&lE3c[g - si], q3 * (-bb - (1. - cc)));
        atomicAdd(&lE3c[g - si - sj], q3 * (-(1. - bb) - (1. - dd)));
    }
}

// Centre EMF of the cells [NG-1, NG+BS], transposed into lam_P (folded).
__global__ void __launch_bounds__(256) k_adj_ecent(AdjArgs A, const double *__restrict__ p, const double *__restrict__ lEc,
                                                   double *lP)
{
    const int e = A.bs[0] + 2, f = A.bs[1] + 2, h = A.bs[2] + 2;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)e * f * h) return;
    int a, b, c;
    decode(t, e, f, h, a, b, c);
    const int i = a + NG - 1, j = b + NG - 1, z = c + NG - 1;
    const int64_t K = A.K, g = cell_of(A, i, j, z);
    const double lE[3] = {lEc[g], lEc[K + g], lEc[2 * K + g]};
    if (lE[0] == 0. && lE[1] == 0. && lE[2] == 0.) return;
    double pr[NPR], lpr[NPR];
    for (int k = 0; k < NPR; k++) {
        pr[k] = p[k * K + g];
        lpr[k] = 0.;
    }
    const Geom<double> G = geom_at(A, CENT, (int64_t)i * A.D[1] + j);
    ecent_vjp(pr, G, lE, lpr);
    const int64_t fg = fold(A, i, j, z);
    for (int k = 0; k < NPR; k++)
        if (lpr[k] != 0.) atomicAdd(&lP[k * K + fg], lpr[k]);
}

// One face of direction DIR: HLL, the two face states and the PPM
// reconstruction of every primitive, transposed into lam_P (folded) and the
// normal face field into lam_ps.  Held at two waves per SIMD (a few registers
// spilled): 1.3x faster than one wave with none spilled.
template <int DIR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_adj_flux(AdjArgs A, const double *__restrict__ p, const double *__restrict__ ps,
                                                  const double *__restrict__ lF, double *lP, double *lps)
{
    const int e1 = A.bs[0] + (DIR == 1 ? 1 : 2), e2 = A.bs[1] + (DIR == 2 ? 1 : 2), e3 = A.bs[2] + (DIR == 3 ? 1 : 2);
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)e1 * e2 * e3) return;
    int a, b, c;
    decode(t, e1, e2, e3, a, b, c);
    const int i = a + (DIR == 1 ? NG : NG - 1), j = b + (DIR == 2 ? NG : NG - 1), z = c + (DIR == 3 ? NG : NG - 1);
    const int64_t K = A.K, cell = cell_of(A, i, j, z);
    double lf[NPR];
    bool any = false;
    for (int k = 0; k < NPR; k++) {
        lf[k] = lF[k * K + cell];
        any |= (lf[k] != 0.);
    }
    if (!any) return;
    const int64_t sd = (DIR == 1) ? (int64_t)A.D[1] * A.D[2] : (DIR == 2) ? (int64_t)A.D[2] : 1;
    // reconstruction (primal values only: the Jacobians are formed per
    // primitive at the end, so that they are not live across the HLL part)
    double pl[NPR], pr[NPR];
    for (int k = 0; k < NPR; k++) {
        const double *q = p + k * K + cell;
        const double xm[5] = {q[-3 * sd], q[-2 * sd], q[-sd], q[0], q[sd]};
        const double xc[5] = {q[-2 * sd], q[-sd], q[0], q[sd], q[2 * sd]};
        double l, r, jl[5], jr[5];
        para_jacobian(xm, &l, &r, jl, jr);
        pl[k] = r;
        para_jacobian(xc, &l, &r, jl, jr);
        pr[k] = l;
    }
    const double bn = ps[(DIR - 1) * K + cell];
    pl[B1 + DIR - 1] = bn;
    pr[B1 + DIR - 1] = bn;
    const int pos = DIR == 1 ? FACE1 : DIR == 2 ? FACE2 : FACE3;
    const Geom<double> G = geom_at(A, pos, (int64_t)i * A.D[1] + j);
    // primal HLL
    double FL[NPR], UL[NPR], FR[NPR], UR[NPR], vxl = 0., vnl = 0., vxr = 0., vnr = 0., dum;
    {
        State<double> q;
        get_state(pl, G, q);
        primtoflux(pl, q, DIR, G, A.gam, FL, &vxl, &vnl);
        primtoflux(pl, q, 0, G, A.gam, UL, &dum, &dum);
        get_state(pr, G, q);
        primtoflux(pr, q, DIR, G, A.gam, FR, &vxr, &vnr);
        primtoflux(pr, q, 0, G, A.gam, UR, &dum, &dum);
    }
    const double m1 = mymax(0., vxl), cmax = fabs(mymax(m1, vxr));
    const double n1 = mymax(0., -vnl), cmin = fabs(mymax(n1, -vnr));
    const double den = cmax + cmin + kSMALL;
    // reverse of F = (cmax FL + cmin FR - cmax cmin (UR - UL)) / den
    double lFL[NPR], lFR[NPR], lUL[NPR], lUR[NPR], lcmax = 0., lcmin = 0.;
    for (int k = 0; k < NPR; k++) {
        const double F = (cmax * FL[k] + cmin * FR[k] - cmax * cmin * (UR[k] - UL[k])) / den;
        lFL[k] = lf[k] * cmax / den;
        lFR[k] = lf[k] * cmin / den;
        lUR[k] = -lf[k] * cmax * cmin / den;
        lUL[k] = -lUR[k];
        lcmax += lf[k] * ((FL[k] - cmin * (UR[k] - UL[k])) - F) / den;
        lcmin += lf[k] * ((FR[k] - cmax * (UR[k] - UL[k])) - F) / den;
    }
    double lvxl = 0., lvxr = 0., lvnl = 0., lvnr = 0.;
    if (m1 > vxr) { if (!(0. > vxl)) lvxl = lcmax; }
    else lvxr = lcmax;
    if (n1 > -vnr) { if (!(0. > -vnl)) lvnl = -lcmin; }
    else lvnr = -lcmin;
    double lpl[NPR], lpr[NPR];
    for (int k = 0; k < NPR; k++) lpl[k] = lpr[k] = 0.;
    face_state_vjp(pl, G, DIR, A.gam, lFL, lUL, lvxl, lvnl, lpl);
    face_state_vjp(pr, G, DIR, A.gam, lFR, lUR, lvxr, lvnr, lpr);
    // normal face field
    {
        const double v = lpl[B1 + DIR - 1] + lpr[B1 + DIR - 1];
        if (v != 0.) atomicAdd(&lps[(DIR - 1) * K + fold(A, i, j, z)], v);
        lpl[B1 + DIR - 1] = lpr[B1 + DIR - 1] = 0.;
    }
    // reconstruction: pl from cells c-3..c+1, pr from c-2..c+2 along DIR
    int64_t fg[6];
    for (int m = -3; m <= 2; m++)
        fg[m + 3] = fold(A, i + (DIR == 1 ? m : 0), j + (DIR == 2 ? m : 0), z + (DIR == 3 ? m : 0));
    for (int k = 0; k < NPR; k++) {
        if (lpl[k] == 0. && lpr[k] == 0.) continue;
        const double *q = p + k * K + cell;
        const double xm[5] = {q[-3 * sd], q[-2 * sd], q[-sd], q[0], q[sd]};
        const double xc[5] = {q[-2 * sd], q[-sd], q[0], q[sd], q[2 * sd]};
        double l, r, jl[5], Jr[5], Jl[5];
        para_jacobian(xm, &l, &r, jl, Jr);
        para_jacobian(xc, &l, &r, Jl, jl);
        for (int m = -3; m <= 2; m++) {
            double v = 0.;
            if (m <= 1) v += lpl[k] * Jr[m + 3];
            if (m >= -2) v += lpr[k] * Jl[m + 2];
            if (v != 0.) atomicAdd(&lP[k * K + fg[m + 3]], v);
        }
    }
}

// ---------------------------------------------------------------------------
static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Zeroes n doubles of each of a, b and c (the accumulators of the adjoint).  A
// kernel rather than hipMemsetAsync, so that a caller may capture the adjoint
// in a HIP graph.
__global__ void __launch_bounds__(256) k_adj_zero3(double *a, double *b, double *c, int64_t n)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    a[t] = 0.;
    b[t] = 0.;
    c[t] = 0.;
}

void launch_adjoint(const KParams &P, const BlockDev &B, double Dt, const double *p, const double *ps,
                    const double *F1, const double *F2, const double *F3, const double *w, double *lP, double *lps,
                    double *lpsf, double *lE, double *lEc, double *lF1, double *lF2, double *lF3, hipStream_t s)
{
    AdjArgs A;
    for (int k = 0; k < 3; k++) {
        A.D[k] = B.D[k];
        A.bs[k] = B.bs[k];
        A.dx[k] = B.dx[k];
    }
    A.K = B.ks;
    A.Dt = Dt;
    A.gam = P.gam;
    A.kcol = B.kcol;
    A.gcov = P.gcov + B.moff * NPOS * 10;
    A.gcon = P.gcon + B.moff * NPOS * 10;
    A.gdet = P.gdet + B.moff * NPOS;
    const int64_t N = (int64_t)A.D[0] * A.D[1] * A.D[2];
    hipLaunchKernelGGL(k_adj_zero3, g1(3 * A.K), dim3(256), 0, s, lps, lE, lEc, 3 * A.K);
    hipLaunchKernelGGL(k_adj_seed, g1(N), dim3(256), 0, s, A, w, lpsf);
    hipLaunchKernelGGL(k_adj_ct, g1(N), dim3(256), 0, s, A, lpsf, lE, lps);
    hipLaunchKernelGGL(k_adj_div, g1(N), dim3(256), 0, s, A, w, lF1, lF2, lF3);
    hipLaunchKernelGGL(k_adj_u0, g1(N), dim3(256), 0, s, A, p, w, lP);
    const int64_t nc = (int64_t)(A.bs[0] + 1) * (A.bs[1] + 1) * (A.bs[2] + 1);
    hipLaunchKernelGGL(k_adj_emf, g1(nc), dim3(256), 0, s, A, F1, F2, F3, lE, lF1, lF2, lF3, lEc);
    const int64_t ne = (int64_t)(A.bs[0] + 2) * (A.bs[1] + 2) * (A.bs[2] + 2);
    hipLaunchKernelGGL(k_adj_ecent, g1(ne), dim3(256), 0, s, A, p, lEc, lP);
    const int64_t f1 = (int64_t)(A.bs[0] + 1) * (A.bs[1] + 2) * (A.bs[2] + 2);
    const int64_t f2 = (int64_t)(A.bs[0] + 2) * (A.bs[1] + 1) * (A.bs[2] + 2);
    const int64_t f3 = (int64_t)(A.bs[0] + 2) * (A.bs[1] + 2) * (A.bs[2] + 1);
    hipLaunchKernelGGL(k_adj_flux<1>, g1(f1), dim3(256), 0, s, A, p, ps, lF1, lP, lps);
    hipLaunchKernelGGL(k_adj_flux<2>, g1(f2), dim3(256), 0, s, A, p, ps, lF2, lP, lps);
    hipLaunchKernelGGL(k_adj_flux<3>, g1(f3), dim3(256), 0, s, A, p, ps, lF3, lP, lps);
}

}  // namespace hamr
