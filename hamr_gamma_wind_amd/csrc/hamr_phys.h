// hamr_phys.h -- pointwise GRMHD physics for the gfx950 step kernels.
//
// Ideal-gas MHD in modified Kerr-Schild code coordinates: 4-velocity and
// magnetic 4-vector from the primitives, fluxes and conserved variables,
// magnetosonic speeds, geometric source terms, PPM reconstruction and the
// Noble et al. (2006) 2D / 1D-v^2 / 1D inversions with the reference's return
// codes.  Expressions keep the reference's association order (the FP64 kernels
// are built without FMA contraction) so that results agree with the reference's
// GPU arithmetic to the last ulps.  Reference: GPU_phys_mhd.h, GPU_metric.h,
// GPU_interp.h, GPU_utoprim_2d.h, GPU_utoprim_1dvsq2fix1.h, GPU_utoprim_1dfix1.h.
#pragma once

#include <hip/hip_runtime.h>

#define HD __host__ __device__ __forceinline__

namespace hamr {

enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7, KTOT = 8, NPR = 9 };
enum { FACE1 = 0, FACE2 = 1, CORN = 2, CENT = 3, FACE3 = 4, NPOS = 5 };

constexpr double kSMALL = 1.e-20;
constexpr double kNEWT_TOL = 1.0e-10;
constexpr double kMIN_NEWT_TOL = 1.0e-10;
constexpr int kMAX_NEWT_ITER = 30;
constexpr int kEXTRA_NEWT_ITER = 2;
constexpr double kNEWT_TOL2 = 1.0e-15;
constexpr double kMIN_NEWT_TOL2 = 1.0e-10;
constexpr double kW_TOO_BIG = 1.e20;
constexpr double kUTSQ_TOO_BIG = 1.e20;
constexpr double kFAIL_VAL = 1.e30;
constexpr double kGAMMAMAX = 80.;

template <class T> HD T mymax(T a, T b) { return (a > b) ? a : b; }
template <class T> HD T mymin(T a, T b) { return (a < b) ? a : b; }

// Division, reciprocal square root and power in the arithmetic of R: IEEE
// double operations for the FP64 step (the reference's), the hardware's
// single-precision reciprocal, reciprocal square root and exp2/log2 (about
// one ulp each) for the FP32 path.
template <class R> HD R dv(R a, R b) { return a / b; }
template <class R> HD R rsq(R a) { return R(1.) / sqrt(a); }
template <class R> HD R rpow(R a, R e) { return pow(a, e); }
#if defined(__HIP_DEVICE_COMPILE__)
template <> HD float dv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> HD float rsq<float>(float a) { return __builtin_amdgcn_rsqf(a); }
template <> HD float rpow<float>(float a, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(a)); }
#endif

// Local geometry: packed symmetric gcov/gcon (00,01,02,03,11,12,13,22,23,33) and sqrt(-g).
// oma = 1 - alpha (lapse alpha = 1/sqrt(-g^00)), formed in FP64 by the FP32 path only.
template <class R> struct Geom { R gcov[10], gcon[10], g, oma; };
// opu = 1 + u_t, formed without cancellation by the FP32 path only.
template <class R> struct State { R ucon[4], ucov[4], bcon[4], bcov[4], opu; };

// 1 + u_t for the FP32 path.  With u_t = -alpha gamma + beta_i U^i (beta_i = g_0i,
// U^i the relative 4-velocity), 1 + u_t = (1 - alpha) - alpha (gamma - 1)
// + beta_i U^i and gamma - 1 = q^2 / (gamma + 1): every term is small where
// 1 + u_t is, so FP32 keeps its relative precision (the direct sum of
// T^i_t and rho u^i in the energy flux cancels the rest-mass part).
template <class R> HD R one_plus_ut(const R *pr, const Geom<R> &G)
{
    R qsq = G.gcov[4] * pr[U1] * pr[U1] + G.gcov[7] * pr[U2] * pr[U2] + G.gcov[9] * pr[U3] * pr[U3]
        + R(2.) * (G.gcov[5] * pr[U1] * pr[U2] + G.gcov[6] * pr[U1] * pr[U3] + G.gcov[8] * pr[U2] * pr[U3]);
    if (qsq < R(0.)) qsq = (fabs(qsq) > R(1.E-10)) ? R(0.) : R(1.E-10);
    const R gamma = sqrt(R(1.) + qsq);
    const R alpha = R(1.) - G.oma;
    return G.oma - alpha * dv(qsq, gamma + R(1.)) + (G.gcov[1] * pr[U1] + G.gcov[2] * pr[U2] + G.gcov[3] * pr[U3]);
}

// GPU_metric.h:3-34
template <class R> HD void lower10(const R *u, const R *g, R *o)
{
    o[0] = g[0] * u[0] + g[1] * u[1] + g[2] * u[2] + g[3] * u[3];
    o[1] = g[1] * u[0] + g[4] * u[1] + g[5] * u[2] + g[6] * u[3];
    o[2] = g[2] * u[0] + g[5] * u[1] + g[7] * u[2] + g[8] * u[3];
    o[3] = g[3] * u[0] + g[6] * u[1] + g[8] * u[2] + g[9] * u[3];
}

// GPU_phys_mhd.h:755
template <class R> HD int gamma_calc(const R *pr, const R *gcov, R *gamma)
{
    R qsq = gcov[4] * pr[U1] * pr[U1] + gcov[7] * pr[U2] * pr[U2] + gcov[9] * pr[U3] * pr[U3]
        + R(2.) * (gcov[5] * pr[U1] * pr[U2] + gcov[6] * pr[U1] * pr[U3] + gcov[8] * pr[U2] * pr[U3]);
    if (qsq < R(0.)) {
        if (fabs(qsq) > R(1.E-10)) { *gamma = R(1.); return 1; }
        qsq = R(1.E-10);
    }
    *gamma = sqrt(R(1.) + qsq);
    return 0;
}

// GPU_phys_mhd.h:714
template <class R> HD void ucon_calc(const R *pr, const Geom<R> &G, R *ucon)
{
    R alpha, gamma, beta[4];
    alpha = rsq(-G.gcon[0]);
    for (int j = 1; j < 4; j++) beta[j] = G.gcon[j] * alpha * alpha;
    gamma_calc(pr, G.gcov, &gamma);
    ucon[0] = dv(gamma, alpha);
    for (int j = 1; j < 4; j++) ucon[j] = pr[U1 + j - 1] - dv(gamma * beta[j], alpha);
}

// GPU_phys_mhd.h:738
template <class R> HD void bcon_calc(const R *pr, const R *ucon, const R *ucov, R *bcon)
{
    bcon[0] = pr[B1] * ucov[1] + pr[B2] * ucov[2] + pr[B3] * ucov[3];
    for (int j = 1; j < 4; j++) bcon[j] = dv(pr[B1 - 1 + j] + bcon[0] * ucon[j], ucon[0]);
}

// GPU_phys_mhd.h:692
template <class R> HD void get_state(const R *pr, const Geom<R> &G, State<R> &q)
{
    ucon_calc(pr, G, q.ucon);
    lower10(q.ucon, G.gcov, q.ucov);
    bcon_calc(pr, q.ucon, q.ucov, q.bcon);
    lower10(q.bcon, G.gcov, q.bcov);
    if constexpr (sizeof(R) == 4) q.opu = one_plus_ut(pr, G);
}

template <class R> HD R dot4(const R *a, const R *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

// Fast magnetosonic speeds along dir (the vchar part of GPU_phys_mhd.h:550-617)
// given w = rho + P + u, bsq and eta = w + bsq of the state.
template <class R> HD void vchar_coord(const R *pr, const State<R> &q, int dir, const Geom<R> &G, R gam, R w,
                                       R bsq, R eta, R *vmax, R *vmin)
{
    R discr, vp, vm, va2, cs2, cms2, Acon_0, Acon_js, Asq, Bsq, Au, Bu, AB, Au2, Bu2, AuBu, A, B, C;
    if (dir == 1) { Acon_0 = G.gcon[1]; Acon_js = G.gcon[4]; }
    else if (dir == 2) { Acon_0 = G.gcon[2]; Acon_js = G.gcon[7]; }
    else { Acon_0 = G.gcon[3]; Acon_js = G.gcon[9]; }
    cs2 = dv(gam * (gam - R(1.)) * pr[UU], w);
    va2 = dv(bsq, eta);
    cms2 = cs2 + va2 - cs2 * va2;
    if (cms2 < R(0.)) cms2 = R(kSMALL);
    if (cms2 > R(1.)) cms2 = R(1.);
    Asq = Acon_js;
    Bsq = G.gcon[0];
    Au = q.ucon[dir];
    Bu = q.ucon[0];
    AB = Acon_0;
    Au2 = Au * Au;
    Bu2 = Bu * Bu;
    AuBu = Au * Bu;
    A = Bu2 - (Bsq + Bu2) * cms2;
    B = R(2.) * (AuBu - (AB + AuBu) * cms2);
    C = Au2 - (Asq + Au2) * cms2;
    discr = B * B - R(4.) * A * C;
    if ((discr < R(0.0)) && (discr > R(-1.e-10))) discr = R(0.0);
    else if (discr < R(-1.e-10)) discr = R(0.);
    discr = sqrt(discr);
    vp = dv(-(-B + discr), R(2.) * A);
    vm = dv(-(-B - discr), R(2.) * A);
    *vmax = mymax(vp, vm);
    *vmin = mymin(vp, vm);
}

// The standalone vchar (GPU_phys_mhd.h:779-885): the same speeds with the
// enthalpy formed as w = rho + gam u.
template <class R> HD void vchar_state(const R *pr, const State<R> &q, int dir, const Geom<R> &G, R gam, R *vmax,
                                       R *vmin)
{
    const R w = pr[RHO] + gam * pr[UU];
    const R bsq = dot4(q.bcon, q.bcov);
    vchar_coord(pr, q, dir, G, gam, w, bsq, w + bsq, vmax, vmin);
}

// GPU_phys_mhd.h:432-623 (ideal gas).  Speeds are written only for dir != 0.
template <class R> HD void primtoflux(const R *pr, const State<R> &q, int dir, const Geom<R> &G, R gam,
                                      R *flux, R *vmax, R *vmin)
{
    R mhd[4], P, u, w, bsq, eta, ptot;
    u = pr[UU];
    P = (gam - R(1.)) * u;
    w = pr[RHO] + P + u;
    bsq = dot4(q.bcon, q.bcov);
    eta = w + bsq;
    ptot = P + R(0.5) * bsq;
    flux[RHO] = pr[RHO] * q.ucon[dir];
    for (int j = 0; j < 4; j++)
        mhd[j] = eta * q.ucon[dir] * q.ucov[j] + ptot * (dir == j ? R(1.) : R(0.)) - q.bcon[dir] * q.bcov[j];
    if constexpr (sizeof(R) == 4)  // same quantity, without the rest-mass cancellation
        flux[UU] = flux[RHO] * q.opu + (u + P + bsq) * q.ucon[dir] * q.ucov[0] + ptot * (dir == 0 ? R(1.) : R(0.))
            - q.bcon[dir] * q.bcov[0];
    else
        flux[UU] = mhd[0] + flux[RHO];
    flux[U1] = mhd[1];
    flux[U2] = mhd[2];
    flux[U3] = mhd[3];
    flux[B1] = q.bcon[1] * q.ucon[dir] - q.bcon[dir] * q.ucon[1];
    flux[B2] = q.bcon[2] * q.ucon[dir] - q.bcon[dir] * q.ucon[2];
    flux[B3] = q.bcon[3] * q.ucon[dir] - q.bcon[dir] * q.ucon[3];
    {
        R Pe = (gam - R(1.0)) * pr[UU];
        R entr = Pe * rpow(pr[RHO], -gam);
        flux[KTOT] = flux[RHO] * entr;
    }
    for (int k = 0; k < NPR; k++) flux[k] *= G.g;
    if (dir != 0) vchar_coord(pr, q, dir, G, gam, w, bsq, eta, vmax, vmin);
}

// GPU_phys_mhd.h:34-197 (ideal gas): dU_nu = gdet T^kap_lam Gamma^lam_{nu kap}.
// conn(lam, s) returns Gamma^lam for the packed pair s of (nu, kap).
template <class R, class ConnF> HD void source(const R *ph, const State<R> &q, ConnF conn, R g, R gam, R *dU)
{
    R mhd[4][4], c, P, u, w, bsq, eta, ptot;
    P = (gam - R(1.)) * ph[UU];
    u = ph[UU];
    w = P + ph[RHO] + u;
    bsq = dot4(q.bcon, q.bcov);
    eta = w + bsq;
    ptot = P + R(0.5) * bsq;
    for (int dir = 0; dir < 4; dir++)
        for (int j = 0; j < 4; j++)
            mhd[dir][j] = eta * q.ucon[dir] * q.ucov[j] + ptot * (dir == j ? R(1.) : R(0.)) - q.bcon[dir] * q.bcov[j];
    for (int k = 0; k < NPR; k++) dU[k] = R(0.);
    for (int k = 0; k < 4; k++) {
        dU[UU] += mhd[0][k] * conn(k, 0);
        dU[U1] += mhd[1][k] * conn(k, 4);
        dU[U2] += mhd[2][k] * conn(k, 7);
        dU[U3] += mhd[3][k] * conn(k, 9);
        c = conn(k, 1); dU[UU] += mhd[1][k] * c; dU[U1] += mhd[0][k] * c;
        c = conn(k, 2); dU[UU] += mhd[2][k] * c; dU[U2] += mhd[0][k] * c;
        c = conn(k, 3); dU[UU] += mhd[3][k] * c; dU[U3] += mhd[0][k] * c;
        c = conn(k, 5); dU[U1] += mhd[2][k] * c; dU[U2] += mhd[1][k] * c;
        c = conn(k, 6); dU[U1] += mhd[3][k] * c; dU[U3] += mhd[1][k] * c;
        c = conn(k, 8); dU[U2] += mhd[3][k] * c; dU[U3] += mhd[2][k] * c;
    }
    dU[UU] *= g;
    dU[U1] *= g;
    dU[U2] *= g;
    dU[U3] *= g;
    dU[KTOT] *= g;
}

// MC slope with factor 2 (GPU_interp.h:2-25): the minmod of 2 dq_-, 2 dq_+ and
// the centred difference, written with min and copysign (the same value as
// the reference's selects, up to the sign of a zero slope).
template <class R> HD R mc_slope(R ym, R y0, R yp)
{
    const R Dqm = R(2.) * (y0 - ym), Dqp = R(2.) * (yp - y0), Dqc = R(0.5) * (yp - ym);
    const R m = fmin(fabs(Dqm), fabs(Dqp));
    const R v = copysign(fmin(fabs(Dqc), m), Dqc);
    return (Dqm * Dqp > R(0.)) ? v : R(0.);
}

// PPM interface value between cells with values ym, yp and MC slopes dqm, dqp
// (GPU_interp.h:89-90): a_{c+1/2} = (y_c + y_{c+1})/2 - (dq_{c+1} - dq_c)/6.
template <class R> HD R ppm_face(R ym, R yp, R dqm, R dqp)
{
    return R(0.5) * (yp + ym) - dv(dqp - dqm, R(6.0));
}

// CW84 monotonicity limits of the parabola of a cell with mean x3 and interface
// values l, r (GPU_interp.h:92-108).
template <class R> HD void ppm_limit(R x3, R l, R r, R &lout, R &rout)
{
    const R qa = (r - x3) * (x3 - l);
    const R qd = (r - l);
    const R qe = R(6.0) * (x3 - R(0.5) * (l + r));
    if (qa <= R(0.)) { l = x3; r = x3; }
    if (qd * (qd - qe) < R(0.0)) l = R(3.0) * x3 - R(2.0) * r;
    else if (qd * (qd + qe) < R(0.0)) r = R(3.0) * x3 - R(2.0) * l;
    lout = l;
    rout = r;
}

// GPU_interp.h:60-111: CW84 PPM with MC (factor 2) slopes; l = a_L, r = a_R.
template <class R> HD void para(R x1, R x2, R x3, R x4, R x5, R &lout, R &rout)
{
    const R dq1 = mc_slope(x1, x2, x3), dq2 = mc_slope(x2, x3, x4), dq3 = mc_slope(x3, x4, x5);
    ppm_limit(x3, ppm_face(x2, x3, dq1, dq2), ppm_face(x3, x4, dq2, dq3), lout, rout);
}

// ---------------------------------------------------------------------------
// Noble 2006 inversions (always FP64: the Newton solve needs it).
struct InvPre {
    double Bcon[4], Qcon[4], ncon[4];
    double Bsq, QdotB, QdotBsq, Qtsq, Qdotn, D;
};

HD void inv_prepare(const double *Ut, const double *gcov, const double *gcon, InvPre &s)
{
    double Bcov[4], Qcov[4], ncov, Qsq;
    s.Bcon[0] = 0.;
    for (int i = 1; i < 4; i++) s.Bcon[i] = Ut[B1 + i - 1];
    lower10(s.Bcon, gcov, Bcov);
    for (int i = 0; i < 4; i++) Qcov[i] = Ut[1 + i];
    lower10(Qcov, gcon, s.Qcon);
    s.Bsq = 0.;
    for (int i = 1; i < 4; i++) s.Bsq += s.Bcon[i] * Bcov[i];
    s.QdotB = 0.;
    for (int i = 0; i < 4; i++) s.QdotB += Qcov[i] * s.Bcon[i];
    s.QdotBsq = s.QdotB * s.QdotB;
    ncov = -sqrt(-1. / gcon[0]);
    s.ncon[0] = gcon[0] * ncov;
    s.ncon[1] = gcon[1] * ncov;
    s.ncon[2] = gcon[2] * ncov;
    s.ncon[3] = gcon[3] * ncov;
    s.Qdotn = s.Qcon[0] * ncov;
    Qsq = 0.;
    for (int i = 0; i < 4; i++) Qsq += Qcov[i] * s.Qcon[i];
    s.Qtsq = Qsq + s.Qdotn * s.Qdotn;
    s.D = Ut[RHO];
}

HD double utsq_of(const double *prim, const double *gcov)
{
    double utsq = gcov[4] * prim[U1] * prim[U1];
    utsq += 2. * gcov[5] * prim[U2] * prim[U1];
    utsq += 2. * gcov[6] * prim[U3] * prim[U1];
    utsq += gcov[7] * prim[U2] * prim[U2];
    utsq += 2 * gcov[8] * prim[U3] * prim[U2];
    utsq += gcov[9] * prim[U3] * prim[U3];
    return utsq;
}

HD void set_utcon(const InvPre &s, double gamma, double W, double *prim)
{
    for (int i = 1; i < 4; i++) {
        double Qtcon = s.Qcon[i] + s.ncon[i] * s.Qdotn;
        prim[U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon + s.QdotB * s.Bcon[i] / W);
    }
}

HD double vsq_calc(double W, double Bsq, double Qtsq, double QdotBsq)
{
    double Wsq = W * W, Xsq = (Bsq + W) * (Bsq + W);
    return ((Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * Xsq));
}

// GPU_utoprim_2d.h general_newton_raphson + func_vsq + validate_x
HD int gnr_2d(double *x, const InvPre &s, double gam)
{
    double dx[2], x_old[2], errx = 1., f = 1., df = 1.;
    int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
    const double Bsq = s.Bsq, Qtsq = s.Qtsq, QdotBsq = s.QdotBsq, Qdotn = s.Qdotn, D = s.D;
    while (keep) {
        double Wsq, p_tmp, dPdvsq, dPdW, gtmp, t11, t16, t18, t2, t21, t23, t24, t25, t3, t35, t36, t4, t40, t9;
        Wsq = x[0] * x[0];
        gtmp = 1. - x[1];
        p_tmp = (gam - 1.) * (fma(x[0], gtmp, -D * sqrt(gtmp))) / gam;
        dPdW = (gam - 1.) * (1. - x[1]) / gam;
        dPdvsq = (gam - 1.) * (fma(0.5, D / sqrt(1. - x[1]), -x[0])) / gam;
        t2 = -0.5 * Bsq + dPdvsq;
        t3 = Bsq + x[0];
        t4 = t3 * t3;
        t9 = 1. / Wsq;
        t11 = Qtsq - x[1] * t4 + QdotBsq * (Bsq + 2.0 * x[0]) * t9;
        t16 = QdotBsq * t9;
        t18 = -Qdotn - 0.5 * Bsq * (1.0 + x[1]) + 0.5 * t16 - x[0] + p_tmp;
        t21 = 1. / t3;
        t23 = 1. / x[0];
        t24 = t16 * t23;
        t25 = -1.0 + dPdW - t24;
        t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + x[1] * Wsq * x[0]) * t9 * t23;
        t36 = 1. / t35;
        dx[0] = -(t2 * t11 + t4 * t18) * t21 * t36;
        t40 = (x[1] + t24) * t3;
        dx[1] = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
        df = -t11 * t11 - t18 * t18;  // resid = (t11, t18); f = |resid|^2 / 2 (utoprim_2d.c:635-640)
        f = -0.5 * df;
        x_old[0] = x[0];
        x_old[1] = x[1];
        x[0] += dx[0];
        x[1] += dx[1];
        x[0] = fabs(x[0]);
        x[0] = (x[0] > kW_TOO_BIG) ? x_old[0] : x[0];
        x[1] = (x[1] < 0.) ? 0. : x[1];
        x[1] = (x[1] > 1.) ? (1. - 1.e-15) : x[1];
        errx = (x[0] == 0.) ? fabs(x[0] - x_old[0]) : fabs((x[0] - x_old[0]) / x[0]);
        if ((fabs(errx) <= kNEWT_TOL) && (doing_extra == 0) && (kEXTRA_NEWT_ITER > 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= kNEWT_TOL) && (doing_extra == 0)) || (i_extra > kEXTRA_NEWT_ITER) || (n_iter >= (kMAX_NEWT_ITER - 1))) keep = 0;
        n_iter++;
    }
    // untrapped divergences of the last residual (utoprim_2d.c:511)
    if ((isfinite(f) == 0) || (isfinite(df) == 0)) return 2;
    if (fabs(errx) > mymin(kNEWT_TOL, kMIN_NEWT_TOL)) return 1;
    return 0;
}

// GPU_utoprim_2d.h Utoprim_new_body
HD int body_2d(const double *Ut, const double *gcov, const double *gcon, double *prim, double gam)
{
    InvPre s;
    double x[2], utsq, gammasq, gamma, rho0, u, p, w, W_last, W, vsq, gtmp, xv;
    int i_increase, retval;
    inv_prepare(Ut, gcov, gcon, s);
    utsq = utsq_of(prim, gcov);
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > kUTSQ_TOO_BIG) return 2;
    gammasq = 1. + utsq;
    gamma = sqrt(gammasq);
    rho0 = s.D / gamma;
    u = prim[UU];
    p = (gam - 1.) * u;
    w = rho0 + u + p;
    W_last = w * gammasq;
    i_increase = 0;
    while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq)) <= W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) && (i_increase < 10)) {
        W_last *= 10.;
        i_increase++;
    }
    x[0] = fabs(W_last);
    xv = fabs(vsq_calc(W_last, s.Bsq, s.Qtsq, s.QdotBsq));
    x[1] = (xv > 1.) ? (1.0 - 1.e-15) : xv;
    retval = gnr_2d(x, s, gam);
    W = x[0];
    vsq = x[1];
    if ((retval != 0) || (W == kFAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > kW_TOO_BIG) return 3;
    if (vsq >= 1.) return 4;
    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    rho0 = s.D * gtmp;
    w = W * (1. - vsq);
    p = (gam - 1.) * (w - rho0) / gam;
    u = w - (rho0 + p);
    if (rho0 <= 0.) return 5;
    if (u <= 0.) return 6;
    prim[RHO] = rho0;
    prim[UU] = u;
    set_utcon(s, gamma, W, prim);
    return 0;
}

// GPU_utoprim_1dvsq2fix1.h W_of_vsq2
HD double W_of_vsq2(double vsq, double &p, double &rho, double &u, double D, double K, double gam)
{
    double gtmp = (1. - vsq);
    rho = D * sqrt(gtmp);
    p = K * pow(rho, gam);
    u = p / (gam - 1.);
    return ((rho + u + p) / gtmp);
}

// GPU_utoprim_1dvsq2fix1.h general_newton_raphson2 + func_1d_gnr2 + validate_x2
HD int gnr_1dvsq2(double *x, const InvPre &s, double K, double gam)
{
    double dx, x_old, errx = 1., W = 0., W_old = 0., p, rho, u, f = 1., df = 1.;
    int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1;
    while (keep) {
        double Wf, Wsq, dWdvsq, fact_tmp, resid, jac;
        Wf = W_of_vsq2(x[0], p, rho, u, s.D, K, gam);
        Wsq = Wf * Wf;
        dWdvsq = (gam * (2. - gam) * p + (gam - 1.) * rho) / (2. * (gam - 1.) * (1. - x[0]) * (1. - x[0]));
        fact_tmp = (s.Bsq + Wf);
        resid = s.Qtsq - x[0] * fact_tmp * fact_tmp + s.QdotBsq * (s.Bsq + 2. * Wf) / Wsq;
        jac = -fact_tmp * (fact_tmp + 2. * dWdvsq * (x[0] + s.QdotBsq / (Wf * Wsq)));
        dx = -resid / jac;
        f = 0.5 * resid * resid;
        df = -2. * f;
        x_old = x[0];
        x[0] += dx;
        x[0] = (x[0] >= 1.0) ? (0.5 * (x_old + 1.)) : x[0];
        x[0] = (x[0] < -1.e-10) ? (0.5 * x_old) : x[0];
        x[0] = fabs(x[0]);
        W_old = W;
        W = W_of_vsq2(x[0], p, rho, u, s.D, K, gam);
        errx = (W == 0.) ? fabs(W - W_old) : fabs((W - W_old) / W);
        errx += (x[0] == 0.) ? fabs(x[0] - x_old) : fabs((x[0] - x_old) / x[0]);
        if ((fabs(errx) <= kNEWT_TOL) && (doing_extra == 0) && (kEXTRA_NEWT_ITER > 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= kNEWT_TOL) && (doing_extra == 0)) || (i_extra > kEXTRA_NEWT_ITER) || (n_iter >= (kMAX_NEWT_ITER - 1))) keep = 0;
        n_iter++;
    }
    // untrapped divergences of the last residual (utoprim_1dvsq2fix1.c:488)
    if ((isfinite(f) == 0) || (isfinite(df) == 0)) return 2;
    if (fabs(errx) > mymin(kNEWT_TOL, kMIN_NEWT_TOL)) return 1;
    return 0;
}

// GPU_utoprim_1dvsq2fix1.h Utoprim_new_body2
HD int body_1dvsq2(const double *Ut, const double *gcov, const double *gcon, double *prim, double K, double gam)
{
    InvPre s;
    double x[1], utsq, gammasq, gamma, rho0, u, p, W, vsq, gtmp;
    int retval;
    inv_prepare(Ut, gcov, gcon, s);
    utsq = utsq_of(prim, gcov);
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > kUTSQ_TOO_BIG) return 2;
    gammasq = 1. + utsq;
    x[0] = 1. - 1. / gammasq;
    retval = gnr_1dvsq2(x, s, K, gam);
    if (retval != 0) return retval * 100 + 1;
    vsq = x[0];
    if ((vsq >= 1.) || (vsq < 0.)) return 4;
    W = W_of_vsq2(vsq, p, rho0, u, s.D, K, gam);
    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    if (rho0 <= 0.) return 5;
    if (u <= 0.) return 6;
    prim[RHO] = rho0;
    prim[UU] = u;
    set_utcon(s, gamma, W, prim);
    return 0;
}

// GPU_utoprim_1dfix1.h gnr2 + func_gnr2_rho: rho from W at fixed entropy
HD int gnr_rho(double *x, double D, double K, double Wg, double gam)
{
    double dx, x_old, errx = 1., f = 1., df = 1.;
    int n_iter = 0, keep = 1;
    while (keep) {
        double A, B, C, rho, W, B0, resid, jac;
        A = D * D;
        B0 = A * gam * K;
        B = B0 / (gam - 1.);
        rho = x[0];
        W = Wg;
        C = pow(rho, gam - 1.);
        resid = rho * W - A - B * C;
        jac = W - B0 * C / rho;
        dx = -resid / jac;
        f = 0.5 * resid * resid;
        df = -2. * f;
        x_old = x[0];
        x[0] += dx;
        x[0] = fabs(x[0]);
        errx = (x[0] == 0.) ? fabs(x[0] - x_old) : fabs((x[0] - x_old) / x[0]);
        if ((fabs(errx) <= kNEWT_TOL2) || (n_iter >= (kMAX_NEWT_ITER - 1))) keep = 0;
        n_iter++;
    }
    if ((isfinite(f) == 0) || (isfinite(df) == 0) || (isfinite(x[0]) == 0)) return 2;
    if (fabs(errx) > kMIN_NEWT_TOL2) return 1;
    return 0;
}

// GPU_utoprim_1dfix1.h general_newton_raphson3 + func_1d_orig1
HD int gnr_1dfix(double *x, const InvPre &s, double K, double rho_g0, double gam)
{
    double dx, x_old, errx = 1., f = 1., df = 1.;
    const double Bsq = s.Bsq, Qtsq = s.Qtsq, QdotBsq = s.QdotBsq, D = s.D;
    int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1, i_increase;
    while (keep) {
        double Dc, t1, t10, t2, t21, t23, t26, t29, t3, t30, t32, t33, t34, t38, t5, t51, t67, t8, x_rho[1], rho, rho_g, resid, jac;
        int ntries = 0;
        rho_g = x_rho[0] = rho_g0;
        while ((gnr_rho(x_rho, D, K, x[0], gam)) && (ntries++ < 10)) {
            rho_g *= 10.;
            x_rho[0] = rho_g;
        }
        rho = x_rho[0];
        Dc = D;
        t1 = Dc * Dc;
        t2 = QdotBsq * t1;
        t3 = t2 * Bsq;
        t5 = Bsq * Bsq;
        t8 = t1 * Bsq;
        t10 = t1 * x[0];
        t21 = x[0] * x[0];
        t23 = rho * rho;
        t26 = 1. / t1;
        resid = (t3 + (2.0 * t2 + ((Qtsq - t5) * t1 + (-2.0 * t8 - t10) * x[0]) * x[0]) * x[0] + (t5 + (2.0 * Bsq + x[0]) * x[0]) * t21 * t23) * t26 / t21;
        t29 = t1 * t1;
        t30 = QdotBsq * t29;
        t32 = gam * K;
        t33 = pow(rho, 1.0 * gam);
        t34 = t32 * t33;
        t38 = t23 * t33;
        t51 = gam * t1 * K * t33;
        t67 = t21 * x[0];
        jac = -2.0 * (t30 * Bsq * t34 + (t30 * t34 + ((-t38 * Bsq * t32 + Bsq * gam * t1 * K * t33) * t1 + (-t38 * gam * K + t51) * t1 * x[0]) * t21) * x[0] + ((-t3 + (-t2 + (-t8 - t10) * t21) * x[0]) * x[0] + (-t5 - Bsq * x[0]) * t67 * t23) * t23) * t26 / (t51 - x[0] * t23) / t67;
        dx = -resid / jac;
        f = 0.5 * resid * resid;
        df = -2. * f;
        x_old = x[0];
        x[0] += dx;
        i_increase = 0;
        while (((x[0] * x[0] * x[0] * (x[0] + 2. * Bsq) - QdotBsq * (2. * x[0] + Bsq)) <= x[0] * x[0] * (Qtsq - Bsq * Bsq)) && (i_increase < 10)) {
            x[0] -= (1. * i_increase) * dx / 10.;
            i_increase++;
        }
        x[0] = fabs(x[0]);
        errx = (x[0] == 0.) ? fabs(x[0] - x_old) : fabs((x[0] - x_old) / x[0]);
        if ((fabs(errx) <= kNEWT_TOL) && (doing_extra == 0) && (kEXTRA_NEWT_ITER > 0)) doing_extra = 1;
        if (doing_extra == 1) i_extra++;
        if (((fabs(errx) <= kNEWT_TOL) && (doing_extra == 0)) || (i_extra > kEXTRA_NEWT_ITER) || (n_iter >= (kMAX_NEWT_ITER - 1))) keep = 0;
        n_iter++;
    }
    if ((isfinite(f) == 0) || (isfinite(df) == 0) || (isfinite(x[0]) == 0)) return 2;
    if (fabs(errx) > mymin(kNEWT_TOL, kMIN_NEWT_TOL)) return 1;
    return 0;
}

// GPU_utoprim_1dfix1.h Utoprim_new_body3
HD int body_1dfix(const double *Ut, const double *gcov, const double *gcon, double *prim, double K, double gam)
{
    InvPre s;
    double x[1], utsq, gammasq, gamma, rho0, u, p, w, W_last, W, vsq, gtmp;
    int i_increase, retval;
    inv_prepare(Ut, gcov, gcon, s);
    utsq = utsq_of(prim, gcov);
    if ((utsq < 0.) && (fabs(utsq) < 1.0e-13)) utsq = fabs(utsq);
    if (utsq < 0. || utsq > kUTSQ_TOO_BIG) return 2;
    gammasq = 1. + utsq;
    gamma = sqrt(gammasq);
    rho0 = s.D / gamma;
    p = K * pow(rho0, gam);
    u = p / (gam - 1.);
    w = rho0 + u + p;
    W_last = w * gammasq;
    i_increase = 0;
    while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq)) <= W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) && (i_increase < 10)) {
        W_last *= 10.;
        i_increase++;
    }
    x[0] = W_last;
    retval = gnr_1dfix(x, s, K, rho0, gam);
    W = x[0];
    if ((retval != 0) || (W == kFAIL_VAL)) return retval * 100 + 1;
    if (W <= 0. || W > kW_TOO_BIG) return 3;
    vsq = vsq_calc(W, s.Bsq, s.Qtsq, s.QdotBsq);
    if (vsq >= 1.) return 4;
    gtmp = sqrt(1. - vsq);
    gamma = 1. / gtmp;
    rho0 = s.D * gtmp;
    p = K * pow(rho0, gam);
    u = p / (gam - 1.);
    if (rho0 <= 0.) return 5;
    if (u <= 0.) return 6;
    prim[RHO] = rho0;
    prim[UU] = u;
    set_utcon(s, gamma, W, prim);
    return 0;
}

// Utoprim_2d / _1dvsq2fix1 / _1dfix1 wrappers: normal-observer frame transform,
// body, write-back only on success (face fields always rewritten).
HD int utoprim(int which, const double *U, const double *gcov, const double *gcon, double gdet, double gam,
               double *prim)
{
    double Ut[NPR], pt[5], alpha, K = 0.;
    int ret;
    if (U[0] <= 0.) return -100;
    for (int i = B1; i <= B3; i++) prim[i] = U[i] / gdet;
    alpha = 1.0 / sqrt(-gcon[0]);
    Ut[RHO] = alpha * U[RHO] / gdet;
    Ut[UU] = alpha * (U[UU] - U[RHO]) / gdet;
    for (int i = U1; i <= U3; i++) Ut[i] = alpha * U[i] / gdet;
    for (int i = B1; i <= B3; i++) Ut[i] = alpha * U[i] / gdet;
    for (int i = 0; i < B1; i++) pt[i] = prim[i];
    if (which != 0) K = U[KTOT] / U[RHO];
    if (which == 0) ret = body_2d(Ut, gcov, gcon, pt, gam);
    else if (which == 1) ret = body_1dvsq2(Ut, gcov, gcon, pt, K, gam);
    else ret = body_1dfix(Ut, gcov, gcon, pt, K, gam);
    if (ret == 0)
        for (int i = 0; i < B1; i++) prim[i] = pt[i];
    return ret;
}

// inflow_check, GPU_bounds.h:1368 (type 0: inner edge, type 1: outer edge).
HD void inflow_check(double *pr, const Geom<double> &G, int type)
{
    double ucon[4], alpha, beta1, gamma, vsq;
    ucon_calc(pr, G, ucon);
    if (((ucon[1] > 0.) && (type == 0)) || ((ucon[1] < 0.) && (type == 1))) {
        gamma_calc(pr, G.gcov, &gamma);
        pr[U1] /= gamma;
        pr[U2] /= gamma;
        pr[U3] /= gamma;
        alpha = 1. / sqrt(-G.gcon[0]);
        beta1 = G.gcon[1] * alpha * alpha;
        pr[UU + 1] = beta1 / alpha;
        const double *gcov = G.gcov;
        vsq = gcov[4] * pr[U1] * pr[U1];
        vsq += 2. * gcov[5] * pr[U2] * pr[U1];
        vsq += 2. * gcov[6] * pr[U3] * pr[U1];
        vsq += gcov[7] * pr[U2] * pr[U2];
        vsq += 2 * gcov[8] * pr[U3] * pr[U2];
        vsq += gcov[9] * pr[U3] * pr[U3];
        vsq = mymax(1.e-13, vsq);
        if (vsq >= 1.) vsq = 1. - 1. / (kGAMMAMAX * kGAMMAMAX);
        gamma = 1. / sqrt(1. - vsq);
        pr[U1] *= gamma;
        pr[U2] *= gamma;
        pr[U3] *= gamma;
    }
}

}  // namespace hamr
