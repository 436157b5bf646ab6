// hamr_adjoint.h -- hand-written reverse-mode derivatives of the pointwise
// physics of one forward stage (FP64): the MHD state and fluxes of a
// reconstructed face state, the HLL combination, the cell-centre EMF, and the
// forward-mode Jacobian of the PPM reconstruction through the limiter's active
// branch.  Every function recomputes the primal quantities it needs from its
// inputs (nothing of the forward pass is stored).
//
// Conventions follow hamr_phys.h: prims RHO, UU, U1-U3 (relative 4-velocity),
// B1-B3, KTOT; packed symmetric metric; fluxes and conserved variables times
// gdet.  Branch points (the clamps of gamma_calc and of the magnetosonic speed,
// the max/min of the HLL speeds, the limiter's cases) take the derivative of the
// branch the primal takes.
#pragma once

#include "hamr_phys.h"

namespace hamr {

// the double versions stay visible beside the Dual overloads below
using ::copysign;
using ::fabs;
using ::fmin;

// ---------------------------------------------------------------------------
// Forward-mode number with N tangents (PPM Jacobian w.r.t. its 5 inputs).
template <int N> struct Dual {
    double v, d[N];
    HD Dual() : v(0.) { for (int k = 0; k < N; k++) d[k] = 0.; }
    HD Dual(double x) : v(x) { for (int k = 0; k < N; k++) d[k] = 0.; }
};
template <int N> HD Dual<N> operator+(const Dual<N> &a, const Dual<N> &b)
{ Dual<N> r; r.v = a.v + b.v; for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k]; return r; }
template <int N> HD Dual<N> operator-(const Dual<N> &a, const Dual<N> &b)
{ Dual<N> r; r.v = a.v - b.v; for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k]; return r; }
template <int N> HD Dual<N> operator-(const Dual<N> &a)
{ Dual<N> r; r.v = -a.v; for (int k = 0; k < N; k++) r.d[k] = -a.d[k]; return r; }
template <int N> HD Dual<N> operator*(const Dual<N> &a, const Dual<N> &b)
{ Dual<N> r; r.v = a.v * b.v; for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
// The value is the IEEE quotient; the tangents use one reciprocal of b (an
// FP64 division per tangent would cost as much as the rest of the parabola).
template <int N> HD Dual<N> operator/(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1. / b.v;
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N> HD bool operator<(const Dual<N> &a, const Dual<N> &b) { return a.v < b.v; }
template <int N> HD bool operator>(const Dual<N> &a, const Dual<N> &b) { return a.v > b.v; }
template <int N> HD bool operator<=(const Dual<N> &a, const Dual<N> &b) { return a.v <= b.v; }
template <int N> HD Dual<N> fabs(const Dual<N> &a) { return (a.v < 0.) ? -a : a; }
template <int N> HD Dual<N> fmin(const Dual<N> &a, const Dual<N> &b) { return (b.v < a.v) ? b : a; }
template <int N> HD Dual<N> copysign(const Dual<N> &a, const Dual<N> &s)
{
    const bool neg = __builtin_signbit(s.v), aneg = __builtin_signbit(a.v);
    return (neg == aneg) ? a : -a;
}
template <int N> HD Dual<N> dv(Dual<N> a, Dual<N> b) { return a / b; }

// d(l, r)/d(x1..x5) of para for one cell: Jl[5], Jr[5] (forward mode over
// all five inputs; the reference form of para_jacobian below).
HD void para_jacobian_dual(const double *x, double *l, double *r, double *Jl, double *Jr)
{
    Dual<5> y[5];
    for (int k = 0; k < 5; k++) {
        y[k] = Dual<5>(x[k]);
        y[k].d[k] = 1.;
    }
    Dual<5> lo, ro;
    para(y[0], y[1], y[2], y[3], y[4], lo, ro);
    *l = lo.v;
    *r = ro.v;
    for (int k = 0; k < 5; k++) {
        Jl[k] = lo.d[k];
        Jr[k] = ro.d[k];
    }
}

// Tangent of mc_slope(ym, y0, yp) along the branch it takes (the choices of
// para_jacobian_dual's fabs, fmin and copysign): s * dX, X the selected one of
// Dqm = 2 (y0 - ym), Dqp = 2 (yp - y0), Dqc = 0.5 (yp - ym); t over (ym, y0, yp).
HD void mc_slope_tangent(double ym, double y0, double yp, double t[3])
{
    const double Dqm = 2. * (y0 - ym), Dqp = 2. * (yp - y0), Dqc = 0.5 * (yp - ym);
    t[0] = t[1] = t[2] = 0.;
    if (!(Dqm * Dqp > 0.)) return;
    const double am = fabs(Dqm), ap = fabs(Dqp), ac = fabs(Dqc);
    const bool pick_p = ap < am;                // fmin(|Dqm|, |Dqp|)
    const double m = pick_p ? ap : am;
    const double sgn = __builtin_signbit(Dqc) ? -1. : 1.;  // copysign(., Dqc) of a positive value
    if (m < ac) {
        if (pick_p) {
            const double f = sgn * (Dqp < 0. ? -2. : 2.);
            t[1] = -f;
            t[2] = f;
        } else {
            const double f = sgn * (Dqm < 0. ? -2. : 2.);
            t[0] = -f;
            t[1] = f;
        }
    } else {
        const double f = sgn * (Dqc < 0. ? -0.5 : 0.5);
        t[0] = -f;
        t[2] = f;
    }
}

// d(l, r)/d(x1..x5) of para for one cell: Jl[5], Jr[5], in closed form along
// the limiter's active branches (the slopes' tangents have three entries, the
// interface values' four): the values are para's; the tangents equal
// para_jacobian_dual's up to the rounding of the division by 6.
HD void para_jacobian(const double *x, double *l, double *r, double *Jl, double *Jr)
{
    double dq[3], t[3][5];
    for (int s = 0; s < 3; s++) {
        dq[s] = mc_slope(x[s], x[s + 1], x[s + 2]);
        double ts[3];
        mc_slope_tangent(x[s], x[s + 1], x[s + 2], ts);
        for (int k = 0; k < 5; k++) t[s][k] = (k >= s && k < s + 3) ? ts[k - s] : 0.;
    }
    const double sixth = 1. / 6.;
    // interface values below and above cell x3 (ppm_face) and their tangents
    double lo = ppm_face(x[1], x[2], dq[0], dq[1]), hi = ppm_face(x[2], x[3], dq[1], dq[2]);
    double tl[5], tr[5];
    for (int k = 0; k < 5; k++) {
        tl[k] = 0.5 * ((k == 2 ? 1. : 0.) + (k == 1 ? 1. : 0.)) - (t[1][k] - t[0][k]) * sixth;
        tr[k] = 0.5 * ((k == 3 ? 1. : 0.) + (k == 2 ? 1. : 0.)) - (t[2][k] - t[1][k]) * sixth;
    }
    // ppm_limit on values and tangents
    const double x3 = x[2];
    const double qa = (hi - x3) * (x3 - lo), qd = hi - lo, qe = 6.0 * (x3 - 0.5 * (lo + hi));
    if (qa <= 0.) {
        lo = x3;
        hi = x3;
        for (int k = 0; k < 5; k++) tl[k] = tr[k] = (k == 2 ? 1. : 0.);
    }
    if (qd * (qd - qe) < 0.) {
        lo = 3.0 * x3 - 2.0 * hi;
        for (int k = 0; k < 5; k++) tl[k] = 3.0 * (k == 2 ? 1. : 0.) - 2.0 * tr[k];
    } else if (qd * (qd + qe) < 0.) {
        hi = 3.0 * x3 - 2.0 * lo;
        for (int k = 0; k < 5; k++) tr[k] = 3.0 * (k == 2 ? 1. : 0.) - 2.0 * tl[k];
    }
    *l = lo;
    *r = hi;
    for (int k = 0; k < 5; k++) {
        Jl[k] = tl[k];
        Jr[k] = tr[k];
    }
}

// ---------------------------------------------------------------------------
// Adjoints of the MHD state of one set of primitives (ucon, ucov, bcon, bcov,
// rho, u), accumulated by the flux and speed adjoints and pulled back to the
// primitives by state_pullback.
struct StateAdj {
    double ucon[4], ucov[4], bcon[4], bcov[4], rho, u, bsq;
};

HD void state_adj_zero(StateAdj &a)
{
    for (int k = 0; k < 4; k++) a.ucon[k] = a.ucov[k] = a.bcon[k] = a.bcov[k] = 0.;
    a.rho = a.u = a.bsq = 0.;
}

// primtoflux (GPU_phys_mhd.h:432) in direction dir, reverse: lF are the
// adjoints of the nine (gdet-weighted) outputs.
HD void primtoflux_vjp(const double *pr, const State<double> &q, int dir, const Geom<double> &G, double gam,
                       const double *lF, StateAdj &a)
{
    const double rho = pr[RHO], u = pr[UU], P = (gam - 1.) * u;
    const double bsq = dot4(q.bcon, q.bcov), eta = rho + P + u + bsq;
    const double ud = q.ucon[dir], bd = q.bcon[dir];
    double lf[NPR];
    for (int k = 0; k < NPR; k++) lf[k] = G.g * lF[k];
    // f[KTOT] = f[RHO] * entr, entr = P rho^-gam; f[UU] = mhd[0] + f[RHO]
    const double rmg = pow(rho, -gam), entr = P * rmg, fr = rho * ud;
    const double lfr = lf[RHO] + lf[UU] + lf[KTOT] * entr;
    const double lentr = lf[KTOT] * fr;
    a.rho += lfr * ud - lentr * gam * entr / rho;
    a.u += lentr * (gam - 1.) * rmg;
    a.ucon[dir] += lfr * rho;
    // mhd[j] = eta ud u_j + ptot delta(dir, j) - bd b_j
    double leta = 0., lptot = 0.;
    for (int j = 0; j < 4; j++) {
        const double lm = (j == 0) ? lf[UU] : lf[U1 + j - 1];
        leta += lm * ud * q.ucov[j];
        a.ucon[dir] += lm * eta * q.ucov[j];
        a.ucov[j] += lm * eta * ud;
        if (j == dir) lptot += lm;
        a.bcon[dir] -= lm * q.bcov[j];
        a.bcov[j] -= lm * bd;
    }
    // f[Bj] = b^j ud - bd u^j
    for (int j = 1; j < 4; j++) {
        const double l = lf[B1 + j - 1];
        a.bcon[j] += l * ud;
        a.ucon[dir] += l * q.bcon[j];
        a.bcon[dir] -= l * q.ucon[j];
        a.ucon[j] -= l * bd;
    }
    // eta = rho + gam u + bsq, ptot = (gam - 1) u + bsq / 2
    a.rho += leta;
    a.u += leta * gam + lptot * (gam - 1.);
    a.bsq += leta + 0.5 * lptot;
}

// Magnetosonic speeds of primtoflux (GPU_phys_mhd.h:550-617), reverse.
HD void speeds_vjp(const double *pr, const State<double> &q, int dir, const Geom<double> &G, double gam, double lvmax,
                   double lvmin, StateAdj &a)
{
    const double rho = pr[RHO], u = pr[UU], P = (gam - 1.) * u, w = rho + P + u;
    const double bsq = dot4(q.bcon, q.bcov), eta = w + bsq;
    double Acon_0, Acon_js;
    if (dir == 1) { Acon_0 = G.gcon[1]; Acon_js = G.gcon[4]; }
    else if (dir == 2) { Acon_0 = G.gcon[2]; Acon_js = G.gcon[7]; }
    else { Acon_0 = G.gcon[3]; Acon_js = G.gcon[9]; }
    const double cs2 = gam * (gam - 1.) * u / w, va2 = bsq / eta;
    double cms2 = cs2 + va2 - cs2 * va2;
    bool clamped = false;
    if (cms2 < 0.) { cms2 = kSMALL; clamped = true; }
    if (cms2 > 1.) { cms2 = 1.; clamped = true; }
    const double Asq = Acon_js, Bsq = G.gcon[0], Au = q.ucon[dir], Bu = q.ucon[0], AB = Acon_0;
    const double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
    const double A = Bu2 - (Bsq + Bu2) * cms2;
    const double B = 2. * (AuBu - (AB + AuBu) * cms2);
    const double C = Au2 - (Asq + Au2) * cms2;
    double discr = B * B - 4. * A * C;
    const bool dclamp = !(discr > 0.);
    if (dclamp) discr = 0.;
    const double sd = sqrt(discr);
    const double vp = -(-B + sd) / (2. * A), vm = -(-B - sd) / (2. * A);
    // vmax = max(vp, vm), vmin = min(vp, vm)
    double lvp, lvm;
    if (vp > vm) { lvp = lvmax; lvm = lvmin; }
    else { lvp = lvmin; lvm = lvmax; }
    // vp = (B - sd) / (2A), vm = (B + sd) / (2A)
    double lB = (lvp + lvm) / (2. * A);
    const double lsd = (lvm - lvp) / (2. * A);
    double lA = -(lvp * vp + lvm * vm) / A;
    double lC = 0.;
    if (!dclamp) {
        const double ld = lsd / (2. * sd);
        lB += ld * 2. * B;
        lA += -4. * C * ld;
        lC += -4. * A * ld;
    }
    const double om = 1. - cms2;
    a.ucon[0] += lA * 2. * Bu * om + lB * 2. * Au * om;
    a.ucon[dir] += lB * 2. * Bu * om + lC * 2. * Au * om;
    if (clamped) return;
    const double lcms2 = -lA * (Bsq + Bu2) - lB * 2. * (AB + AuBu) - lC * (Asq + Au2);
    const double lcs2 = lcms2 * (1. - va2), lva2 = lcms2 * (1. - cs2);
    // cs2 = gam (gam - 1) u / w, w = rho + gam u
    a.u += lcs2 * (gam * (gam - 1.) / w - cs2 * gam / w);
    a.rho += -lcs2 * cs2 / w;
    // va2 = bsq / eta, eta = rho + gam u + bsq
    const double leta = -lva2 * va2 / eta;
    a.bsq += lva2 / eta + leta;
    a.rho += leta;
    a.u += leta * gam;
}

// get_state (GPU_phys_mhd.h:692-760), reverse: adds the primitives' adjoints to lpr.
HD void state_pullback(const double *pr, const State<double> &q, const Geom<double> &G, StateAdj &a, double *lpr)
{
    // bsq = bcon . bcov
    for (int m = 0; m < 4; m++) {
        a.bcon[m] += a.bsq * q.bcov[m];
        a.bcov[m] += a.bsq * q.bcon[m];
    }
    // bcov = g bcon
    {
        double t[4];
        lower10(a.bcov, G.gcov, t);
        for (int m = 0; m < 4; m++) a.bcon[m] += t[m];
    }
    // bcon[j] = (B_j + b0 u^j) / u0, b0 = B . ucov(1..3)
    const double u0 = q.ucon[0], b0 = q.bcon[0];
    double lb0 = a.bcon[0];
    for (int j = 1; j < 4; j++) {
        const double l = a.bcon[j];
        lpr[B1 + j - 1] += l / u0;
        lb0 += l * q.ucon[j] / u0;
        a.ucon[j] += l * b0 / u0;
        a.ucon[0] -= l * q.bcon[j] / u0;
    }
    for (int i = 1; i < 4; i++) {
        lpr[B1 + i - 1] += lb0 * q.ucov[i];
        a.ucov[i] += lb0 * pr[B1 + i - 1];
    }
    // ucov = g ucon
    {
        double t[4];
        lower10(a.ucov, G.gcov, t);
        for (int m = 0; m < 4; m++) a.ucon[m] += t[m];
    }
    // ucon[0] = gamma / alpha, ucon[j] = U_j - gamma beta_j / alpha
    const double alpha = 1. / sqrt(-G.gcon[0]);
    double lgam = a.ucon[0] / alpha;
    for (int j = 1; j < 4; j++) {
        const double beta = G.gcon[j] * alpha * alpha;
        lgam -= a.ucon[j] * beta / alpha;
        lpr[U1 + j - 1] += a.ucon[j];
    }
    // gamma = sqrt(1 + qsq), qsq = U . g_spatial . U (no derivative where clamped)
    const double *g = G.gcov;
    const double U[3] = {pr[U1], pr[U2], pr[U3]};
    const double qsq = g[4] * U[0] * U[0] + g[7] * U[1] * U[1] + g[9] * U[2] * U[2]
        + 2. * (g[5] * U[0] * U[1] + g[6] * U[0] * U[2] + g[8] * U[1] * U[2]);
    if (qsq >= 0.) {
        const double gamma = sqrt(1. + qsq);
        const double lq = lgam / (2. * gamma);
        lpr[U1] += lq * 2. * (g[4] * U[0] + g[5] * U[1] + g[6] * U[2]);
        lpr[U2] += lq * 2. * (g[5] * U[0] + g[7] * U[1] + g[8] * U[2]);
        lpr[U3] += lq * 2. * (g[6] * U[0] + g[8] * U[1] + g[9] * U[2]);
    }
    lpr[RHO] += a.rho;
    lpr[UU] += a.u;
}

// One face state: fluxes F (dir), conserved U (dir 0) and speeds, reverse.
HD void face_state_vjp(const double *pr, const Geom<double> &G, int dir, double gam, const double *lF, const double *lU,
                       double lvmax, double lvmin, double *lpr)
{
    State<double> q;
    get_state(pr, G, q);
    StateAdj a;
    state_adj_zero(a);
    primtoflux_vjp(pr, q, dir, G, gam, lF, a);
    primtoflux_vjp(pr, q, 0, G, gam, lU, a);
    speeds_vjp(pr, q, dir, G, gam, lvmax, lvmin, a);
    state_pullback(pr, q, G, a, lpr);
}

// Conserved variables U(P) (primtoflux, dir 0), reverse.
HD void conserved_vjp(const double *pr, const Geom<double> &G, double gam, const double *lU, double *lpr)
{
    State<double> q;
    get_state(pr, G, q);
    StateAdj a;
    state_adj_zero(a);
    primtoflux_vjp(pr, q, 0, G, gam, lU, a);
    state_pullback(pr, q, G, a, lpr);
}

// Cell-centre EMF E_i = -gdet (u^j b^k - u^k b^j) (GPU_const_trans.h:63-119), reverse.
HD void ecent_vjp(const double *pr, const Geom<double> &G, const double *lE, double *lpr)
{
    State<double> q;
    get_state(pr, G, q);
    StateAdj a;
    state_adj_zero(a);
    const double *u = q.ucon, *b = q.bcon;
    // E1 = -g (u2 b3 - u3 b2), E2 = -g (u3 b1 - u1 b3), E3 = -g (u1 b2 - u2 b1)
    const double l1 = -G.g * lE[0], l2 = -G.g * lE[1], l3 = -G.g * lE[2];
    a.ucon[2] += l1 * b[3]; a.bcon[3] += l1 * u[2]; a.ucon[3] -= l1 * b[2]; a.bcon[2] -= l1 * u[3];
    a.ucon[3] += l2 * b[1]; a.bcon[1] += l2 * u[3]; a.ucon[1] -= l2 * b[3]; a.bcon[3] -= l2 * u[1];
    a.ucon[1] += l3 * b[2]; a.bcon[2] += l3 * u[1]; a.ucon[2] -= l3 * b[1]; a.bcon[1] -= l3 * u[2];
    state_pullback(pr, q, G, a, lpr);
}

}  // namespace hamr
