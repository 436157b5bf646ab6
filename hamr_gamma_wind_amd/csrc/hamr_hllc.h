// hamr_hllc.h -- HLLC Riemann solver in the face's orthonormal frame
// (fluxcalc2D_FT, GPU_hllc.h:3-247, with calc_HLLC_hydro :250 and calc_HLLC
// :317; ideal gas, entropy tracer).
//
// At a face of direction DIR the reconstructed states are moved to the
// orthonormal tetrad of the face (host-built, face_tetrad), where the
// flat-space HLL state and flux select between the hydrodynamic HLLC (normal
// field negligible) and the relativistic MHD HLLC of the contact; the
// frame-space state and flux are mapped back to the coordinate flux through the
// face.  The expressions keep the reference's operand order: the FP64 step
// reproduces it rounding for rounding.
//
// TAU (the FP32 form, on by default for R = float): the frame energy entries of
// U, F and the HLL state carry T^(a)_(0) + rho u^(a), formed with u^(0) - 1 =
// q^2 / (u^(0) + 1); the star state's energy takes T^(0)_(dir) - rho u^(dir);
// the contact speed is the same root written 2C / (-B + sqrt(D)); and the
// coordinate energy flux T^dir_0 + rho u^dir is the spatial frame rows plus
// alpha times the frame energy plus (1 - alpha, formed in FP64) times the mass
// flux.  No step subtracts the rest mass from a quantity that holds it, which
// in FP32 would leave the energy flux with few correct digits.
#pragma once

#include "hamr_phys.h"

namespace hamr {

// primtoflux_FT (GPU_hllc.h:903-997): flat-space row `dir` of the stress tensor
// (T^dir_nu, no rest mass added), number and entropy-tracer flux and the dual
// Maxwell row; dir = 0 gives the conserved densities.
template <class R, bool TAU = (sizeof(R) == 4)>
HD void primtoflux_ft(const R *pr, const R *ucon, const R *bcon, int dir, R gam, R *flux)
{
    const R P = (gam - R(1.)) * pr[UU];
    const R w = P + pr[RHO] + pr[UU];
    const R bsq = -bcon[0] * bcon[0] + bcon[1] * bcon[1] + bcon[2] * bcon[2] + bcon[3] * bcon[3];
    const R eta = w + bsq;
    const R ptot = P + R(0.5) * bsq;
    flux[RHO] = pr[RHO] * ucon[dir];
    if constexpr (TAU) {
        // FP32: the energy row with the rest-mass term added (T^dir_0 + rho u^dir),
        // formed with u^0 - 1 = q^2 / (u^0 + 1) so that nothing cancels
        const R um1 = dv(ucon[1] * ucon[1] + ucon[2] * ucon[2] + ucon[3] * ucon[3], ucon[0] + R(1.));
        flux[UU] = -(flux[RHO] * um1) - (pr[UU] + P + bsq) * ucon[dir] * ucon[0] + ptot * (dir == 0 ? R(1.) : R(0.))
            + bcon[dir] * bcon[0];
    } else {
        flux[UU] = -eta * ucon[dir] * ucon[0] + ptot * (dir == 0 ? R(1.) : R(0.)) + bcon[dir] * bcon[0];
    }
    for (int j = 1; j < 4; j++)
        flux[UU + j] = eta * ucon[dir] * ucon[j] + ptot * (dir == j ? R(1.) : R(0.)) - bcon[dir] * bcon[j];
    for (int k = B1; k <= B3; k++) flux[k] = bcon[k - 4] * ucon[dir] - bcon[dir] * ucon[k - 4];
    flux[KTOT] = flux[RHO] * P * rpow(pr[RHO], -gam);
}

// vchar_FT (GPU_hllc.h:999-1111, FULL_DISP off): fast magnetosonic speeds
// along dir seen in the flat frame.
template <class R> HD void vchar_ft(const R *pr, const R *ucon, const R *bcon, int dir, R gam, R *vmax, R *vmin)
{
    const R bsq = -bcon[0] * bcon[0] + bcon[1] * bcon[1] + bcon[2] * bcon[2] + bcon[3] * bcon[3];
    const R EF = pr[RHO] + gam * pr[UU];
    const R cs2 = dv(gam * (gam - R(1.)) * pr[UU], EF);
    const R EE = bsq + EF;
    const R va2 = dv(bsq, EE);
    R cms2 = cs2 + va2 - cs2 * va2;
    if (cms2 < R(0.)) cms2 = R(kSMALL);
    if (cms2 > R(1.)) cms2 = R(1.);
    const R Asq = R(1.), Bsq = R(-1.);
    const R Au = ucon[dir], Bu = ucon[0];
    const R Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
    const R A = Bu2 - (Bsq + Bu2) * cms2;
    const R B = R(2.) * (AuBu - (AuBu)*cms2);
    const R C = Au2 - (Asq + Au2) * cms2;
    R discr = B * B - R(4.) * A * C;
    if ((discr < R(0.0)) && (discr > R(-1.e-10))) discr = R(0.0);
    else if (discr < R(-1.e-10)) discr = R(0.);
    discr = sqrt(discr);
    const R vp = dv(-(-B + discr), R(2.) * A);
    const R vm = dv(-(-B - discr), R(2.) * A);
    if (vp > vm) { *vmax = vp; *vmin = vm; }
    else { *vmax = vm; *vmin = vp; }
}

// The face's HLL state and flux (index 0: state, 1: flux) in the flat frame and
// the left / right states, fluxes and 4-velocities the two HLLC variants use.
// In FP32 the energy entries (UU) of U, F and HLL hold T^(a)_(0) + rho u^(a)
// (primtoflux_ft), and smd = T^(0)_(dir) - rho u^(dir), formed the same way,
// enters the star state's energy.
template <class R> struct FtFace {
    R U_l[NPR], U_r[NPR], F_l[NPR], F_r[NPR];
    R HLL[2][NPR];
    R lu0, lud, ru0, rud;  // u^(0), u^(dir) of the left and right states
    R smd_l, smd_r;
};

// T^(0)_(dir) - rho u^(dir) of a state in the flat frame (FP32 path)
template <class R> HD R ft_smd(const R *pr, const R *ucon, const R *bcon, int dir, R gam)
{
    const R P = (gam - R(1.)) * pr[UU];
    const R bsq = -bcon[0] * bcon[0] + bcon[1] * bcon[1] + bcon[2] * bcon[2] + bcon[3] * bcon[3];
    const R um1 = dv(ucon[1] * ucon[1] + ucon[2] * ucon[2] + ucon[3] * ucon[3], ucon[0] + R(1.));
    return pr[RHO] * ucon[dir] * um1 + (pr[UU] + P + bsq) * ucon[0] * ucon[dir] - bcon[0] * bcon[dir];
}

// Cyclic component order starting at the face direction: g[0..2] = dir, next, last.
template <int DIR> struct Cyc {
    static constexpr int g1 = DIR, g2 = DIR % 3 + 1, g3 = (DIR % 3 + 1) % 3 + 1;
    static constexpr int U_1 = UU + g1, U_2 = UU + g2, U_3 = UU + g3;
    static constexpr int B_1 = B1 - 1 + g1, B_2 = B1 - 1 + g2, B_3 = B1 - 1 + g3;
};

// calc_HLLC_hydro (GPU_hllc.h:250-315): the contact speed from the HLL state
// and flux of a hydrodynamic (field-free) Riemann fan, then the star state of
// the upwind side; HLL where the contact is not admissible.
template <int DIR, class R, bool TAU>
HD void hllc_hydro(const FtFace<R> &f, R int_velocity, R cmin_roe, R cmax_roe, R (&FT)[2][NPR])
{
    using Y = Cyc<DIR>;
    // the contact from the total energy (FP32: the stored entry minus rho u)
    R H[2][NPR];
    for (int a = 0; a < 2; a++)
        for (int k = 0; k < NPR; k++) H[a][k] = f.HLL[a][k];
    if constexpr (TAU) {
        H[0][UU] = f.HLL[0][UU] - f.HLL[0][RHO];
        H[1][UU] = f.HLL[1][UU] - f.HLL[1][RHO];
    }
    const R A = -H[1][UU];
    const R B = -H[1][Y::U_1] + H[0][UU];
    const R C = H[0][Y::U_1];
    const R D = B * B - R(4.) * A * C;
    // FP32: the same root as 2C / (-B + sqrt(D)) (-B ~ the energy density, so
    // -B - sqrt(D) would cancel)
    const R vcon = TAU ? dv(R(2.) * C, -B + R(sqrt(D))) : dv(-B - R(sqrt(D)), R(2.) * A);
    const R ptot = -(-H[1][UU]) * vcon + H[1][Y::U_1];
    const bool fail = !(fabs(ptot) > R(0.)) || (vcon < cmin_roe) || (vcon > cmax_roe) || !(fabs(vcon) > R(0.));
    if (cmax_roe > int_velocity && vcon <= int_velocity && !fail) {
        const R s = dv(cmax_roe - dv(f.rud, f.ru0), cmax_roe - vcon + R(kSMALL));
        FT[0][RHO] = s * f.U_r[RHO];
        if constexpr (TAU) {
            FT[0][UU] = dv(cmax_roe * f.U_r[UU] + f.smd_r - ptot * vcon, cmax_roe - vcon + R(kSMALL));
            FT[0][Y::U_1] = (-(FT[0][UU] - FT[0][RHO]) + ptot) * vcon;
        } else {
            FT[0][UU] = dv(cmax_roe * f.U_r[UU] + f.U_r[Y::U_1] - ptot * vcon, cmax_roe - vcon + R(kSMALL));
            FT[0][Y::U_1] = (-FT[0][UU] + ptot) * vcon;
        }
        FT[0][Y::U_2] = s * f.U_r[Y::U_2];
        FT[0][Y::U_3] = s * f.U_r[Y::U_3];
        FT[0][Y::B_1] = H[0][Y::B_1];
        FT[0][Y::B_2] = s * f.U_r[Y::B_2];
        FT[0][Y::B_3] = s * f.U_r[Y::B_3];
        FT[0][KTOT] = s * f.U_r[KTOT];
        for (int k = 0; k < NPR; k++) FT[1][k] = (f.F_r[k] + cmax_roe * (FT[0][k] - f.U_r[k]));
    } else if (cmin_roe < int_velocity && vcon >= int_velocity && !fail) {
        const R s = dv(cmin_roe - dv(f.lud, f.lu0), cmin_roe - vcon + R(kSMALL));
        FT[0][RHO] = s * f.U_l[RHO];
        if constexpr (TAU) {
            FT[0][UU] = dv(cmin_roe * f.U_l[UU] + f.smd_l - ptot * vcon, cmin_roe - vcon + R(kSMALL));
            FT[0][Y::U_1] = (-(FT[0][UU] - FT[0][RHO]) + ptot) * vcon;
        } else {
            FT[0][UU] = dv(cmin_roe * f.U_l[UU] + f.U_l[Y::U_1] - ptot * vcon, cmin_roe - vcon + R(kSMALL));
            FT[0][Y::U_1] = (-FT[0][UU] + ptot) * vcon;
        }
        FT[0][Y::U_2] = s * f.U_l[Y::U_2];
        FT[0][Y::U_3] = s * f.U_l[Y::U_3];
        FT[0][Y::B_1] = H[0][Y::B_1];
        FT[0][Y::B_2] = s * f.U_l[Y::B_2];
        FT[0][Y::B_3] = s * f.U_l[Y::B_3];
        FT[0][KTOT] = s * f.U_l[KTOT];
        for (int k = 0; k < NPR; k++) FT[1][k] = (f.F_l[k] + cmin_roe * (FT[0][k] - f.U_l[k]));
    } else {
        for (int k = 0; k < NPR; k++) {
            FT[0][k] = f.HLL[0][k];
            FT[1][k] = f.HLL[1][k];
        }
    }
}

// calc_HLLC (GPU_hllc.h:317-400): the relativistic MHD HLLC of the contact
// (normal field from the HLL state): contact velocity from the quadratic of
// the HLL state and flux, its transverse components from the field, total
// pressure, and the star state of the upwind side; HLL where the contact is
// not admissible.
template <int DIR, class R, bool TAU>
HD void hllc_mhd(const FtFace<R> &f, R int_velocity, R cmin_roe, R cmax_roe, R (&FT)[2][NPR])
{
    using Y = Cyc<DIR>;
    R H[2][NPR];
    for (int a = 0; a < 2; a++)
        for (int k = 0; k < NPR; k++) H[a][k] = f.HLL[a][k];
    if constexpr (TAU) {
        H[0][UU] = f.HLL[0][UU] - f.HLL[0][RHO];
        H[1][UU] = f.HLL[1][UU] - f.HLL[1][RHO];
    }
    R vcon[4];
    const R A = -H[1][UU] - (H[0][Y::B_2] * H[1][Y::B_2] + H[0][Y::B_3] * H[1][Y::B_3]);
    const R B = -H[1][Y::U_1] + H[0][UU] + (H[0][Y::B_2] * H[0][Y::B_2] + H[0][Y::B_3] * H[0][Y::B_3])
        + (H[1][Y::B_2] * H[1][Y::B_2] + H[1][Y::B_3] * H[1][Y::B_3]);
    const R C = H[0][Y::U_1] - (H[0][Y::B_2] * H[1][Y::B_2] + H[0][Y::B_3] * H[1][Y::B_3]);
    const R D = B * B - R(4.) * A * C;
    vcon[0] = R(0.);
    vcon[Y::g1] = TAU ? dv(R(2.) * C, -B + R(sqrt(D))) : dv(-B - R(sqrt(D)), R(2.) * A);
    vcon[Y::g2] = dv(H[0][Y::B_2] * vcon[Y::g1] - H[1][Y::B_2], H[0][Y::B_1]);
    vcon[Y::g3] = dv(H[0][Y::B_3] * vcon[Y::g1] - H[1][Y::B_3], H[0][Y::B_1]);
    const R gammasq = dv(R(1.), R(1.) - (vcon[1] * vcon[1] + vcon[2] * vcon[2] + vcon[3] * vcon[3]));
    const R v_dot_B = (vcon[1] * H[0][B1] + vcon[2] * H[0][B2] + vcon[3] * H[0][B3]);
    const R ptot = -(-H[1][UU] - H[0][Y::B_1] * (v_dot_B)) * vcon[DIR] + H[1][Y::U_1]
        + dv(H[0][Y::B_1] * H[0][Y::B_1], gammasq);
    const R vn = vcon[DIR];
    const bool fail = !(fabs(ptot) > R(0.)) || (vn < cmin_roe) || (vn > cmax_roe) || !(fabs(vn) > R(0.));
    if (cmax_roe > int_velocity && vn <= int_velocity && !fail) {
        const R den = cmax_roe - vn + R(kSMALL);
        FT[0][RHO] = dv(cmax_roe - dv(f.rud, f.ru0), den) * f.U_r[RHO];
        if constexpr (TAU) {
            FT[0][UU] = dv(cmax_roe * f.U_r[UU] + f.smd_r - ptot * vn + v_dot_B * H[0][Y::B_1], den);
            FT[0][Y::U_1] = (-(FT[0][UU] - FT[0][RHO]) + ptot) * vn - v_dot_B * H[0][Y::B_1];
        } else {
            FT[0][UU] = dv(cmax_roe * f.U_r[UU] + f.U_r[Y::U_1] - ptot * vn + v_dot_B * H[0][Y::B_1], den);
            FT[0][Y::U_1] = (-FT[0][UU] + ptot) * vn - v_dot_B * H[0][Y::B_1];
        }
        FT[0][Y::U_2] = dv(-H[0][Y::B_1] * (dv(H[0][Y::B_2], gammasq) + v_dot_B * vcon[Y::g2])
                               + cmax_roe * f.U_r[Y::U_2] - f.F_r[Y::U_2], den);
        FT[0][Y::U_3] = dv(-H[0][Y::B_1] * (dv(H[0][Y::B_3], gammasq) + v_dot_B * vcon[Y::g3])
                               + cmax_roe * f.U_r[Y::U_3] - f.F_r[Y::U_3], den);
        FT[0][Y::B_1] = H[0][Y::B_1];
        FT[0][Y::B_2] = H[0][Y::B_2];
        FT[0][Y::B_3] = H[0][Y::B_3];
        FT[0][KTOT] = dv(cmax_roe - dv(f.rud, f.ru0), den) * f.U_r[KTOT];
        for (int k = 0; k < NPR; k++) FT[1][k] = (f.F_r[k] + cmax_roe * (FT[0][k] - f.U_r[k]));
    } else if (cmin_roe < int_velocity && vn >= int_velocity && !fail) {
        const R den = cmin_roe - vn + R(kSMALL);
        FT[0][RHO] = dv(cmin_roe - dv(f.lud, f.lu0), den) * f.U_l[RHO];
        if constexpr (TAU) {
            FT[0][UU] = dv(cmin_roe * f.U_l[UU] + f.smd_l - ptot * vn + v_dot_B * H[0][Y::B_1], den);
            FT[0][Y::U_1] = (-(FT[0][UU] - FT[0][RHO]) + ptot) * vn - v_dot_B * H[0][Y::B_1];
        } else {
            FT[0][UU] = dv(cmin_roe * f.U_l[UU] + f.U_l[Y::U_1] - ptot * vn + v_dot_B * H[0][Y::B_1], den);
            FT[0][Y::U_1] = (-FT[0][UU] + ptot) * vn - v_dot_B * H[0][Y::B_1];
        }
        FT[0][Y::U_2] = dv(-H[0][Y::B_1] * (dv(H[0][Y::B_2], gammasq) + v_dot_B * vcon[Y::g2])
                               + cmin_roe * f.U_l[Y::U_2] - f.F_l[Y::U_2], den);
        FT[0][Y::U_3] = dv(-H[0][Y::B_1] * (dv(H[0][Y::B_3], gammasq) + v_dot_B * vcon[Y::g3])
                               + cmin_roe * f.U_l[Y::U_3] - f.F_l[Y::U_3], den);
        FT[0][Y::B_1] = H[0][Y::B_1];
        FT[0][Y::B_2] = H[0][Y::B_2];
        FT[0][Y::B_3] = H[0][Y::B_3];
        FT[0][KTOT] = dv(cmin_roe - dv(f.lud, f.lu0), den) * f.U_l[KTOT];
        for (int k = 0; k < NPR; k++) FT[1][k] = (f.F_l[k] + cmin_roe * (FT[0][k] - f.U_l[k]));
    } else {
        for (int k = 0; k < NPR; k++) {
            FT[0][k] = f.HLL[0][k];
            FT[1][k] = f.HLL[1][k];
        }
    }
}

// The frame part of fluxcalc2D_FT (GPU_hllc.h:113-230) for one face: from the
// coordinate 4-velocities and fields of the two states, the tetrad mud (e_(a)^mu
// at [mu*4 + a]) and its inverse mudi ([a*4 + mu]), and the face's coordinate
// velocity int_velocity in the frame, the 9 coordinate fluxes (before the gdet
// factor; the energy flux with the rest-mass flux added, as the step's
// conserved variables are).
template <int DIR, class R, bool TAU = (sizeof(R) == 4)>
HD void hllc_frame_flux(const R *pl, const R *pr, const R *ucl, const R *bcl, const R *ucr, const R *bcr, const R *mud,
                        const R *mudi, R int_velocity, R gam, R oma, R *F1)
{
    R l_ucon[4], l_bcon[4], r_ucon[4], r_bcon[4];
    for (int i1 = 0; i1 < 4; i1++) {
        l_ucon[i1] = R(0.0);
        l_bcon[i1] = R(0.0);
        r_ucon[i1] = R(0.0);
        r_bcon[i1] = R(0.0);
        for (int j1 = 0; j1 < 4; j1++) {
            l_ucon[i1] += ucl[j1] * mudi[i1 * 4 + j1];
            l_bcon[i1] += bcl[j1] * mudi[i1 * 4 + j1];
            r_ucon[i1] += ucr[j1] * mudi[i1 * 4 + j1];
            r_bcon[i1] += bcr[j1] * mudi[i1 * 4 + j1];
        }
    }
    FtFace<R> f;
    primtoflux_ft<R, TAU>(pl, l_ucon, l_bcon, DIR, gam, f.F_l);
    primtoflux_ft<R, TAU>(pr, r_ucon, r_bcon, DIR, gam, f.F_r);
    primtoflux_ft<R, TAU>(pl, l_ucon, l_bcon, 0, gam, f.U_l);
    primtoflux_ft<R, TAU>(pr, r_ucon, r_bcon, 0, gam, f.U_r);
    R cmax_l, cmin_l, cmax_r, cmin_r;
    vchar_ft(pl, l_ucon, l_bcon, DIR, gam, &cmax_l, &cmin_l);
    vchar_ft(pr, r_ucon, r_bcon, DIR, gam, &cmax_r, &cmin_r);
    const R cmax_roe = mymax(cmax_r, cmax_l);
    const R cmin_roe = mymin(cmin_r, cmin_l);
    f.lu0 = l_ucon[0];
    f.lud = l_ucon[DIR];
    f.ru0 = r_ucon[0];
    f.rud = r_ucon[DIR];
    if constexpr (TAU) {
        f.smd_l = ft_smd(pl, l_ucon, l_bcon, DIR, gam);
        f.smd_r = ft_smd(pr, r_ucon, r_bcon, DIR, gam);
    }

    R FT[2][NPR];
    if (cmax_roe <= int_velocity) {
        for (int k = 0; k < NPR; k++) FT[0][k] = f.U_r[k];
        for (int k = 0; k < NPR; k++) FT[1][k] = f.F_r[k];
    } else if (cmin_roe >= int_velocity) {
        for (int k = 0; k < NPR; k++) FT[0][k] = f.U_l[k];
        for (int k = 0; k < NPR; k++) FT[1][k] = f.F_l[k];
    } else {
        for (int k = 0; k < NPR; k++)
            f.HLL[0][k] = dv(f.F_l[k] - f.F_r[k] + cmax_roe * f.U_r[k] - cmin_roe * f.U_l[k],
                             cmax_roe - cmin_roe + R(kSMALL));
        for (int k = 0; k < NPR; k++)
            f.HLL[1][k] = dv(cmax_roe * f.F_l[k] - cmin_roe * f.F_r[k] + cmax_roe * cmin_roe * (f.U_r[k] - f.U_l[k]),
                             cmax_roe - cmin_roe + R(kSMALL));
        const R bn = f.HLL[0][DIR + B1 - 1];
        const R e0 = TAU ? f.HLL[0][UU] - f.HLL[0][RHO] : f.HLL[0][UU];  // the total energy
        const bool do_hydro = (fabs(bn * bn * l_ucon[0] * r_ucon[0]) < R(1.e-10) * fabs(e0));
        if (do_hydro) hllc_hydro<DIR, R, TAU>(f, int_velocity, cmin_roe, cmax_roe, FT);
        else hllc_mhd<DIR, R, TAU>(f, int_velocity, cmin_roe, cmax_roe, FT);
    }

    // stress-energy row: T^dir_nu = e^dir_(0) T^(0)_(b) e^(b)_nu + e^dir_(dir) T^(dir)_(b) e^(b)_nu
    const R m0 = mud[DIR * 4 + 0], md = mud[DIR * 4 + DIR];
    // FP32: the frame energy entries carry + rho u; the momentum rows take the
    // total (a sum of like-signed terms), the energy row is formed below
    R te[2] = {FT[0][UU], FT[1][UU]};
    if constexpr (TAU) {
        FT[0][UU] = te[0] - FT[0][RHO];
        FT[1][UU] = te[1] - FT[1][RHO];
    }
    for (int j1 = 0; j1 < 4; j1++) {
        F1[j1 + UU] = R(0.);
        for (int j2 = 0; j2 < 4; j2++) {
            F1[j1 + UU] += FT[0][j2 + UU] * m0 * mudi[j2 * 4 + j1];
            F1[j1 + UU] += FT[1][j2 + UU] * md * mudi[j2 * 4 + j1];
        }
    }
    // TAU: T^dir_0 + rho u^dir less (1 - alpha) rho u^dir, from the spatial
    // frame rows and the frame energy with rest mass (e^(0)_0 = alpha); the
    // rest-mass flux times 1 - alpha (formed in FP64) is added below, so no
    // term cancels the rest mass
    R e_tau = R(0.);
    if constexpr (TAU) {
        for (int j2 = 1; j2 < 4; j2++) {
            e_tau += FT[0][j2 + UU] * m0 * mudi[j2 * 4];
            e_tau += FT[1][j2 + UU] * md * mudi[j2 * 4];
        }
        e_tau += te[0] * m0 * mudi[0];
        e_tau += te[1] * md * mudi[0];
    }
    // dual Maxwell row; its time component from antisymmetry
    FT[0][U3] = R(0.);
    FT[1][U3] = -FT[0][B1 - 1 + DIR];
    for (int j1 = 1; j1 < 4; j1++) {
        F1[j1 + U3] = R(0.);
        for (int j2 = 0; j2 < 4; j2++) {
            F1[j1 + U3] += FT[0][j2 + U3] * m0 * mud[j1 * 4 + j2];
            F1[j1 + U3] += FT[1][j2 + U3] * md * mud[j1 * 4 + j2];
        }
    }
    F1[RHO] = FT[0][RHO] * m0;
    F1[RHO] += FT[1][RHO] * md;
    F1[KTOT] = FT[0][KTOT] * m0;
    F1[KTOT] += FT[1][KTOT] * md;
    if constexpr (TAU) F1[UU] = e_tau + F1[RHO] * oma;
    else F1[UU] += F1[RHO];
}

// The face's coordinate velocity in the frame, g^{0 dir} / sqrt((g^{0 dir})^2 -
// g^00 g^{dir dir}) (GPU_hllc.h:88).
template <int DIR, class R> HD R face_velocity(const R *gcon)
{
    constexpr int dd = DIR == 1 ? 4 : DIR == 2 ? 7 : 9;
    return dv(gcon[DIR], R(sqrt(gcon[DIR] * gcon[DIR] - gcon[0] * gcon[dd])));
}

}  // namespace hamr
