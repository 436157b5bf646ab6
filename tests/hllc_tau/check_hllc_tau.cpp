// Host check of the HLLC solver's rest-mass-free energy form (the TAU switch of
// hllc_frame_flux, the form the FP32 fluxes use): evaluated in FP64 it equals
// the reference-order FP64 solver (fluxcalc2D_FT, GPU_hllc.h:3-247) on faces
// of the Kerr-Schild metric (a = 0.9375, r 1.5-31.5, theta 0.2-2.9) with
// random left and right states, hydrodynamic and magnetised.  Faces where the
// solver itself is ill-conditioned are counted, not compared: those where the
// same face with rho and u scaled by 1 + 1e-9 and B by its square root (which
// scales every stress and number flux by exactly 1 + 1e-9) gives fluxes off
// that scaling by more than 1e-11 of |F_rho| + |F|, as near the switch between
// the star states and HLL.
// Exit status 0 when every compared flux agrees within 1e-9 of (|F_rho| + |F|).
#include <cmath>
#include <cstdio>
#include <random>
#include "hamr_hllc.h"
extern "C" int hamr_face_tetrad(const double *gcov10, const double *gcon10, int dir, double *mud, double *mudi);
using namespace hamr;

static int pair10(int x, int y)
{
    static const int t[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
    return t[x][y];
}

// inverse of a 4x4 matrix by Gauss-Jordan elimination with partial pivoting
static void inverse4(const double *a, double *o)
{
    double m[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) m[i][j] = j < 4 ? a[i * 4 + j] : (j - 4 == i ? 1. : 0.);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
        for (int j = 0; j < 8; j++) std::swap(m[c][j], m[p][j]);
        const double d = m[c][c];
        for (int j = 0; j < 8; j++) m[c][j] /= d;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                const double f = m[r][c];
                for (int j = 0; j < 8; j++) m[r][j] -= f * m[c][j];
            }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o[i * 4 + j] = m[i][j + 4];
}

// Kerr-Schild g_mu nu (t, r, theta, phi)
static void kerr_schild(double r, double th, double a, double *g)
{
    const double c = std::cos(th), s2 = std::sin(th) * std::sin(th), rho2 = r * r + a * a * c * c, z = 2. * r / rho2;
    for (int k = 0; k < 16; k++) g[k] = 0.;
    g[0] = -(1. - z);
    g[1] = g[4] = z;
    g[3] = g[12] = -z * a * s2;
    g[5] = 1. + z;
    g[7] = g[13] = -a * s2 * (1. + z);
    g[10] = rho2;
    g[15] = s2 * (rho2 + a * a * s2 * (1. + z));
}

struct Face {
    double gcov[10], gcon[10], mud[16], mudi[16], oma;
};

template <int DIR, bool TAU> static void flux(const Face &F, const double *pl0, const double *pr0, double scale, double *out)
{
    Geom<double> G;
    for (int k = 0; k < 10; k++) {
        G.gcov[k] = F.gcov[k];
        G.gcon[k] = F.gcon[k];
    }
    G.g = 1.;
    G.oma = 0.;
    double pl[NPR], pr[NPR];
    // rho, u -> s rho, s u and B -> sqrt(s) B leave the velocities and every
    // speed unchanged and scale the stress and number fluxes by s
    for (int k = 0; k < NPR; k++) {
        const double f = (k == RHO || k == UU) ? scale : (k >= B1 && k <= B3) ? std::sqrt(scale) : 1.;
        pl[k] = pl0[k] * f;
        pr[k] = pr0[k] * f;
    }
    State<double> ql, qr;
    get_state(pl, G, ql);
    get_state(pr, G, qr);
    hllc_frame_flux<DIR, double, TAU>(pl, pr, ql.ucon, ql.bcon, qr.ucon, qr.bcon, F.mud, F.mudi,
                                      face_velocity<DIR>(G.gcon), 13. / 9., F.oma, out);
}

template <int DIR> static void run(std::mt19937_64 &rng, double *worst, long *compared, long *skipped)
{
    std::uniform_real_distribution<double> U(0., 1.);
    for (int it = 0; it < 40000; it++) {
        const double r = 1.5 + 30. * U(rng), th = 0.2 + 2.7 * U(rng);
        double g4[16], gi4[16];
        kerr_schild(r, th, 0.9375, g4);
        inverse4(g4, gi4);
        Face F;
        for (int x = 0; x < 4; x++)
            for (int y = x; y < 4; y++) {
                F.gcov[pair10(x, y)] = g4[x * 4 + y];
                F.gcon[pair10(x, y)] = gi4[x * 4 + y];
            }
        hamr_face_tetrad(F.gcov, F.gcon, DIR, F.mud, F.mudi);
        F.oma = 1. - 1. / std::sqrt(-F.gcon[0]);
        const double vs = (it % 3 == 0) ? 0.02 : 0.4, bs = (it % 4 == 0) ? 0. : (it % 4 == 1 ? 0.01 : 0.5);
        double p[2][NPR];
        for (int s = 0; s < 2; s++) {
            p[s][RHO] = 0.1 + U(rng);
            p[s][UU] = p[s][RHO] * (0.001 + 0.1 * U(rng));
            for (int k = U1; k <= U3; k++) p[s][k] = vs * (2. * U(rng) - 1.);
            for (int k = B1; k <= B3; k++) p[s][k] = bs * (2. * U(rng) - 1.);
            p[s][KTOT] = 0.;
        }
        p[1][B1 + DIR - 1] = p[0][B1 + DIR - 1];
        double ref[NPR], tau[NPR], nudge[NPR];
        flux<DIR, false>(F, p[0], p[1], 1., ref);
        flux<DIR, true>(F, p[0], p[1], 1., tau);
        flux<DIR, false>(F, p[0], p[1], 1. + 1e-9, nudge);  // the same face scaled by 1 + 1e-9
        bool well = true;
        for (int k = 0; k < B1; k++) {
            const double expect = ref[k] * (1. + 1e-9);
            if (std::fabs(nudge[k] - expect) > 1e-11 * (std::fabs(ref[RHO]) + std::fabs(ref[k]))) well = false;
        }
        if (!well) {
            (*skipped)++;
            continue;
        }
        (*compared)++;
        for (int k = 0; k < NPR - 1; k++) {
            const double e = std::fabs(tau[k] - ref[k]) / (std::fabs(ref[RHO]) + std::fabs(ref[k]) + 1e-300);
            if (e > worst[k]) worst[k] = e;
        }
    }
}

int main()
{
    std::mt19937_64 rng(11);
    double worst[NPR] = {0.};
    long compared = 0, skipped = 0;
    run<1>(rng, worst, &compared, &skipped);
    run<2>(rng, worst, &compared, &skipped);
    run<3>(rng, worst, &compared, &skipped);
    double w = 0.;
    for (int k = 0; k < NPR - 1; k++) {
        printf("var %d: max |F_tau - F| / (|F_rho| + |F|) = %.3e\n", k, worst[k]);
        w = std::fmax(w, worst[k]);
    }
    printf("%ld faces compared, %ld ill-conditioned faces skipped\n", compared, skipped);
    return (w <= 1e-9 && compared > 60000) ? 0 : 1;
}
