// hamr_setup.cpp -- host set-up: block table, coordinates, Kerr-Schild metric in
// code coordinates, and the Fishbone-Moncrief torus initial state.
//
// Coordinates follow coord.c (coord :26, bl_coord :111, vofx_matthewcoords :147
// with the uniform-theta parameters BRAVO=0 TANGO=1 CHARLIE=0, set_points :327,
// set_gridparam :357); the metric follows metric.c (gcov_func_spherical :13,
// numerical dx/dx' Jacobian dxdxp_func :501, gdet_func :392, gcon_func :410,
// conn_func :430 with EPS = 1e-5).  The torus follows init.c:808-1110
// (init_torus with rin = 6, rmax = 12; lfish_calc of init_torus.c:501;
// coord_transform of init.c:3218); the field is a poloidal loop from
// A_phi ~ max(rho/rho_max - 0.2, 0) on cell corners, curled onto the faces.
#include "hamr_grid.h"
#include "hamr_phys.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hamr {

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

Layout make_layout(const GridDesc &d)
{
    Layout L;
    for (int k = 0; k < 3; k++) L.D[k] = d.bs[k] + 2 * NG;
    L.ncell = (int64_t)L.D[0] * L.D[1] * L.D[2];
    L.ksize = round_up(L.ncell, 64);
    L.ncol = (int64_t)L.D[0] * L.D[1];
    L.kcol = round_up(L.ncol, 64);
    L.nblk = d.nb[0] * d.nb[1] * d.nb[2];
    L.ngeo = d.nb[0] * d.nb[1];
    return L;
}

std::vector<BlockDev> build_block_table(const GridDesc &d, int *ngeo_out)
{
    const int nb1 = d.nb[0], nb2 = d.nb[1], nb3 = d.nb[2];
    auto id = [&](int a, int b, int c) { return (a * nb2 + b) * nb3 + c; };
    std::vector<BlockDev> t(nb1 * nb2 * nb3);
    for (int c1 = 0; c1 < nb1; c1++)
        for (int c2 = 0; c2 < nb2; c2++)
            for (int c3 = 0; c3 < nb3; c3++) {
                BlockDev &b = t[id(c1, c2, c3)];
                std::memset(&b, 0, sizeof b);
                b.geo = c1 * nb2 + c2;
                b.c1 = c1;
                b.c2 = c2;
                b.c3 = c3;
                const bool flat = d.minkowski != 0;
                // x1
                if (c1 > 0) { b.nbr[0] = id(c1 - 1, c2, c3); b.kind[0] = FK_COPY; }
                else if (flat) { b.nbr[0] = id(nb1 - 1, c2, c3); b.kind[0] = FK_COPY; }
                else { b.nbr[0] = -1; b.kind[0] = FK_OUTFLOW; }
                if (c1 < nb1 - 1) { b.nbr[1] = id(c1 + 1, c2, c3); b.kind[1] = FK_COPY; }
                else if (flat) { b.nbr[1] = id(0, c2, c3); b.kind[1] = FK_COPY; }
                else { b.nbr[1] = -1; b.kind[1] = FK_OUTFLOW; }
                // x2: polar axis uses the partner half a turn away in phi
                const int cpi = (c3 + nb3 / 2) % nb3;
                if (c2 > 0) { b.nbr[2] = id(c1, c2 - 1, c3); b.kind[2] = FK_COPY; }
                else if (flat) { b.nbr[2] = id(c1, nb2 - 1, c3); b.kind[2] = FK_COPY; }
                else { b.nbr[2] = id(c1, 0, cpi); b.kind[2] = FK_POLE; b.pole |= 1; }
                if (c2 < nb2 - 1) { b.nbr[3] = id(c1, c2 + 1, c3); b.kind[3] = FK_COPY; }
                else if (flat) { b.nbr[3] = id(c1, 0, c3); b.kind[3] = FK_COPY; }
                else { b.nbr[3] = id(c1, nb2 - 1, cpi); b.kind[3] = FK_POLE; b.pole |= 2; }
                // x3 periodic
                b.nbr[4] = id(c1, c2, (c3 + nb3 - 1) % nb3);
                b.kind[4] = FK_COPY;
                b.nbr[5] = id(c1, c2, (c3 + 1) % nb3);
                b.kind[5] = FK_COPY;
                b.push3[0] = b.push3[1] = -1;
            }
    // per-block layout: uniform sizes, slabs in natural order
    const Layout L = make_layout(d);
    const GridCoords gc = grid_coords(d);
    for (int n = 0; n < (int)t.size(); n++) {
        BlockDev &b = t[n];
        b.level = 0;
        b.o[0] = b.c1 * d.bs[0];
        b.o[1] = b.c2 * d.bs[1];
        b.o[2] = b.c3 * d.bs[2];
        for (int k = 0; k < 3; k++) {
            b.bs[k] = d.bs[k];
            b.D[k] = L.D[k];
            b.dx[k] = gc.dx[k];
        }
        b.ks = (int)L.ksize;
        b.kcol = (int)L.kcol;
        b.off = (int64_t)n * L.ksize;
        b.moff = (int64_t)b.geo * L.kcol;
    }
    if (ngeo_out) *ngeo_out = nb1 * nb2;
    return t;
}

std::vector<int> uniform_push_table(std::vector<BlockDev> &t)
{
    std::vector<int> tab;
    for (BlockDev &b : t) {
        const int ntile = ((b.bs[0] + kPushTile - 1) / kPushTile) * ((b.bs[1] + kPushTile - 1) / kPushTile);
        for (int side = 0; side < 2; side++) {
            b.push3[side] = (int)tab.size();
            tab.insert(tab.end(), ntile, b.nbr[4 + side]);
        }
    }
    return tab;
}

std::vector<PushEntry> push_entries(std::vector<BlockDev> &blk, const std::vector<int> &tab)
{
    std::vector<PushEntry> out(tab.size(), PushEntry{-1, 0, 0, 0, 0});
    for (BlockDev &B : blk) {
        const int ntile = ((B.bs[0] + kPushTile - 1) / kPushTile) * ((B.bs[1] + kPushTile - 1) / kPushTile);
        for (int side = 0; side < 2; side++) {
            if (B.push3[side] < 0) continue;
            B.push3[side] &= ~kPushWhole;
            for (int t = 0; t < ntile; t++) {
                const int m = tab[B.push3[side] + t];
                if (m < 0) continue;
                const BlockDev &M = blk[m];
                PushEntry &E = out[B.push3[side] + t];
                E.m = m;
                E.ks = M.ks;
                E.off = M.off;
                E.delta = kPushGeneral;
                if (M.D[1] == B.D[1] && M.D[2] == B.D[2]) {
                    const int64_t d = ((int64_t)(B.o[0] - M.o[0]) * B.D[1] + (B.o[1] - M.o[1])) * B.D[2] +
                                      (side ? -B.bs[2] : M.bs[2]);
                    if (d > -2147483647 && d < 2147483647) E.delta = (int)d;
                }
            }
            bool whole = true;
            const PushEntry &F = out[B.push3[side]];
            for (int t = 1; t < ntile; t++) {
                const PushEntry &E = out[B.push3[side] + t];
                whole = whole && E.m == F.m && E.delta == F.delta;
            }
            if (whole && F.delta != kPushGeneral) B.push3[side] |= kPushWhole;
        }
    }
    return out;
}

// The reference's block numbering of a single-level grid: the level-0 Z-order
// of AMR_set_coord (AMR.c:188-275).  Block number k's coordinates follow from
// k alone: walking the power-of-two scales p = 2^s from the largest down, the
// counter is compared with the volume of the (clipped) p-cube at the current
// corner, and each of the seven sub-cube moves of that scale (+x3, +x2, +x2x3,
// +x1, +x1x3, +x1x2, +x1x2x3, each undoing the earlier moves it replaces) is
// taken when the counter covers it, subtracting the blocks it skips.  A move
// along a direction is possible only when a full p-step fits before the last
// block along it.  The steps and their order are the reference's, so that the
// numbers (and the rdump<k> file names they give) are the same.
static void ref_block_coords(int k, const int nb[3], int c[3])
{
    const int maxs = (int)(log((double)std::max(nb[0], std::max(nb[1], nb[2]))) / log(2.));
    int cnt = k;
    c[0] = c[1] = c[2] = 0;
    for (int s = maxs; s >= 0; s--) {
        const int p = 1 << s;
        const int in1 = std::min(p, nb[0] - c[0] - 1), in2 = std::min(p, nb[1] - c[1] - 1),
                  in3 = std::min(p, nb[2] - c[2] - 1);
        const bool f1 = in1 == p, f2 = in2 == p, f3 = in3 == p;
        auto w = [&](int d) { return std::min(p, nb[d] - c[d]); };
        auto cube = [&]() { return w(2) * w(1) * w(0); };
        if (f3 && cnt >= cube()) { cnt -= in3 * w(1) * w(0); c[2] += in3; }
        if (f2 && cnt >= cube()) { cnt -= w(2) * in2 * w(0); if (f3) c[2] -= in3; c[1] += in2; }
        if (f3 && f2 && cnt >= cube()) { cnt -= in3 * w(1) * w(0); c[2] += in3; }
        if (f1 && cnt >= cube()) {
            cnt -= w(2) * w(1) * in1;
            if (f3) c[2] -= in3;
            if (f2) c[1] -= in2;
            c[0] += in1;
        }
        if (f1 && f3 && cnt >= cube()) { cnt -= in3 * w(1) * w(0); c[2] += in3; }
        if (f1 && f2 && cnt >= cube()) { cnt -= w(2) * in2 * w(0); if (f3) c[2] -= in3; c[1] += in2; }
        if (f1 && f2 && f3 && cnt >= cube()) { cnt -= in3 * w(1) * w(0); c[2] += in3; }
    }
}

std::vector<int> zorder(const GridDesc &d)
{
    const int nblk = d.nb[0] * d.nb[1] * d.nb[2];
    std::vector<int> out(nblk, -1), seen(nblk, 0);
    for (int k = 0; k < nblk; k++) {
        int c[3];
        ref_block_coords(k, d.nb, c);
        for (int a = 0; a < 3; a++)
            if (c[a] < 0 || c[a] >= d.nb[a]) return {};
        const int n = (c[0] * d.nb[1] + c[1]) * d.nb[2] + c[2];
        if (seen[n]++) return {};
        out[k] = n;
    }
    return out;
}

uint64_t hash64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// ---------------------------------------------------------------------------
GridCoords grid_coords(const GridDesc &d)
{
    GridCoords g;
    const int N1 = d.nb[0] * d.bs[0], N2 = d.nb[1] * d.bs[1], N3 = d.nb[2] * d.bs[2];
    if (d.minkowski) {
        for (int k = 0; k < 3; k++) g.startx[k] = 0.;
        g.dx[0] = 1. / N1;
        g.dx[1] = 1. / N2;
        g.dx[2] = 1. / N3;
        g.rin = g.rout = 0.;
        g.fractheta = 1.;
        return g;
    }
    g.rin = (d.rin > 0.) ? d.rin : 0.9 * (1. + std::sqrt(1. - d.a * d.a));
    g.rout = d.rout;
    g.fractheta = 1.0 - 2.0 / ((double)N2);
    g.startx[0] = std::log(g.rin);
    g.startx[1] = -1. + 1. * (1. - g.fractheta);
    g.startx[2] = 0.;
    g.dx[0] = (std::log(g.rout) - std::log(g.rin)) / (double)N1;
    g.dx[1] = 2. * g.fractheta / (double)N2;
    g.dx[2] = 2. * M_PI / (double)N3;
    return g;
}

// vofx_matthewcoords (coord.c:147) with the uniform-theta parameters, plus the
// axis excision of bl_coord (COORDSINGFIX, SINGSMALL = 1e-20).
void bl_coord(const GridDesc &, const double *X, double *r, double *th)
{
    const double A1 = 1., A2 = 1., A3 = 1.;
    double V1 = std::exp(X[1]);
    double sign = 1., X_2 = (X[2] + 1.0) / 2.0, Xc = std::sqrt(X_2 * X_2), V2;
    if (X_2 < 0.0) sign = -1.;
    if (X_2 > 1.0) { sign = -1.; Xc = 2. - Xc; }
    if (X_2 >= 0.5) {
        Xc = 1. - Xc;
        V2 = M_PI - sign * (A1 * M_PI * Xc + M_PI * (1. - A1) * (A3 * std::pow(Xc, A2) + 0.50 / M_PI * std::sin(M_PI + 2. * M_PI * (A3 * std::pow(Xc, A2)))));
    } else {
        V2 = sign * (A1 * M_PI * Xc + M_PI * (1. - A1) * (A3 * std::pow(Xc, A2) + 0.50 / M_PI * std::sin(M_PI + 2. * M_PI * (A3 * std::pow(Xc, A2)))));
    }
    if (std::fabs(V2) < 1e-20) V2 = (V2 >= 0.0) ? 1e-20 : -1e-20;
    if (std::fabs(M_PI - V2) < 1e-20) V2 = (V2 >= M_PI) ? M_PI + 1e-20 : M_PI - 1e-20;
    *r = V1;
    *th = V2;
}

static void ks_gcov(double a, double r, double th, double g[4][4])
{
    double cth = std::cos(th), sth = std::sin(th), s2 = sth * sth, rho2 = r * r + a * a * cth * cth;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) g[i][j] = 0.;
    g[0][0] = (-1. + 2. * r / rho2);
    g[0][1] = (2. * r / rho2);
    g[0][3] = (-2. * a * r * s2 / rho2);
    g[1][0] = g[0][1];
    g[1][1] = (1. + 2. * r / rho2);
    g[1][3] = (-a * s2 * (1. + 2. * r / rho2));
    g[2][2] = rho2;
    g[3][0] = g[0][3];
    g[3][1] = g[1][3];
    g[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
}

static void dxdxp_fd(const GridDesc &d, const double *X, double J[4][4])
{
    for (int k = 0; k < 4; k++) {
        double Xh[4], Xl[4], Vh[4], Vl[4];
        for (int l = 0; l < 4; l++) Xh[l] = Xl[l] = X[l];
        Xh[k] += 0.00001;
        Xl[k] -= 0.00001;
        Vh[0] = Xh[0];
        Vl[0] = Xl[0];
        bl_coord(d, Xh, &Vh[1], &Vh[2]);
        bl_coord(d, Xl, &Vl[1], &Vl[2]);
        Vh[3] = Xh[3];
        Vl[3] = Xl[3];
        for (int j = 0; j < 4; j++) J[j][k] = (Vh[j] - Vl[j]) / (Xh[k] - Xl[k]);
    }
}

void metric_code(const GridDesc &d, const double *X, double gp[4][4])
{
    if (d.minkowski) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) gp[i][j] = (i == j) ? (i == 0 ? -1. : 1.) : 0.;
        return;
    }
    double r, th, g[4][4], J[4][4];
    bl_coord(d, X, &r, &th);
    ks_gcov(d.a, r, th, g);
    dxdxp_fd(d, X, J);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            gp[i][j] = 0.;
            for (int k = 0; k < 4; k++)
                for (int l = 0; l < 4; l++) gp[i][j] += g[k][l] * J[k][i] * J[l][j];
        }
}

// LU decomposition with partial pivoting (Doolittle, in place); returns 1 if singular.
static int lu4(double a[4][4], int perm[4], double *sign)
{
    *sign = 1.;
    for (int i = 0; i < 4; i++) perm[i] = i;
    for (int k = 0; k < 4; k++) {
        int p = k;
        double big = std::fabs(a[k][k]);
        for (int i = k + 1; i < 4; i++)
            if (std::fabs(a[i][k]) > big) { big = std::fabs(a[i][k]); p = i; }
        if (big == 0.) return 1;
        if (p != k) {
            for (int j = 0; j < 4; j++) std::swap(a[k][j], a[p][j]);
            std::swap(perm[k], perm[p]);
            *sign = -*sign;
        }
        for (int i = k + 1; i < 4; i++) {
            a[i][k] /= a[k][k];
            for (int j = k + 1; j < 4; j++) a[i][j] -= a[i][k] * a[k][j];
        }
    }
    return 0;
}

double gdet_of(const double g[4][4])
{
    double a[4][4], s;
    int perm[4];
    std::memcpy(a, g, sizeof a);
    if (lu4(a, perm, &s)) return 0.;
    double det = s;
    for (int i = 0; i < 4; i++) det *= a[i][i];
    return std::sqrt(std::fabs(det));
}

static void lu_solve4(const double a[4][4], const int perm[4], const double *b, double *x)
{
    double y[4];
    for (int i = 0; i < 4; i++) {
        double v = b[perm[i]];
        for (int j = 0; j < i; j++) v -= a[i][j] * y[j];
        y[i] = v;
    }
    for (int i = 3; i >= 0; i--) {
        double v = y[i];
        for (int j = i + 1; j < 4; j++) v -= a[i][j] * x[j];
        x[i] = v / a[i][i];
    }
}

// Inverse by LU with one step of iterative refinement (the metric's entries
// span several decades at large r).
int invert4(const double g[4][4], double inv[4][4])
{
    double a[4][4], s;
    int perm[4];
    std::memcpy(a, g, sizeof a);
    if (lu4(a, perm, &s)) return 1;
    for (int c = 0; c < 4; c++) {
        double e[4] = {0., 0., 0., 0.}, x[4], r[4], dx[4];
        e[c] = 1.;
        lu_solve4(a, perm, e, x);
        for (int i = 0; i < 4; i++) {
            double v = e[i];
            for (int j = 0; j < 4; j++) v -= g[i][j] * x[j];
            r[i] = v;
        }
        lu_solve4(a, perm, r, dx);
        for (int i = 0; i < 4; i++) inv[i][c] = x[i] + dx[i];
    }
    return 0;
}

void conn_code(const GridDesc &d, const double *X, const double gcon[4][4], double conn[4][4][4])
{
    double c[4][4][4], tmp[4][4][4];
    for (int k = 0; k < 4; k++) {
        double Xh[4], Xl[4], gh[4][4], gl[4][4];
        for (int l = 0; l < 4; l++) Xh[l] = Xl[l] = X[l];
        Xh[k] += 1.e-5;
        Xl[k] -= 1.e-5;
        metric_code(d, Xh, gh);
        metric_code(d, Xl, gl);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) c[i][j][k] = (gh[i][j] - gl[i][j]) / (Xh[k] - Xl[k]);
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) tmp[i][j][k] = 0.5 * (c[j][i][k] + c[k][i][j] - c[k][j][i]);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) {
                conn[i][j][k] = 0.;
                for (int l = 0; l < 4; l++) conn[i][j][k] += gcon[i][l] * tmp[l][j][k];
            }
}

void coord_of(const GridDesc &d, const GridCoords &gc, int i, int j, int pos, double X[4])
{
    coord_of_level(d, gc, 0, i, j, pos, X);
}

void coord_of_level(const GridDesc &d, const GridCoords &gc, int level, int i, int j, int pos, double X[4])
{
    // spacings of the level: exact halvings of the base spacing
    const double sc = std::ldexp(1.0, -level);
    const double dx0 = gc.dx[0] * sc, dx1 = gc.dx[1] * sc, dx2 = gc.dx[2] * sc;
    const int N2 = (d.nb[1] * d.bs[1]) << level;
    X[0] = 0.;
    X[3] = gc.startx[2] + 0.5 * dx2;
    if (d.minkowski) {
        X[1] = gc.startx[0] + ((pos == FACE1 || pos == CORN) ? i : i + 0.5) * dx0;
        X[2] = gc.startx[1] + ((pos == FACE2 || pos == CORN) ? j : j + 0.5) * dx1;
        return;
    }
    int jl = j;
    if (j < 0) jl = -j - 1;
    if (j >= N2) jl = 2 * N2 - 1 - j;
    if (j == N2 && pos == FACE2) jl = j;
    if (pos == FACE1) { X[1] = gc.startx[0] + i * dx0; X[2] = gc.startx[1] + (jl + 0.5) * dx1; }
    else if (pos == FACE2) { X[1] = gc.startx[0] + (i + 0.5) * dx0; X[2] = gc.startx[1] + jl * dx1; }
    else if (pos == CORN) { X[1] = gc.startx[0] + i * dx0; X[2] = gc.startx[1] + jl * dx1; }
    else { X[1] = gc.startx[0] + (i + 0.5) * dx0; X[2] = gc.startx[1] + (jl + 0.5) * dx1; }
    if (j < 0) {
        X[2] = X[2] + 1;
        X[2] = -X[2];
        X[2] = X[2] - 1;
    } else if (j == N2 && pos == FACE2) {
    } else if (j >= N2) {
        X[2] = X[2] + 1;
        X[2] = 4. - X[2];
        X[2] = X[2] - 1;
    }
}

static void pack10(const double g[4][4], double *o, int64_t stride)
{
    static const int pi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3}, pj[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    for (int s = 0; s < 10; s++) o[s * stride] = g[pi[s]][pj[s]];
}

void fill_geometry(const GridDesc &d, const Layout &L, int c1, int c2, double *gcov, double *gcon,
                   double *gdet, double *conn, double *rcen)
{
    fill_geometry_level(d, 0, c1 * d.bs[0], c2 * d.bs[1], L.D[0], L.D[1], L.kcol, gcov, gcon, gdet, conn, rcen);
}

void fill_geometry_level(const GridDesc &d, int level, int o1, int o2, int D1, int D2, int64_t kcol, double *gcov,
                         double *gcon, double *gdet, double *conn, double *rcen)
{
    GridCoords gc = grid_coords(d);
    static const int pi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3}, pj[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int il = 0; il < D1; il++)
        for (int jl = 0; jl < D2; jl++) {
            const int i = o1 + il - NG, j = o2 + jl - NG;
            const int64_t col = (int64_t)il * D2 + jl;
            for (int pos = 0; pos < NPOS; pos++) {
                double X[4], g[4][4], gi[4][4];
                coord_of_level(d, gc, level, i, j, pos, X);
                metric_code(d, X, g);
                invert4(g, gi);
                pack10(g, gcov + ((int64_t)pos * 10) * kcol + col, kcol);
                pack10(gi, gcon + ((int64_t)pos * 10) * kcol + col, kcol);
                gdet[(int64_t)pos * kcol + col] = gdet_of(g);
                if (pos == CENT) {
                    double cn[4][4][4], r, th;
                    conn_code(d, X, gi, cn);
                    for (int lam = 0; lam < 4; lam++)
                        for (int s = 0; s < 10; s++) conn[((int64_t)lam * 10 + s) * kcol + col] = cn[lam][pi[s]][pj[s]];
                    if (d.minkowski) rcen[col] = 1.;
                    else {
                        bl_coord(d, X, &r, &th);
                        rcen[col] = r;
                    }
                }
            }
        }
}

// ---------------------------------------------------------------------------
// Fishbone-Moncrief torus
// ---------------------------------------------------------------------------
// Face tetrads (set_Mud, hllc.c:20-178).  For face direction a the other two
// directions are taken cyclically, b = a % 3 + 1 and c = b % 3 + 1:
//   e_(0) = A g^{0mu} (A = -(-g^00)^-1/2: the normal observer),
//   e_(a) from the normal and g^{a mu}, e_(b) in the (b, c) plane orthogonal to
//   e_(c) = g_cc^-1/2 d_c.
// The products keep the reference's operand order (same roundings).
static int pair10(int x, int y)
{
    static const int t[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
    return t[x][y];
}

void face_tetrad(const double *gcov, const double *gcon, int dir, double *mud, double *mudi)
{
    const int a = dir, b = dir % 3 + 1, c = b % 3 + 1;
    auto gu = [&](int x, int y) { return gcon[pair10(x, y)]; };
    auto gd = [&](int x, int y) { return gcov[pair10(x, y)]; };
    const double A = -std::pow(-gu(0, 0), -0.5);
    const double B = std::pow((gu(0, 0)) * (gu(0, 0) * gu(a, a) - gu(0, a) * gu(0, a)), -0.5);
    const double C = std::pow(gd(c, c), -0.5);
    const double D = std::pow((gd(c, c)) * (gd(b, b) * gd(c, c) - gd(b, c) * gd(b, c)), -0.5);
    for (int k = 0; k < 16; k++) mud[k] = mudi[k] = 0.;
    double *M = mud, *Mi = mudi;
    M[0 * 4 + 0] = A * gu(0, 0);
    M[a * 4 + 0] = A * gu(0, a);
    M[a * 4 + a] = B * (gu(0, a) * gu(0, a) - gu(0, 0) * gu(a, a));
    M[b * 4 + 0] = A * gu(0, b);
    M[b * 4 + a] = B * (gu(0, a) * gu(0, b) - gu(0, 0) * gu(a, b));
    M[b * 4 + b] = D * gd(c, c);
    M[c * 4 + 0] = A * gu(0, c);
    M[c * 4 + a] = B * (gu(0, a) * gu(0, c) - gu(0, 0) * gu(a, c));
    M[c * 4 + b] = -D * gd(b, c);
    M[c * 4 + c] = C;
    const double E = gu(0, a) * gu(a, b) - gu(a, a) * gu(0, b);
    const double F = gu(0, a) * gu(0, b) - gu(0, 0) * gu(a, b);
    const double G = gu(0, a) * gu(a, c) - gu(a, a) * gu(0, c);
    const double H = gu(0, a) * gu(0, c) - gu(0, 0) * gu(a, c);
    Mi[0 * 4 + 0] = -A;
    Mi[a * 4 + 0] = B * gu(0, a);
    Mi[a * 4 + a] = -B * gu(0, 0);
    Mi[b * 4 + 0] = B * B * E * gu(0, 0) / (D * gd(c, c));
    Mi[b * 4 + a] = B * B * F * gu(0, 0) / (D * gd(c, c));
    Mi[b * 4 + b] = 1. / (D * gd(c, c));
    Mi[c * 4 + 0] = (B * B / C) * gu(0, 0) * (G + E * gd(b, c) / gd(c, c));
    Mi[c * 4 + a] = (B * B / C) * gu(0, 0) * (H + F * gd(b, c) / gd(c, c));
    Mi[c * 4 + b] = (1. / C) * gd(b, c) / gd(c, c);
    Mi[c * 4 + c] = 1. / C;
}

void fill_tetrads(const double *gcov, const double *gcon, int64_t ncol, int64_t kcol, double *tet)
{
    std::memset(tet, 0, sizeof(double) * 96 * (size_t)kcol);
    static const int posof[3] = {FACE1, FACE2, FACE3};
    for (int t = 0; t < 3; t++)
        for (int64_t col = 0; col < ncol; col++) {
            double gc[10], gi[10], m[16], mi[16];
            for (int s = 0; s < 10; s++) {
                gc[s] = gcov[((int64_t)posof[t] * 10 + s) * kcol + col];
                gi[s] = gcon[((int64_t)posof[t] * 10 + s) * kcol + col];
            }
            face_tetrad(gc, gi, t + 1, m, mi);
            for (int e = 0; e < 16; e++) {
                tet[((int64_t)t * 32 + e) * kcol + col] = m[e];
                tet[((int64_t)t * 32 + 16 + e) * kcol + col] = mi[e];
            }
        }
}

static double lfish_calc(double a, double r)
{
    return (((a * a - 2. * a * std::sqrt(r) + r * r) *
             ((-2. * a * r * (a * a - 2. * a * std::sqrt(r) + r * r)) / std::sqrt(2. * a * std::sqrt(r) + (-3. + r) * r) +
              ((a + (-2. + r) * std::sqrt(r)) * (r * r * r + a * a * (2. + r))) / std::sqrt(1 + (2. * a) / std::pow(r, 1.5) - 3. / r))) /
            (r * r * r * std::sqrt(2. * a * std::sqrt(r) + (-3. + r) * r) * (a * a + (-2. + r) * r)));
}

struct TorusEval {
    double a, gam, rin, kappa, l, lnh_in;
};

static double lnh_at(const TorusEval &T, double r, double th)
{
    const double a = T.a, l = T.l;
    double sth = std::sin(th), cth = std::cos(th);
    double DD = r * r - 2. * r + a * a;
    double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
    double SS = r * r + a * a * cth * cth;
    return 0.5 * std::log((1. + std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * sth * AA * sth))) / (SS * DD / AA)) -
           0.5 * std::sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) - 2. * a * r * l / AA - T.lnh_in;
}

static TorusEval make_torus(const GridDesc &d, const TorusDesc &t)
{
    TorusEval T;
    T.a = d.a;
    T.gam = d.gam;
    T.rin = t.rin;
    T.kappa = t.kappa;
    T.l = lfish_calc(d.a, t.rmax);
    T.lnh_in = 0.;
    T.lnh_in = lnh_at(T, t.rin, M_PI / 2.);
    return T;
}

// torus density (unnormalised) at BL (r, th); negative outside the torus
static double torus_rho(const TorusEval &T, double r, double th, double *u_out)
{
    double lnh = (r >= T.rin) ? lnh_at(T, r, th) : 1.;
    if (lnh < 0. || r < T.rin) {
        if (u_out) *u_out = 0.;
        return -1.;
    }
    double hm1 = std::exp(lnh) - 1.;
    double rho = std::pow(hm1 * (T.gam - 1.) / (T.kappa * T.gam), 1. / (T.gam - 1.));
    if (u_out) *u_out = T.kappa * std::pow(rho, T.gam) / (T.gam - 1.);
    return rho;
}

// BL u^phi of the torus -> relative 4-velocity in code coordinates (coord_transform)
static void torus_velocity(const GridDesc &d, const TorusEval &T, const double *X, double r, double th,
                           const double gcov[4][4], const double gcon[4][4], double *utcon)
{
    const double a = T.a, l = T.l;
    double sth = std::sin(th), cth = std::cos(th);
    double DD = r * r - 2. * r + a * a;
    double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
    double SS = r * r + a * a * cth * cth;
    double expm2chi = SS * SS * DD / (AA * AA * sth * sth);
    double up1 = std::sqrt((-1. + std::sqrt(1. + 4. * l * l * expm2chi)) / 2.);
    double up = 2. * a * r * std::sqrt(1. + up1 * up1) / std::sqrt(AA * SS * DD) + std::sqrt(SS / AA) * up1 / sth;
    // Boyer-Lindquist metric
    double rho2 = SS, s2 = sth * sth;
    double bg00 = -(1. - 2. * r / rho2), bg03 = -2. * a * r * s2 / rho2;
    double bg33 = s2 * (r * r + a * a + 2. * a * a * r * s2 / rho2);
    double ucon[4] = {0., 0., 0., up};
    double Aq = bg00, Bq = 2. * (bg03 * up), Cq = 1. + bg33 * up * up;
    double discr = Bq * Bq - 4. * Aq * Cq;
    ucon[0] = (-Bq - std::sqrt(discr)) / (2. * Aq);
    // BL -> KS (u^r = 0, so only u^t and u^phi pick up terms that vanish here)
    double tmp[4] = {ucon[0] + 2. * r / DD * ucon[1], ucon[1], ucon[2], ucon[3] + a / DD * ucon[1]};
    // KS -> code coordinates
    double J[4][4], Ji[4][4], uc[4];
    dxdxp_fd(d, X, J);
    invert4(J, Ji);
    for (int i = 0; i < 4; i++) {
        uc[i] = 0.;
        for (int j = 0; j < 4; j++) uc[i] += Ji[i][j] * tmp[j];
    }
    double alpha = 1. / std::sqrt(-gcon[0][0]);
    double gamma = alpha * uc[0];
    for (int i = 1; i < 4; i++) {
        double beta = gcon[0][i] * alpha * alpha;
        utcon[i] = uc[i] + gamma * beta / alpha;
    }
    (void)gcov;
}

std::string init_torus(const GridDesc &d, const std::vector<BlockDev> &blocks, const TorusDesc &t,
                       const double *gcovA, const double *gconA, const double *gdetA, double *p, double *ps)
{
    if (d.minkowski) return "init_torus: grid is a Minkowski box";
    const GridCoords gc = grid_coords(d);
    const TorusEval T = make_torus(d, t);
    const int N2 = d.nb[1] * d.bs[1], N3 = d.nb[2] * d.bs[2];
    auto cell = [&](const BlockDev &B, int i, int j, int z) { return ((int64_t)i * B.D[1] + j) * B.D[2] + z; };
    auto P = [&](const BlockDev &B, int k) { return p + NPR * B.off + (int64_t)k * B.ks; };
    auto PS = [&](const BlockDev &B, int k) { return ps + 3 * B.off + (int64_t)k * B.ks; };
    auto gdet_at = [&](const BlockDev &B, int pos, int64_t col) { return gdetA[B.moff * NPOS + (int64_t)pos * B.kcol + col]; };
    auto metric_at = [&](const BlockDev &B, int pos, int64_t col, double g[4][4], double gi[4][4]) {
        static const int pi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3}, pj[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
        for (int s = 0; s < 10; s++) {
            const int64_t e = B.moff * NPOS * 10 + ((int64_t)pos * 10 + s) * B.kcol + col;
            g[pi[s]][pj[s]] = g[pj[s]][pi[s]] = gcovA[e];
            gi[pi[s]][pj[s]] = gi[pj[s]][pi[s]] = gconA[e];
        }
    };
    const int nblk = (int)blocks.size();
    // pass 1: densities and internal energies, rho_max
    double rhomax = 0.;
    for (int b = 0; b < nblk; b++) {
        const BlockDev &B = blocks[b];
        for (int il = NG; il < NG + B.bs[0]; il++)
            for (int jl = NG; jl < NG + B.bs[1]; jl++) {
                const int i = B.o[0] + il - NG, j = B.o[1] + jl - NG;
                double X[4], r, th, u;
                coord_of_level(d, gc, B.level, i, j, CENT, X);
                bl_coord(d, X, &r, &th);
                double rho = torus_rho(T, r, th, &u);
                for (int zl = NG; zl < NG + B.bs[2]; zl++) {
                    const int64_t c = cell(B, il, jl, zl);
                    P(B, RHO)[c] = rho;
                    P(B, UU)[c] = u;
                }
                rhomax = std::max(rhomax, rho);
            }
    }
    if (!(rhomax > 0.)) return "init_torus: the torus does not intersect the grid";
    // pass 2: velocities, perturbation, normalisation, atmosphere
    for (int b = 0; b < nblk; b++) {
        const BlockDev &B = blocks[b];
        const int N2l = N2 << B.level, N3l = N3 << B.level;
        // the perturbation's cell key: global cell index on the block's level
        const uint64_t lkey = B.level ? hash64(0xA5A5A5A5ULL + (uint64_t)B.level) : 0ULL;
        for (int il = NG; il < NG + B.bs[0]; il++)
            for (int jl = NG; jl < NG + B.bs[1]; jl++) {
                const int i = B.o[0] + il - NG, j = B.o[1] + jl - NG;
                const int64_t col = (int64_t)il * B.D[1] + jl;
                double X[4], r, th, g[4][4], gi[4][4], ut[4] = {0., 0., 0., 0.};
                coord_of_level(d, gc, B.level, i, j, CENT, X);
                bl_coord(d, X, &r, &th);
                metric_at(B, CENT, col, g, gi);
                const double rho_t = P(B, RHO)[cell(B, il, jl, NG)];
                const bool in = rho_t > 0.;
                if (in) torus_velocity(d, T, X, r, th, g, gi, ut);
                const double sr = std::sqrt(r);
                const double rhoflr = d.rho_floor / (r * sr), uflr = d.u_floor / (r * r * sr);
                for (int zl = NG; zl < NG + B.bs[2]; zl++) {
                    const int64_t c = cell(B, il, jl, zl);
                    const int zg = B.o[2] + zl - NG;
                    double rho, u;
                    if (in) {
                        const uint64_t gid = ((uint64_t)(i + 8) * (uint64_t)(N2l + 16) + (uint64_t)(j + 8)) * (uint64_t)(N3l + 16) + (uint64_t)zg;
                        const double U01 = (double)(hash64(t.seed ^ (gid * 0x9E3779B97F4A7C15ULL) ^ lkey) >> 11) * (1.0 / 9007199254740992.0);
                        rho = P(B, RHO)[c] / rhomax;
                        u = P(B, UU)[c] * (1. + t.perturb * (U01 - 0.5)) / rhomax;
                    } else {
                        rho = 0.;
                        u = 0.;
                    }
                    if (!(rho >= rhoflr)) rho = rhoflr;
                    if (!(u >= uflr)) u = uflr;
                    P(B, RHO)[c] = rho;
                    P(B, UU)[c] = u;
                    P(B, U1)[c] = in ? ut[1] : 0.;
                    P(B, U2)[c] = in ? ut[2] : 0.;
                    P(B, U3)[c] = in ? ut[3] : 0.;
                    P(B, KTOT)[c] = (d.gam - 1.) * u * std::pow(rho, -d.gam);
                }
            }
    }
    // pass 3: vector potential on corners and face fields (unnormalised)
    auto aphi = [&](int level, int i, int j) {
        double X[4], r, th;
        coord_of_level(d, gc, level, i, j, CORN, X);
        bl_coord(d, X, &r, &th);
        double rho = torus_rho(T, r, th, nullptr);
        if (rho < 0.) return 0.;
        return std::max(rho / rhomax - t.aphi_cut, 0.);
    };
    for (int b = 0; b < nblk; b++) {
        const BlockDev &B = blocks[b];
        const double sc = std::ldexp(1.0, -B.level);
        const double dx1 = gc.dx[0] * sc, dx2 = gc.dx[1] * sc;
        for (int il = NG; il <= NG + B.bs[0]; il++)
            for (int jl = NG; jl <= NG + B.bs[1]; jl++) {
                const int i = B.o[0] + il - NG, j = B.o[1] + jl - NG;
                const int64_t col = (int64_t)il * B.D[1] + jl;
                const double A00 = aphi(B.level, i, j), A01 = aphi(B.level, i, j + 1), A10 = aphi(B.level, i + 1, j);
                const double gF1 = gdet_at(B, FACE1, col);
                const double gF2 = gdet_at(B, FACE2, col);
                const double b1 = (A01 - A00) / (dx2 * gF1);
                const double b2 = -(A10 - A00) / (dx1 * gF2);
                for (int zl = NG; zl <= NG + B.bs[2]; zl++) {
                    const int64_t c = cell(B, il, jl, zl);
                    if (jl < NG + B.bs[1] && zl < NG + B.bs[2]) PS(B, 0)[c] = b1;
                    if (il < NG + B.bs[0] && zl < NG + B.bs[2]) PS(B, 1)[c] = b2;
                    if (il < NG + B.bs[0] && jl < NG + B.bs[1]) PS(B, 2)[c] = 0.;
                }
            }
    }
    // cell-centred field from the faces, as the update kernel forms it
    auto centre_field = [&](const BlockDev &B) {
        for (int il = NG; il < NG + B.bs[0]; il++)
            for (int jl = NG; jl < NG + B.bs[1]; jl++) {
                const int64_t col = (int64_t)il * B.D[1] + jl;
                const double gF1a = gdet_at(B, FACE1, col), gF1b = gdet_at(B, FACE1, col + B.D[1]);
                const double gF2a = gdet_at(B, FACE2, col), gF2b = gdet_at(B, FACE2, col + 1);
                const double gF3 = gdet_at(B, FACE3, col), gC = gdet_at(B, CENT, col);
                for (int zl = NG; zl < NG + B.bs[2]; zl++) {
                    const int64_t c = cell(B, il, jl, zl);
                    const int64_t si = (int64_t)B.D[1] * B.D[2], sj = B.D[2];
                    double U1f = (PS(B, 0)[c] * gF1a + PS(B, 0)[c + si] * gF1b) / 2.0;
                    double U2f = (PS(B, 1)[c] * gF2a + PS(B, 1)[c + sj] * gF2b) / 2.0;
                    double U3f = (PS(B, 2)[c] * gF3 + PS(B, 2)[c + 1] * gF3) / 2.0;
                    P(B, B1)[c] = U1f / gC;
                    P(B, B2)[c] = U2f / gC;
                    P(B, B3)[c] = U3f / gC;
                }
            }
    };
    for (int b = 0; b < nblk; b++) centre_field(blocks[b]);
    // pass 4: normalise min beta over field-carrying cells
    double beta_min = 1e300;
    for (int b = 0; b < nblk; b++) {
        const BlockDev &B = blocks[b];
        for (int il = NG; il < NG + B.bs[0]; il++)
            for (int jl = NG; jl < NG + B.bs[1]; jl++) {
                const int64_t col = (int64_t)il * B.D[1] + jl;
                double g[4][4], gi[4][4];
                metric_at(B, CENT, col, g, gi);
                Geom<double> G;
                static const int pi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3}, pj[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
                for (int s = 0; s < 10; s++) { G.gcov[s] = g[pi[s]][pj[s]]; G.gcon[s] = gi[pi[s]][pj[s]]; }
                G.g = 1.;
                for (int zl = NG; zl < NG + B.bs[2]; zl++) {
                    const int64_t c = cell(B, il, jl, zl);
                    double pr[NPR];
                    for (int k = 0; k < NPR; k++) pr[k] = P(B, k)[c];
                    State<double> q;
                    get_state(pr, G, q);
                    double bsq = dot4(q.bcon, q.bcov);
                    if (bsq > 0.) beta_min = std::min(beta_min, (d.gam - 1.) * pr[UU] / (0.5 * bsq));
                }
            }
    }
    if (beta_min < 1e299 && t.beta_min > 0.) {
        const double norm = std::sqrt(beta_min / t.beta_min);
        for (int b = 0; b < nblk; b++)
            for (int k = 0; k < 3; k++) {
                double *f = PS(blocks[b], k);
                for (int64_t c = 0; c < blocks[b].ks; c++) f[c] *= norm;
            }
        for (int b = 0; b < nblk; b++) centre_field(blocks[b]);
    }
    return "";
}

}  // namespace hamr
