// hamr_capi.hip -- the C ABI of include/hamr_hip.h: context, resident block
// set, metric and state staging, per-block stages and the batched step.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hamr_hip.h"
#include "hamr_grid.h"
#include "hamr_kernels.h"

using namespace hamr;

struct hamr_ctx {
    GridDesc d;
    Layout L;
    std::vector<BlockDev> blocks;
    std::vector<int> order;
    std::string err;
    hipStream_t s = nullptr;
    int device = 0;
    KParams P;
    // device buffers
    double *p = nullptr, *ph = nullptr, *ps = nullptr, *psh = nullptr;
    double *F[3] = {nullptr, nullptr, nullptr};
    double *E = nullptr, *Ec = nullptr, *U0 = nullptr, *Un = nullptr;  // Un: updated conserved variables (scratch)
    int *pflag = nullptr, *failcnt = nullptr;
    unsigned long long *dtmin = nullptr;
    double *dt = nullptr;          // [0] dt of the next step, [1] previous
    double *gcov = nullptr, *gcon = nullptr, *gdet = nullptr, *conn = nullptr, *rcen = nullptr;
    BlockDev *blk = nullptr;
    int *ident = nullptr;          // 0..nblk-1 (single-block lists are offsets into it)
    int *zlist = nullptr;          // all blocks in Z-order
    int *ring = nullptr;           // [side][ring][c3] block ids of the pole rings
    int nring = 0;
    hipGraphExec_t graph = nullptr;
    int prec = HAMR_PREC_FP64;     // arithmetic of reconstruction, fluxes and EMFs
    // adjoint scratch (allocated by the first adjoint call)
    double *aw = nullptr, *alP = nullptr, *alps = nullptr, *alpsf = nullptr, *alE = nullptr, *alEc = nullptr;
    double *alF[3] = {nullptr, nullptr, nullptr};
    std::vector<void *> allocs;
};

static int fail(hamr_ctx *c, const std::string &m)
{
    if (c) c->err = m;
    return -1;
}

#define HIPCHK(c, x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail((c), std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int check_launch(hamr_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

template <class T> static int dalloc(hamr_ctx *c, T **ptr, size_t n)
{
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(T) + 256);
    if (e != hipSuccess) return fail(c, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemset(v, 0, n * sizeof(T) + 256);
    if (e != hipSuccess) return fail(c, std::string("hipMemset: ") + hipGetErrorString(e));
    c->allocs.push_back(v);
    *ptr = (T *)v;
    return 0;
}

static bool valid_block(const hamr_ctx *c, int n) { return n >= 0 && n < c->L.nblk; }

// Host<->device copies ordered on the context stream and completed before return.
static int h2d(hamr_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}
static int d2h(hamr_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" {

const char *hamr_last_error(const hamr_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hamr_create(const hamr_desc *desc, hamr_ctx **out)
{
    if (!desc || !out) return -1;
    *out = nullptr;
    hamr_ctx *c = new hamr_ctx();
    GridDesc &d = c->d;
    for (int k = 0; k < 3; k++) {
        d.bs[k] = desc->bs[k];
        d.nb[k] = desc->nb[k];
    }
    d.gam = desc->gam;
    d.a = desc->spin;
    d.cour = desc->cour;
    d.rin = desc->rin;
    d.rout = desc->rout;
    d.rho_floor = desc->rho_floor;
    d.u_floor = desc->u_floor;
    d.minkowski = desc->minkowski;
    c->device = desc->device;
    *out = c;
    for (int k = 0; k < 3; k++)
        if (d.bs[k] < 2 * NG || d.nb[k] < 1) return fail(c, "hamr_create: each block needs >= 6 cells and >= 1 block per direction");
    if (!d.minkowski) {
        if (d.nb[2] > 1 && (d.nb[2] % 2)) return fail(c, "hamr_create: the polar axis needs an even number of phi-blocks (or one)");
        if (d.nb[2] == 1 && (d.bs[2] % 2)) return fail(c, "hamr_create: the polar axis needs an even number of phi cells");
        if (!(d.rout > 0.) || !(d.gam > 1.)) return fail(c, "hamr_create: need rout > 0 and gam > 1");
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    c->L = make_layout(d);
    int ngeo = 0;
    c->blocks = build_block_table(d, &ngeo);
    c->order = zorder(d);
    const Layout &L = c->L;
    const size_t K = (size_t)L.ksize, nb = (size_t)L.nblk;
    if (dalloc(c, &c->p, nb * NPR * K) || dalloc(c, &c->ph, nb * NPR * K) || dalloc(c, &c->ps, nb * 3 * K) ||
        dalloc(c, &c->psh, nb * 3 * K) || dalloc(c, &c->F[0], nb * NPR * K) || dalloc(c, &c->F[1], nb * NPR * K) ||
        dalloc(c, &c->F[2], nb * NPR * K) || dalloc(c, &c->E, nb * 3 * K) || dalloc(c, &c->Ec, nb * 3 * K) ||
        dalloc(c, &c->U0, nb * NPR * K) || dalloc(c, &c->Un, nb * NPR * K) || dalloc(c, &c->pflag, nb * K) || dalloc(c, &c->failcnt, nb * 4) ||
        dalloc(c, &c->dtmin, nb * 3) || dalloc(c, &c->dt, 4))
        return -1;
    const size_t KC = (size_t)L.kcol, ng = (size_t)L.ngeo;
    if (dalloc(c, &c->gcov, ng * NPOS * 10 * KC) || dalloc(c, &c->gcon, ng * NPOS * 10 * KC) ||
        dalloc(c, &c->gdet, ng * NPOS * KC) || dalloc(c, &c->conn, ng * 40 * KC) || dalloc(c, &c->rcen, ng * KC) ||
        dalloc(c, &c->blk, nb) || dalloc(c, &c->ident, nb) || dalloc(c, &c->zlist, nb))
        return -1;
    if (h2d(c, c->blk, c->blocks.data(), nb * sizeof(BlockDev))) return -1;
    std::vector<int> id(nb);
    for (size_t n = 0; n < nb; n++) id[n] = (int)n;
    if (h2d(c, c->ident, id.data(), nb * sizeof(int))) return -1;
    if (h2d(c, c->zlist, c->order.data(), nb * sizeof(int))) return -1;
    // pole rings: for each c1, the blocks of c2 = 0 (side 0) and c2 = NB2-1 (side 1), in c3 order
    if (!d.minkowski) {
        c->nring = d.nb[0];
        std::vector<int> ring((size_t)2 * c->nring * d.nb[2]);
        for (int side = 0; side < 2; side++)
            for (int c1 = 0; c1 < d.nb[0]; c1++)
                for (int c3 = 0; c3 < d.nb[2]; c3++) {
                    int c2 = side ? d.nb[1] - 1 : 0;
                    ring[((size_t)side * c->nring + c1) * d.nb[2] + c3] = (c1 * d.nb[1] + c2) * d.nb[2] + c3;
                }
        if (dalloc(c, &c->ring, ring.size())) return -1;
        if (h2d(c, c->ring, ring.data(), ring.size() * sizeof(int))) return -1;
    }
    const GridCoords gc = grid_coords(d);
    KParams &P = c->P;
    for (int k = 0; k < 3; k++) {
        P.D[k] = L.D[k];
        P.bs[k] = d.bs[k];
        P.dx[k] = gc.dx[k];
    }
    P.ksize = L.ksize;
    P.kcol = L.kcol;
    P.nb3 = d.nb[2];
    P.gam = d.gam;
    P.cour = d.cour;
    P.rho_floor = d.rho_floor;
    P.u_floor = d.u_floor;
    P.blk = c->blk;
    P.gcov = c->gcov;
    P.gcon = c->gcon;
    P.gdet = c->gdet;
    P.conn = c->conn;
    P.rcen = c->rcen;
    launch_dt_reset(L.nblk, c->dtmin, c->s);
    if (check_launch(c, "dt reset")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    c->err.clear();
    return 0;
}

int hamr_destroy(hamr_ctx *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (void *v : c->allocs) (void)hipFree(v);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return 0;
}

int hamr_layout(const hamr_ctx *c, int *D, long long *ncell, long long *ksize, int *nblk, int *ngeo)
{
    if (!c) return -1;
    if (D)
        for (int k = 0; k < 3; k++) D[k] = c->L.D[k];
    if (ncell) *ncell = c->L.ncell;
    if (ksize) *ksize = c->L.ksize;
    if (nblk) *nblk = c->L.nblk;
    if (ngeo) *ngeo = c->L.ngeo;
    return 0;
}

int hamr_block_row(const hamr_ctx *c, int n, int *row)
{
    if (!c || !row || !valid_block(c, n)) return -1;
    const BlockDev &b = c->blocks[n];
    row[0] = b.geo;
    row[1] = b.c1;
    row[2] = b.c2;
    row[3] = b.c3;
    row[4] = b.pole;
    for (int k = 0; k < 6; k++) {
        row[5 + k] = b.nbr[k];
        row[11 + k] = b.kind[k];
    }
    return 0;
}

int hamr_block_order(const hamr_ctx *c, int *order)
{
    if (!c || !order) return -1;
    std::memcpy(order, c->order.data(), c->order.size() * sizeof(int));
    return 0;
}


// ---- host-only set-up (no device needed) ----
static int desc_to_grid(const hamr_desc *desc, GridDesc &d)
{
    if (!desc) return -1;
    for (int k = 0; k < 3; k++) {
        d.bs[k] = desc->bs[k];
        d.nb[k] = desc->nb[k];
        if (d.bs[k] < 2 * NG || d.nb[k] < 1) return -1;
    }
    d.gam = desc->gam;
    d.a = desc->spin;
    d.cour = desc->cour;
    d.rin = desc->rin;
    d.rout = desc->rout;
    d.rho_floor = desc->rho_floor;
    d.u_floor = desc->u_floor;
    d.minkowski = desc->minkowski;
    return 0;
}

int hamr_desc_layout(const hamr_desc *desc, int *D, long long *ncell, long long *ksize, long long *ncol,
                     long long *kcol, int *nblk, int *ngeo)
{
    GridDesc d;
    if (desc_to_grid(desc, d)) return -1;
    Layout L = make_layout(d);
    if (D)
        for (int k = 0; k < 3; k++) D[k] = L.D[k];
    if (ncell) *ncell = L.ncell;
    if (ksize) *ksize = L.ksize;
    if (ncol) *ncol = L.ncol;
    if (kcol) *kcol = L.kcol;
    if (nblk) *nblk = L.nblk;
    if (ngeo) *ngeo = L.ngeo;
    return 0;
}

int hamr_desc_block_table(const hamr_desc *desc, int *rows, int *order)
{
    GridDesc d;
    if (desc_to_grid(desc, d)) return -1;
    std::vector<BlockDev> t = build_block_table(d, nullptr);
    for (size_t n = 0; n < t.size(); n++) {
        int *row = rows + n * 17;
        const BlockDev &b = t[n];
        row[0] = b.geo;
        row[1] = b.c1;
        row[2] = b.c2;
        row[3] = b.c3;
        row[4] = b.pole;
        for (int k = 0; k < 6; k++) {
            row[5 + k] = b.nbr[k];
            row[11 + k] = b.kind[k];
        }
    }
    if (order) {
        std::vector<int> z = zorder(d);
        std::memcpy(order, z.data(), z.size() * sizeof(int));
    }
    return 0;
}

int hamr_desc_coords(const hamr_desc *desc, double *startx, double *dx)
{
    GridDesc d;
    if (desc_to_grid(desc, d)) return -1;
    GridCoords gc = grid_coords(d);
    for (int k = 0; k < 3; k++) {
        if (startx) startx[k] = gc.startx[k];
        if (dx) dx[k] = gc.dx[k];
    }
    return 0;
}

int hamr_desc_metric_at(const hamr_desc *desc, const double *X, double *gcov, double *gcon, double *gdet,
                        double *conn, double *r_th)
{
    GridDesc d;
    if (desc_to_grid(desc, d) || !X) return -1;
    static const int pi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3}, pj[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    double g[4][4], gi[4][4], cn[4][4][4];
    metric_code(d, X, g);
    if (invert4(g, gi)) return -1;
    for (int s = 0; s < 10; s++) {
        if (gcov) gcov[s] = g[pi[s]][pj[s]];
        if (gcon) gcon[s] = gi[pi[s]][pj[s]];
    }
    if (gdet) *gdet = gdet_of(g);
    if (conn) {
        conn_code(d, X, gi, cn);
        for (int lam = 0; lam < 4; lam++)
            for (int s = 0; s < 10; s++) conn[lam * 10 + s] = cn[lam][pi[s]][pj[s]];
    }
    if (r_th) {
        if (d.minkowski) { r_th[0] = 1.; r_th[1] = 0.; }
        else bl_coord(d, X, &r_th[0], &r_th[1]);
    }
    return 0;
}

int hamr_desc_geometry(const hamr_desc *desc, int geo, double *gcov, double *gcon, double *gdet, double *conn,
                       double *rcen)
{
    GridDesc d;
    if (desc_to_grid(desc, d)) return -1;
    Layout L = make_layout(d);
    if (geo < 0 || geo >= L.ngeo) return -1;
    const size_t KC = (size_t)L.kcol;
    std::memset(gcov, 0, NPOS * 10 * KC * sizeof(double));
    std::memset(gcon, 0, NPOS * 10 * KC * sizeof(double));
    std::memset(gdet, 0, NPOS * KC * sizeof(double));
    std::memset(conn, 0, 40 * KC * sizeof(double));
    std::memset(rcen, 0, KC * sizeof(double));
    fill_geometry(d, L, geo / d.nb[1], geo % d.nb[1], gcov, gcon, gdet, conn, rcen);
    return 0;
}

int hamr_desc_torus(const hamr_desc *desc, const hamr_torus_desc *t, double *p, double *ps)
{
    GridDesc d;
    if (desc_to_grid(desc, d) || !t) return -1;
    Layout L = make_layout(d);
    const size_t KC = (size_t)L.kcol, ng = (size_t)L.ngeo, K = (size_t)L.ksize, N = (size_t)L.ncell;
    std::vector<double> gcov(ng * NPOS * 10 * KC, 0.), gcon(ng * NPOS * 10 * KC, 0.), gdet(ng * NPOS * KC, 0.);
    std::vector<double> conn(40 * KC), rcen(KC);
    for (size_t g = 0; g < ng; g++)
        fill_geometry(d, L, (int)g / d.nb[1], (int)g % d.nb[1], gcov.data() + g * NPOS * 10 * KC,
                      gcon.data() + g * NPOS * 10 * KC, gdet.data() + g * NPOS * KC, conn.data(), rcen.data());
    std::vector<BlockDev> blocks = build_block_table(d, nullptr);
    std::vector<double> tp((size_t)L.nblk * NPR * K, 0.), ts((size_t)L.nblk * 3 * K, 0.);
    TorusDesc td{t->rin, t->rmax, t->kappa, t->beta_min, t->aphi_cut, t->perturb, t->seed};
    std::string e = init_torus(d, L, blocks, td, gcov.data(), gcon.data(), gdet.data(), tp.data(), ts.data());
    if (!e.empty()) return -2;
    for (size_t s = 0; s < (size_t)L.nblk * NPR; s++) std::memcpy(p + s * N, tp.data() + s * K, N * sizeof(double));
    for (size_t s = 0; s < (size_t)L.nblk * 3; s++) std::memcpy(ps + s * N, ts.data() + s * K, N * sizeof(double));
    return 0;
}

int hamr_metric_sizes(const hamr_ctx *c, long long *ncol, long long *kcol)
{
    if (!c) return -1;
    if (ncol) *ncol = c->L.ncol;
    if (kcol) *kcol = c->L.kcol;
    return 0;
}

int hamr_set_grid(hamr_ctx *c)
{
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    const Layout &L = c->L;
    const size_t KC = (size_t)L.kcol;
    std::vector<double> gcov(NPOS * 10 * KC), gcon(NPOS * 10 * KC), gdet(NPOS * KC), conn(40 * KC), rcen(KC);
    for (int c1 = 0; c1 < c->d.nb[0]; c1++)
        for (int c2 = 0; c2 < c->d.nb[1]; c2++) {
            std::fill(gcov.begin(), gcov.end(), 0.);
            std::fill(gcon.begin(), gcon.end(), 0.);
            std::fill(gdet.begin(), gdet.end(), 0.);
            std::fill(conn.begin(), conn.end(), 0.);
            std::fill(rcen.begin(), rcen.end(), 0.);
            fill_geometry(c->d, L, c1, c2, gcov.data(), gcon.data(), gdet.data(), conn.data(), rcen.data());
            int r = hamr_upload_metric(c, c1 * c->d.nb[1] + c2, gcov.data(), gcon.data(), gdet.data(), conn.data(),
                                       rcen.data());
            if (r) return r;
        }
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_upload_metric(hamr_ctx *c, int geo, const double *gcov, const double *gcon, const double *gdet,
                       const double *conn, const double *rcen)
{
    if (!c || geo < 0 || geo >= c->L.ngeo) return fail(c, "hamr_upload_metric: bad geometry index");
    const size_t KC = (size_t)c->L.kcol;
    if (h2d(c, c->gcov + (size_t)geo * NPOS * 10 * KC, gcov, NPOS * 10 * KC * sizeof(double))) return -1;
    if (h2d(c, c->gcon + (size_t)geo * NPOS * 10 * KC, gcon, NPOS * 10 * KC * sizeof(double))) return -1;
    if (h2d(c, c->gdet + (size_t)geo * NPOS * KC, gdet, NPOS * KC * sizeof(double))) return -1;
    if (h2d(c, c->conn + (size_t)geo * 40 * KC, conn, 40 * KC * sizeof(double))) return -1;
    if (h2d(c, c->rcen + (size_t)geo * KC, rcen, KC * sizeof(double))) return -1;
    return 0;
}

int hamr_download_metric(hamr_ctx *c, int geo, double *gcov, double *gcon, double *gdet, double *conn, double *rcen)
{
    if (!c || geo < 0 || geo >= c->L.ngeo) return fail(c, "hamr_download_metric: bad geometry index");
    const size_t KC = (size_t)c->L.kcol;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, gcov, c->gcov + (size_t)geo * NPOS * 10 * KC, NPOS * 10 * KC * sizeof(double))) return -1;
    if (d2h(c, gcon, c->gcon + (size_t)geo * NPOS * 10 * KC, NPOS * 10 * KC * sizeof(double))) return -1;
    if (d2h(c, gdet, c->gdet + (size_t)geo * NPOS * KC, NPOS * KC * sizeof(double))) return -1;
    if (d2h(c, conn, c->conn + (size_t)geo * 40 * KC, 40 * KC * sizeof(double))) return -1;
    if (d2h(c, rcen, c->rcen + (size_t)geo * KC, KC * sizeof(double))) return -1;
    return 0;
}

int hamr_init_torus(hamr_ctx *c, const hamr_torus_desc *t)
{
    if (!c || !t) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    const Layout &L = c->L;
    const size_t K = (size_t)L.ksize, nb = (size_t)L.nblk, KC = (size_t)L.kcol, ng = (size_t)L.ngeo;
    std::vector<double> gcov(ng * NPOS * 10 * KC), gcon(ng * NPOS * 10 * KC), gdet(ng * NPOS * KC);
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, gcov.data(), c->gcov, gcov.size() * sizeof(double))) return -1;
    if (d2h(c, gcon.data(), c->gcon, gcon.size() * sizeof(double))) return -1;
    if (d2h(c, gdet.data(), c->gdet, gdet.size() * sizeof(double))) return -1;
    std::vector<double> p(nb * NPR * K, 0.), ps(nb * 3 * K, 0.);
    TorusDesc td{t->rin, t->rmax, t->kappa, t->beta_min, t->aphi_cut, t->perturb, t->seed};
    std::string e = init_torus(c->d, L, c->blocks, td, gcov.data(), gcon.data(), gdet.data(), p.data(), ps.data());
    if (!e.empty()) return fail(c, e);
    if (h2d(c, c->p, p.data(), p.size() * sizeof(double))) return -1;
    if (h2d(c, c->ps, ps.data(), ps.size() * sizeof(double))) return -1;
    if (h2d(c, c->ph, p.data(), p.size() * sizeof(double))) return -1;
    if (h2d(c, c->psh, ps.data(), ps.size() * sizeof(double))) return -1;
    return 0;
}

static int array_info(hamr_ctx *c, int which, double **ptr, int *ncomp)
{
    switch (which) {
    case HAMR_P: *ptr = c->p; *ncomp = NPR; return 0;
    case HAMR_PS: *ptr = c->ps; *ncomp = 3; return 0;
    case HAMR_PH: *ptr = c->ph; *ncomp = NPR; return 0;
    case HAMR_PSH: *ptr = c->psh; *ncomp = 3; return 0;
    case HAMR_F1: *ptr = c->F[0]; *ncomp = NPR; return 0;
    case HAMR_F2: *ptr = c->F[1]; *ncomp = NPR; return 0;
    case HAMR_F3: *ptr = c->F[2]; *ncomp = NPR; return 0;
    case HAMR_EMF: *ptr = c->E; *ncomp = 3; return 0;
    case HAMR_ECENT: *ptr = c->Ec; *ncomp = 3; return 0;
    case HAMR_U0: *ptr = c->U0; *ncomp = NPR; return 0;
    case HAMR_UNEW: *ptr = c->Un; *ncomp = NPR; return 0;
    default: return fail(c, "unknown array selector");
    }
}

// Fluxes and EMFs are float arrays in the mixed step.
static bool is_float_array(const hamr_ctx *c, int which)
{
    return c->prec == HAMR_PREC_MIXED &&
           (which == HAMR_F1 || which == HAMR_F2 || which == HAMR_F3 || which == HAMR_EMF || which == HAMR_ECENT);
}

int hamr_set_array(hamr_ctx *c, int which, const double *src)
{
    if (!c || !src) return -1;
    double *d;
    int nc;
    if (array_info(c, which, &d, &nc)) return -1;
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell, rows = (size_t)c->L.nblk * nc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (is_float_array(c, which)) {
        std::vector<float> t(rows * N);
        for (size_t e = 0; e < t.size(); e++) t[e] = (float)src[e];
        HIPCHK(c, hipMemcpy2DAsync(d, K * sizeof(float), t.data(), N * sizeof(float), N * sizeof(float), rows,
                                   hipMemcpyHostToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        return 0;
    }
    HIPCHK(c, hipMemcpy2DAsync(d, K * sizeof(double), src, N * sizeof(double), N * sizeof(double), rows,
                               hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_get_array(hamr_ctx *c, int which, double *dst)
{
    if (!c || !dst) return -1;
    double *d;
    int nc;
    if (array_info(c, which, &d, &nc)) return -1;
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell, rows = (size_t)c->L.nblk * nc;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (is_float_array(c, which)) {
        std::vector<float> t(rows * N);
        HIPCHK(c, hipMemcpy2DAsync(t.data(), N * sizeof(float), d, K * sizeof(float), N * sizeof(float), rows,
                                   hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        for (size_t e = 0; e < t.size(); e++) dst[e] = (double)t[e];
        return 0;
    }
    HIPCHK(c, hipMemcpy2DAsync(dst, N * sizeof(double), d, K * sizeof(double), N * sizeof(double), rows,
                               hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_set_array_dev(hamr_ctx *c, int which, const double *dsrc)
{
    if (!c || !dsrc) return -1;
    double *d;
    int nc;
    if (array_info(c, which, &d, &nc)) return -1;
    if (is_float_array(c, which)) return fail(c, "hamr_set_array_dev: float array in the mixed step");
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell, rows = (size_t)c->L.nblk * nc;
    HIPCHK(c, hipMemcpy2DAsync(d, K * sizeof(double), dsrc, N * sizeof(double), N * sizeof(double), rows,
                               hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_get_array_dev(hamr_ctx *c, int which, double *ddst)
{
    if (!c || !ddst) return -1;
    double *d;
    int nc;
    if (array_info(c, which, &d, &nc)) return -1;
    if (is_float_array(c, which)) return fail(c, "hamr_get_array_dev: float array in the mixed step");
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell, rows = (size_t)c->L.nblk * nc;
    HIPCHK(c, hipMemcpy2DAsync(ddst, N * sizeof(double), d, K * sizeof(double), N * sizeof(double), rows,
                               hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_set_precision(hamr_ctx *c, int mode)
{
    if (!c) return -1;
    if (mode != HAMR_PREC_FP64 && mode != HAMR_PREC_MIXED) return fail(c, "hamr_set_precision: unknown mode");
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (mode != c->prec && c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    c->prec = mode;
    return 0;
}

int hamr_get_precision(const hamr_ctx *c, int *mode)
{
    if (!c || !mode) return -1;
    *mode = c->prec;
    return 0;
}

int hamr_upload(hamr_ctx *c, int n, const double *p_aos, const double *ps_aos)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_upload: bad block");
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell;
    std::vector<double> tp(NPR * K, 0.), ts(3 * K, 0.);
    for (size_t cell = 0; cell < N; cell++) {
        for (int k = 0; k < NPR; k++) tp[k * K + cell] = p_aos[cell * NPR + k];
        if (ps_aos)
            for (int k = 0; k < 3; k++) ts[k * K + cell] = ps_aos[cell * 4 + 1 + k];
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (h2d(c, c->p + (size_t)n * NPR * K, tp.data(), NPR * K * sizeof(double))) return -1;
    if (ps_aos) if (h2d(c, c->ps + (size_t)n * 3 * K, ts.data(), 3 * K * sizeof(double))) return -1;
    return 0;
}

int hamr_download(hamr_ctx *c, int n, double *p_aos, double *ps_aos)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_download: bad block");
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell;
    std::vector<double> tp(NPR * K), ts(3 * K);
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, tp.data(), c->p + (size_t)n * NPR * K, NPR * K * sizeof(double))) return -1;
    if (d2h(c, ts.data(), c->ps + (size_t)n * 3 * K, 3 * K * sizeof(double))) return -1;
    for (size_t cell = 0; cell < N; cell++) {
        if (p_aos)
            for (int k = 0; k < NPR; k++) p_aos[cell * NPR + k] = tp[k * K + cell];
        if (ps_aos)
            for (int k = 0; k < 3; k++) ps_aos[cell * 4 + 1 + k] = ts[k * K + cell];
    }
    return 0;
}

// ---- per-block stages ----
int hamr_fluxcalc(hamr_ctx *c, int dir, int flag, int n)
{
    if (!c || !valid_block(c, n) || dir < 1 || dir > 3) return fail(c, "hamr_fluxcalc: bad arguments");
    launch_flux(c->P, dir, c->ident + n, 1, flag ? c->ph : c->p, flag ? c->psh : c->ps, c->F[dir - 1], c->dtmin, c->prec, c->s);
    return check_launch(c, "hamr_fluxcalc");
}

int hamr_block_dt(hamr_ctx *c, int n, int dir, double *out)
{
    if (!c || !valid_block(c, n) || dir < 1 || dir > 3 || !out) return fail(c, "hamr_block_dt: bad arguments");
    unsigned long long v;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, &v, c->dtmin + n * 3 + dir - 1, sizeof v)) return -1;
    double d;
    std::memcpy(&d, &v, sizeof d);
    *out = d;
    const unsigned long long init = 0x41CDCD6500000000ULL;
    if (h2d(c, c->dtmin + n * 3 + dir - 1, &init, sizeof init)) return -1;
    return 0;
}

int hamr_ct_emf(hamr_ctx *c, int flag, int n)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_ct_emf: bad block");
    launch_emf(c->P, c->ident + n, 1, flag ? c->ph : c->p, c->F[0], c->F[1], c->F[2], c->Ec, c->E, c->prec, c->s);
    return check_launch(c, "hamr_ct_emf");
}

int hamr_ct_bound(hamr_ctx *c, double dt)
{
    if (!c) return -1;
    launch_pole(c->P, c->ring, c->nring, dt, nullptr, c->E, c->prec, c->s);
    return check_launch(c, "hamr_ct_bound");
}

int hamr_ct_update(hamr_ctx *c, int flag, double Dt, int n)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_ct_update: bad block");
    launch_ct_update(c->P, c->ident + n, 1, Dt, nullptr, 1.0, c->ps, flag ? c->ps : c->psh, c->E, c->prec, c->s);
    return check_launch(c, "hamr_ct_update");
}

int hamr_update(hamr_ctx *c, int flag, int n, double Dt)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_update: bad block");
    if (flag)
        launch_update(c->P, c->ident + n, 1, 1, Dt, nullptr, 1.0, c->ph, c->ph, c->p, c->U0, c->Un, c->ps, c->F[0], c->F[1],
                      c->F[2], c->pflag, c->failcnt, c->prec, c->s);
    else
        launch_update(c->P, c->ident + n, 1, 0, Dt, nullptr, 1.0, c->p, c->p, c->ph, c->U0, c->Un, c->psh, c->F[0], c->F[1],
                      c->F[2], c->pflag, c->failcnt, c->prec, c->s);
    return check_launch(c, "hamr_update");
}

static void bound_all(hamr_ctx *c, double *p, double *ps)
{
    for (int dir = 1; dir <= 3; dir++) launch_bound(c->P, dir, c->zlist, c->L.nblk, p, ps, c->pflag, c->s);
}

int hamr_boundprim(hamr_ctx *c, int flag)
{
    if (!c) return -1;
    if (flag) bound_all(c, c->p, c->ps);
    else bound_all(c, c->ph, c->psh);
    return check_launch(c, "hamr_boundprim");
}

int hamr_flags(hamr_ctx *c, int n, int *pflag, int *failimage)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_flags: bad block");
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (pflag) if (d2h(c, pflag, c->pflag + (size_t)n * K, N * sizeof(int))) return -1;
    if (failimage) if (d2h(c, failimage, c->failcnt + (size_t)n * 4, 4 * sizeof(int))) return -1;
    return 0;
}

int hamr_clear_flags(hamr_ctx *c)
{
    if (!c) return -1;
    HIPCHK(c, hipMemsetAsync(c->failcnt, 0, (size_t)c->L.nblk * 4 * sizeof(int), c->s));
    return 0;
}

// ---- batched step ----
static void enqueue_step(hamr_ctx *c)
{
    const KParams &P = c->P;
    const int nb = c->L.nblk;
    const int *bl = c->zlist;
    hipStream_t s = c->s;
    // predictor: p/ps -> ph/psh with Dt = dt/2
    for (int dir = 1; dir <= 3; dir++) launch_flux(P, dir, bl, nb, c->p, c->ps, c->F[dir - 1], nullptr, c->prec, s);
    launch_emf(P, bl, nb, c->p, c->F[0], c->F[1], c->F[2], c->Ec, c->E, c->prec, s);
    launch_pole(P, c->ring, c->nring, 0., c->dt, c->E, c->prec, s);
    launch_ct_update(P, bl, nb, 0., c->dt, 0.5, c->ps, c->psh, c->E, c->prec, s);
    launch_update(P, bl, nb, 0, 0., c->dt, 0.5, c->p, c->p, c->ph, c->U0, c->Un, c->psh, c->F[0], c->F[1], c->F[2], c->pflag,
                  c->failcnt, c->prec, s);
    bound_all(c, c->ph, c->psh);
    // corrector: U0 -> p/ps with Dt = dt, fluxes from ph/psh, Courant minima recorded
    for (int dir = 1; dir <= 3; dir++) launch_flux(P, dir, bl, nb, c->ph, c->psh, c->F[dir - 1], c->dtmin, c->prec, s);
    launch_emf(P, bl, nb, c->ph, c->F[0], c->F[1], c->F[2], c->Ec, c->E, c->prec, s);
    launch_pole(P, c->ring, c->nring, 0., c->dt, c->E, c->prec, s);
    launch_ct_update(P, bl, nb, 0., c->dt, 1.0, c->ps, c->ps, c->E, c->prec, s);
    launch_update(P, bl, nb, 1, 0., c->dt, 1.0, c->ph, c->ph, c->p, c->U0, c->Un, c->ps, c->F[0], c->F[1], c->F[2], c->pflag,
                  c->failcnt, c->prec, s);
    bound_all(c, c->p, c->ps);
    launch_dt_finalize(nb, c->dtmin, c->dt, s);
}

int hamr_estimate_dt(hamr_ctx *c, double *dt)
{
    if (!c || !dt) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    const double big[2] = {1e300, 1e300};
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (h2d(c, c->dt, big, sizeof big)) return -1;
    launch_dt_reset(c->L.nblk, c->dtmin, c->s);
    for (int dir = 1; dir <= 3; dir++) launch_flux(c->P, dir, c->zlist, c->L.nblk, c->p, c->ps, c->F[dir - 1], c->dtmin, c->prec, c->s);
    launch_dt_finalize(c->L.nblk, c->dtmin, c->dt, c->s);
    if (check_launch(c, "hamr_estimate_dt")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, dt, c->dt, sizeof(double))) return -1;
    return 0;
}

static int set_dt(hamr_ctx *c, const double *dt_inout)
{
    if (dt_inout && *dt_inout > 0.) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (h2d(c, c->dt, dt_inout, sizeof(double))) return -1;
    }
    return 0;
}

int hamr_step(hamr_ctx *c, int nsteps, double *dt_inout)
{
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    if (set_dt(c, dt_inout)) return -1;
    launch_dt_reset(c->L.nblk, c->dtmin, c->s);
    for (int k = 0; k < nsteps; k++) enqueue_step(c);
    if (check_launch(c, "hamr_step")) return -1;
    if (dt_inout) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (d2h(c, dt_inout, c->dt, sizeof(double))) return -1;
    }
    return 0;
}

int hamr_step_graph(hamr_ctx *c, int nsteps, double *dt_inout)
{
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    if (set_dt(c, dt_inout)) return -1;
    launch_dt_reset(c->L.nblk, c->dtmin, c->s);
    if (!c->graph) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        enqueue_step(c);
        HIPCHK(c, hipStreamEndCapture(c->s, &g));
        HIPCHK(c, hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
    }
    for (int k = 0; k < nsteps; k++) HIPCHK(c, hipGraphLaunch(c->graph, c->s));
    if (dt_inout) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (d2h(c, dt_inout, c->dt, sizeof(double))) return -1;
    }
    return 0;
}


int hamr_time_stage(hamr_ctx *c, int stage, int nrep, double *avg_ms)
{
    if (!c || !avg_ms || nrep < 1 || stage < 0 || stage > 7) return fail(c, "hamr_time_stage: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const KParams &P = c->P;
    const int nb = c->L.nblk;
    const int *bl = c->zlist;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->s));
    for (int r = 0; r < nrep; r++) {
        switch (stage) {
        case 0:
        case 1:
        case 2: launch_flux(P, stage + 1, bl, nb, c->p, c->ps, c->F[stage], nullptr, c->prec, c->s); break;
        case 3: launch_emf(P, bl, nb, c->p, c->F[0], c->F[1], c->F[2], c->Ec, c->E, c->prec, c->s); break;
        case 4: launch_pole(P, c->ring, c->nring, 0., c->dt, c->E, c->prec, c->s); break;
        case 5: launch_ct_update(P, bl, nb, 0., c->dt, 0.5, c->ps, c->psh, c->E, c->prec, c->s); break;
        case 6:
            launch_update(P, bl, nb, 0, 0., c->dt, 0.5, c->p, c->p, c->ph, c->U0, c->Un, c->psh, c->F[0], c->F[1], c->F[2],
                          c->pflag, c->failcnt, c->prec, c->s);
            break;
        default: bound_all(c, c->ph, c->psh); break;
        }
    }
    HIPCHK(c, hipEventRecord(e1, c->s));
    if (check_launch(c, "hamr_time_stage")) return -1;
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / nrep;
    return 0;
}

// ---- one forward stage and its adjoint ----
static void enqueue_stage_forward(hamr_ctx *c, double Dt)
{
    const KParams &P = c->P;
    const int nb = c->L.nblk;
    const int *bl = c->zlist;
    bound_all(c, c->p, c->ps);
    for (int dir = 1; dir <= 3; dir++) launch_flux(P, dir, bl, nb, c->p, c->ps, c->F[dir - 1], nullptr, c->prec, c->s);
    launch_emf(P, bl, nb, c->p, c->F[0], c->F[1], c->F[2], c->Ec, c->E, c->prec, c->s);
    launch_pole(P, c->ring, c->nring, Dt, nullptr, c->E, c->prec, c->s);
    launch_ct_update(P, bl, nb, Dt, nullptr, 1.0, c->ps, c->psh, c->E, c->prec, c->s);
    launch_cons(P, bl, nb, Dt, c->p, c->U0, c->Un, c->psh, c->F[0], c->F[1], c->F[2], c->prec, c->s);
}

int hamr_stage_forward(hamr_ctx *c, double Dt)
{
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    enqueue_stage_forward(c, Dt);
    return check_launch(c, "hamr_stage_forward");
}

// w, grad_p, grad_ps: pointers of `kind` (host or device), unpadded layout.
static int stage_adjoint(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps, hipMemcpyKind in,
                         hipMemcpyKind out)
{
    if (!c || !w) return -1;
    if (!c->d.minkowski || c->L.nblk != 1 || c->prec != HAMR_PREC_FP64)
        return fail(c, "hamr_stage_adjoint: needs a flat periodic box of one block in FP64");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t K = (size_t)c->L.ksize, N = (size_t)c->L.ncell;
    if (!c->aw) {
        if (dalloc(c, &c->aw, NPR * K) || dalloc(c, &c->alP, NPR * K) || dalloc(c, &c->alps, 3 * K) ||
            dalloc(c, &c->alpsf, 3 * K) || dalloc(c, &c->alE, 3 * K) || dalloc(c, &c->alEc, 3 * K) ||
            dalloc(c, &c->alF[0], NPR * K) || dalloc(c, &c->alF[1], NPR * K) || dalloc(c, &c->alF[2], NPR * K))
            return -1;
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    HIPCHK(c, hipMemcpy2DAsync(c->aw, K * sizeof(double), w, N * sizeof(double), N * sizeof(double), NPR, in, c->s));
    // the fluxes of the stage (their mass-flux signs select the EMF upwinding)
    bound_all(c, c->p, c->ps);
    for (int dir = 1; dir <= 3; dir++)
        launch_flux(c->P, dir, c->zlist, 1, c->p, c->ps, c->F[dir - 1], nullptr, HAMR_PREC_FP64, c->s);
    launch_adjoint(c->P, c->blocks[0], Dt, c->p, c->ps, c->F[0], c->F[1], c->F[2], c->aw, c->alP, c->alps, c->alpsf,
                   c->alE, c->alEc, c->alF[0], c->alF[1], c->alF[2], c->s);
    if (check_launch(c, "hamr_stage_adjoint")) return -1;
    if (gp)
        HIPCHK(c, hipMemcpy2DAsync(gp, N * sizeof(double), c->alP, K * sizeof(double), N * sizeof(double), NPR, out, c->s));
    if (gps)
        HIPCHK(c, hipMemcpy2DAsync(gps, N * sizeof(double), c->alps, K * sizeof(double), N * sizeof(double), 3, out, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_stage_adjoint(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps)
{
    return stage_adjoint(c, Dt, w, gp, gps, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}

int hamr_stage_adjoint_dev(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps)
{
    return stage_adjoint(c, Dt, w, gp, gps, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice);
}

int hamr_copy_probe(hamr_ctx *c, long long bytes, int nrep, double *avg_ms)
{
    if (!c || bytes < 16 || nrep < 1) return fail(c, "hamr_copy_probe: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, (size_t)bytes));
    HIPCHK(c, hipMalloc(&b, (size_t)bytes));
    HIPCHK(c, hipMemsetAsync(a, 0, (size_t)bytes, c->s));
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->s));
    for (int r = 0; r < nrep; r++) launch_copy(a, b, bytes, c->s);
    HIPCHK(c, hipEventRecord(e1, c->s));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (avg_ms) *avg_ms = (double)ms / nrep;
    return check_launch(c, "hamr_copy_probe");
}

int hamr_sync(hamr_ctx *c)
{
    if (!c) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return check_launch(c, "hamr_sync");
}

}  // extern "C"
