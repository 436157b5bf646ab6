// hamr_capi.hip -- the C ABI of include/hamr_hip.h: context, resident block
// set, metric and state staging, per-block stages, the batched step, the
// multi-level block set and the launch autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/hamr_hip.h"
#include "hamr_amr.h"
#include "hamr_grid.h"
#include "hamr_kernels.h"

using namespace hamr;

namespace {

constexpr int kTiles[3] = {64, 128, 256};   // tile sizes (threads per workgroup) the work lists exist for
constexpr int kDefaultTile = 2;
static_assert(kTiles[kDefaultTile] == kLineT, "line-staged flux tiles use the default tile size");

// Kernels the autotuner configures (each walks the work list of one class).
enum TuneSlot : int {
    TS_FLUX1 = 0, TS_FLUX2, TS_FLUX3, TS_ECENT, TS_ECORNER, TS_CT, TS_CONS, TS_INVERT, TS_FALLBACK,
    TS_BOUND1, TS_BOUND2, TS_BOUND3, TS_EXCHANGE, TS_N
};
constexpr int kSlotClass[TS_N] = {KC_FLUX1, KC_FLUX2, KC_FLUX3, KC_ECENT, KC_CORNER, KC_CORNER, KC_INTERIOR,
                                   KC_INTERIOR, KC_INTERIOR, KC_BOUND1, KC_BOUND2, KC_BOUND3, -1};

struct TuneCfg {
    int tile = kDefaultTile;  // index into kTiles
    int grid = 0;             // workgroups; 0: one per item
    int swz = 0;              // XCD-contiguous item order
    int march = 0;            // TS_CONS only: the x1-marching form (tile, grid, swz unused)
    int alt = 0;              // the slot's second form: flux slots line-staged tiles (KC_LINE lists)
                              // instead of one thread per face; TS_CONS k_cons + k_invert instead of k_update
    float ms = -1.f;          // measured time of the chosen configuration
};

// Work lists of one kernel class over its block list, for each tile size.
struct ClassList {
    int2 *items[3] = {nullptr, nullptr, nullptr};
    int n[3] = {0, 0, 0};
    int lds_cols[3] = {0, 0, 0};
    std::vector<int> start[3], count[3];  // per block id (count 0: not in the list)
};

}  // namespace

struct hamr_ctx {
    GridDesc d;
    Layout L;                      // uniform grid: the common block layout
    bool amr = false;
    std::vector<BlockDev> blocks;
    std::vector<AmrGeo> geos;      // geometry slabs (level, first column, extents, offset)
    std::vector<int> order;        // launch order
    int64_t cells = 0, cols = 0;   // sum of slab lengths / metric slab widths
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
    double *tet = nullptr;         // face tetrads of the HLLC solver (96 per metric column)
    float *g32 = nullptr;          // the metric in FP32 (KParams::g32)
    BlockDev *blk = nullptr;
    int *ring = nullptr, *rside = nullptr;  // polar rings
    int nring = 0, ring_len = 0, ni_max = 0;
    ClassList cl[KC_N];
    TuneCfg tune[TS_N];
    // multi-level exchange tables per stage
    AmrEntry *ex[3] = {nullptr, nullptr, nullptr};
    AmrProlong *pro[3] = {nullptr, nullptr, nullptr};
    int64_t nex[3] = {0, 0, 0}, npro[3] = {0, 0, 0};
    int64_t npush3 = 0;            // stage 3's last entries: the rows that the update writes
    std::vector<int> push3tab;     // push table (host copy: receivers)
    PushEntry *push3 = nullptr;
    AmrReflux *rfx[3] = {nullptr, nullptr, nullptr};  // flux correction at level jumps, per direction
    int64_t nrfx[3] = {0, 0, 0};
    AmrEmf *emfx = nullptr;        // EMF correction at level jumps
    int64_t nemfx = 0;
    hipGraphExec_t graph = nullptr;
    hipGraphExec_t las_graph = nullptr;  // one time-level cycle (hamr_step_las_graph)
    int las_graph_maxtl = 0;
    int prec = HAMR_PREC_FP64;     // arithmetic of reconstruction, fluxes and EMFs
    // adjoint scratch (allocated by the first adjoint call)
    double *aw = nullptr, *alP = nullptr, *alps = nullptr, *alpsf = nullptr, *alE = nullptr, *alEc = nullptr;
    double *alF[3] = {nullptr, nullptr, nullptr};
    std::vector<void *> allocs;
    void *stage = nullptr;         // pinned host staging of whole-array copies
    size_t stage_bytes = 0;
    // restart records: record n of the whole-grid layout starts at roff[n] doubles
    std::vector<int64_t> roff;
    int *io_blist = nullptr;       // 0 .. nblk-1
    int64_t *io_roff = nullptr;
    std::vector<int64_t> doff;     // dump record n starts at doff[n] floats
    int64_t *io_doff = nullptr;
    double *io = nullptr;          // device record of one block (grown on demand)
    size_t io_bytes = 0;
    double *scratch = nullptr;     // this device's address of the file-scope `scratch` array
    // local adaptive time stepping (multi-level sets)
    AmrGrid grid;                  // the set's blocks, rings, levels and correction tables (no exchange tables)
    std::vector<int> tl;           // time level per block (empty: not set)
    int tlmax = 1;                 // largest time level
    double *bdt = nullptr;         // per-block Courant dt of the last corrector that recorded them
    std::vector<void *> las_allocs;
    int2 *las_items = nullptr;     // filtered work lists of every (sub-iteration, phase) set of blocks
    std::vector<int64_t> las_off;  // [(s * 2 + phase - 1) * KC_N + cls] -> {start, count} pairs
    LasCorr *las_f[3] = {nullptr, nullptr, nullptr}, *las_es = nullptr, *las_ej = nullptr;
    int64_t nlas_f[3] = {0, 0, 0}, nlas_es = 0, nlas_ej = 0;
    double *acc_f[3] = {nullptr, nullptr, nullptr}, *acc_es = nullptr, *acc_ej = nullptr;
    // the exchange entries between blocks that both step, per set of stepping time levels
    struct LasExchange {
        AmrEntry *ex[3] = {nullptr, nullptr, nullptr};
        AmrProlong *pro[3] = {nullptr, nullptr, nullptr};
        int64_t nex[3] = {0, 0, 0}, npro[3] = {0, 0, 0};
    };
    std::map<int, LasExchange> las_ex;  // key: bit log2(tl) set for every stepping level
};

// ---- the device scratch of hamr_stage_adjoint_dev_noalloc ----
// The adjoint's work arrays for the entry point that allocates nothing, in one
// file-scope device array that every GPU the library runs on holds from the
// code object's load to the end of the process (kScratchBytes each).  On the
// host the array's name only stands for the symbol: kernels and copies get
// the device address that hipGetSymbolAddress returns for the current device,
// taken once per device (scratch_address) and kept in hamr_ctx::scratch.
constexpr size_t kScratchBytes = (size_t)HAMR_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;  // device -> address of `scratch` on that device

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

// Zero-filled on the context's stream: every later copy and kernel of the
// context is ordered after the fill (the stream is non-blocking, so a fill on
// the null stream could land after an upload to the same memory).
template <class T> static int dalloc(hamr_ctx *c, T **ptr, size_t n)
{
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(T) + 256);
    if (e != hipSuccess) return fail(c, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(v, 0, n * sizeof(T) + 256, c->s);
    if (e != hipSuccess) return fail(c, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    c->allocs.push_back(v);
    *ptr = (T *)v;
    return 0;
}

static int nblk_of(const hamr_ctx *c) { return (int)c->blocks.size(); }

// A named roctx range around the host-side enqueue of a stage (rocprofv3
// --marker-trace shows the ranges; the kernels they launch carry the same
// correlation).  Inside a graph capture the ranges mark the capture.
struct StageRange {
    explicit StageRange(const char *name) { roctxRangePushA(name); }
    ~StageRange() { roctxRangePop(); }
    StageRange(const StageRange &) = delete;
    StageRange &operator=(const StageRange &) = delete;
};

// Captured graphs are dropped when what they captured changes (arithmetic,
// Riemann solver, launch shapes, time levels).
static void drop_graphs(hamr_ctx *c)
{
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->las_graph) (void)hipGraphExecDestroy(c->las_graph);
    c->graph = nullptr;
    c->las_graph = nullptr;
}
static bool valid_block(const hamr_ctx *c, int n) { return n >= 0 && n < nblk_of(c); }
static int64_t ncell_of(const BlockDev &B) { return (int64_t)B.D[0] * B.D[1] * B.D[2]; }

// Pinned host staging buffer of at least `bytes` (grown on demand, kept until
// hamr_destroy).  Asynchronous copies on the context stream touch only this
// buffer, never pageable memory.
static int stage_reserve(hamr_ctx *c, size_t bytes)
{
    if (c->stage_bytes >= bytes) return 0;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->stage) HIPCHK(c, hipHostFree(c->stage));
    c->stage = nullptr;
    c->stage_bytes = 0;
    HIPCHK(c, hipHostMalloc(&c->stage, bytes, hipHostMallocDefault));
    c->stage_bytes = bytes;
    return 0;
}

// Host<->device copies ordered on the context stream and completed before return.
static int h2d(hamr_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (stage_reserve(c, bytes)) return -1;
    std::memcpy(c->stage, src, bytes);
    HIPCHK(c, hipMemcpyAsync(dst, c->stage, bytes, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}
static int d2h(hamr_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (stage_reserve(c, bytes)) return -1;
    HIPCHK(c, hipMemcpyAsync(c->stage, src, bytes, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    std::memcpy(dst, c->stage, bytes);
    return 0;
}

// ---- work lists and launches ----
static int build_class_lists(hamr_ctx *c)
{
    const int nb = nblk_of(c);
    for (int cls = 0; cls < KC_N; cls++) {
        // block list of the class: every block, except the ghost fill of a
        // multi-level set, whose kernels serve only the outflow and polar faces
        std::vector<int> bl;
        for (int n : c->order) {
            const BlockDev &B = c->blocks[n];
            bool use = true;
            if (c->amr) {
                if (cls == KC_BOUND1) use = B.kind[0] == FK_OUTFLOW || B.kind[1] == FK_OUTFLOW;
                else if (cls == KC_BOUND2) use = B.kind[2] == FK_POLE || B.kind[3] == FK_POLE;
                else if (cls == KC_BOUND3) use = false;
            }
            if (use) bl.push_back(n);
        }
        ClassList &L = c->cl[cls];
        for (int ti = 0; ti < 3; ti++) {
            const int T = kTiles[ti];
            std::vector<int2> items;
            L.start[ti].assign(nb, 0);
            L.count[ti].assign(nb, 0);
            int cols = 1;
            for (int n : bl) {
                const BlockDev &B = c->blocks[n];
                int nt;
                if (cls >= KC_LINE1 && cls <= KC_LINE3) {  // kLineT threads per tile whatever T is
                    const LineTiles lt = line_tiles(B.bs[0], B.bs[1], B.bs[2], cls - KC_LINE1 + 1);
                    nt = lt.ntiles;
                    cols = std::max(cols, lt.L * (lt.S + 3));
                } else {
                    nt = (int)((class_work(B, cls) + T - 1) / T);
                    cols = std::max(cols, cons_max_cols(B.bs[2], T));
                }
                L.start[ti][n] = (int)items.size();
                L.count[ti][n] = nt;
                for (int t = 0; t < nt; t++) items.push_back(make_int2(n, t));
            }
            L.n[ti] = (int)items.size();
            L.lds_cols[ti] = cols;
            if (!items.empty()) {
                if (dalloc(c, &L.items[ti], items.size())) return -1;
                if (h2d(c, L.items[ti], items.data(), items.size() * sizeof(int2))) return -1;
            }
        }
    }
    return 0;
}

// The class a tune slot's launches walk: the flux slots' per-face lists
// unless their line-staged form is chosen.
static bool line_flux(const hamr_ctx *c, int slot) { return slot <= TS_FLUX3 && c->tune[slot].alt; }
static int slot_class(const hamr_ctx *c, int slot)
{
    return line_flux(c, slot) ? KC_LINE1 + slot - TS_FLUX1 : kSlotClass[slot];
}

// Launch of a tune slot over its whole block list with the tuned configuration.
static Launch all_launch(const hamr_ctx *c, int slot)
{
    const TuneCfg &t = c->tune[slot];
    const ClassList &L = c->cl[slot_class(c, slot)];
    Launch x;
    x.line = line_flux(c, slot) ? 1 : 0;
    const int ti = x.line ? kDefaultTile : t.tile;  // line-staged tiles are kLineT threads
    x.threads = kTiles[ti];
    x.wl.items = L.items[ti];
    x.wl.n = L.n[ti];
    x.wl.swz = t.swz;
    x.grid = (t.grid > 0 && t.grid < x.wl.n) ? t.grid : x.wl.n;
    x.lds_cols = L.lds_cols[ti];
    return x;
}

// Launch of a class over one block (the per-block entry points).
static Launch block_launch(const hamr_ctx *c, int cls, int n)
{
    const ClassList &L = c->cl[cls];
    Launch x;
    x.threads = kTiles[kDefaultTile];
    x.wl.items = L.items[kDefaultTile] ? L.items[kDefaultTile] + L.start[kDefaultTile][n] : nullptr;
    x.wl.n = L.count[kDefaultTile][n];
    x.wl.swz = 0;
    x.grid = x.wl.n;
    x.lds_cols = L.lds_cols[kDefaultTile];
    x.line = (cls >= KC_LINE1 && cls <= KC_LINE3) ? 1 : 0;
    return x;
}

// The flux launch of direction dir over block n, in the form the flux slot is tuned to.
static Launch flux_block_launch(const hamr_ctx *c, int dir, int n)
{
    return block_launch(c, slot_class(c, TS_FLUX1 + dir - 1), n);
}

// Device geometry slabs and state arrays for the block table already in c.
// c->scratch: the address of `scratch` on the context's device (current here).
static int scratch_address(hamr_ctx *c)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(c->device);
    if (it == g_scratch_ptr.end()) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = g_scratch_ptr.emplace(c->device, (double *)ptr).first;
    }
    c->scratch = it->second;
    return 0;
}

static int init_device(hamr_ctx *c)
{
    const size_t nb = (size_t)nblk_of(c), K = (size_t)c->cells;
    if (scratch_address(c)) return -1;
    if (dalloc(c, &c->p, NPR * K) || dalloc(c, &c->ph, NPR * K) || dalloc(c, &c->ps, 3 * K) ||
        dalloc(c, &c->psh, 3 * K) || dalloc(c, &c->F[0], NPR * K) || dalloc(c, &c->F[1], NPR * K) ||
        dalloc(c, &c->F[2], NPR * K) || dalloc(c, &c->E, 3 * K) || dalloc(c, &c->Ec, 3 * K) ||
        dalloc(c, &c->U0, NPR * K) || dalloc(c, &c->Un, NPR * K) || dalloc(c, &c->pflag, K) ||
        dalloc(c, &c->failcnt, nb * 4) || dalloc(c, &c->dtmin, nb * 3) || dalloc(c, &c->dt, 4))
        return -1;
    const size_t KC = (size_t)c->cols;
    if (dalloc(c, &c->gcov, NPOS * 10 * KC) || dalloc(c, &c->gcon, NPOS * 10 * KC) || dalloc(c, &c->gdet, NPOS * KC) ||
        dalloc(c, &c->conn, 40 * KC) || dalloc(c, &c->rcen, KC) || dalloc(c, &c->tet, 96 * KC) ||
        dalloc(c, &c->g32, NPOS * G32N * KC) || dalloc(c, &c->blk, nb))
        return -1;
    {
        // (sets the blocks' kPushWhole flags before they are uploaded)
        const std::vector<PushEntry> pe = push_entries(c->blocks, c->push3tab);
        if (dalloc(c, &c->push3, std::max<size_t>(pe.size(), 1))) return -1;
        if (!pe.empty() && h2d(c, c->push3, pe.data(), pe.size() * sizeof(PushEntry))) return -1;
    }
    if (h2d(c, c->blk, c->blocks.data(), nb * sizeof(BlockDev))) return -1;
    std::vector<int> bl(nb);
    c->roff.assign(nb + 1, 0);
    c->doff.assign(nb + 1, 0);
    for (size_t n = 0; n < nb; n++) {
        const BlockDev &B = c->blocks[n];
        bl[n] = (int)n;
        c->roff[n + 1] = c->roff[n] + RREC * ncell_of(B);
        c->doff[n + 1] = c->doff[n] + DREC * (int64_t)B.bs[0] * B.bs[1] * B.bs[2];
    }
    if (dalloc(c, &c->io_blist, nb) || dalloc(c, &c->io_roff, nb + 1) || dalloc(c, &c->io_doff, nb + 1) ||
        dalloc(c, &c->bdt, nb))
        return -1;
    if (h2d(c, c->io_blist, bl.data(), nb * sizeof(int))) return -1;
    if (h2d(c, c->io_roff, c->roff.data(), (nb + 1) * sizeof(int64_t))) return -1;
    if (h2d(c, c->io_doff, c->doff.data(), (nb + 1) * sizeof(int64_t))) return -1;
    if (build_class_lists(c)) return -1;
    KParams &P = c->P;
    P.gam = c->d.gam;
    P.cour = c->d.cour;
    P.rho_floor = c->d.rho_floor;
    P.u_floor = c->d.u_floor;
    P.blk = c->blk;
    P.gcov = c->gcov;
    P.gcon = c->gcon;
    P.gdet = c->gdet;
    P.conn = c->conn;
    P.rcen = c->rcen;
    P.tet = c->tet;
    P.g32 = c->g32;
    P.push3 = c->push3;
    P.tiled_push = c->amr ? 1 : 0;
    P.xp = c->p;
    P.xph = c->ph;
    P.xps = c->ps;
    P.xpsh = c->psh;
    P.riemann = HAMR_RIEMANN_HLL;
    P.las_s = -1;
    P.las_want = LAS_IDLE;
    launch_dt_reset((int)nb, c->dtmin, c->s);
    if (check_launch(c, "dt reset")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

static int upload_rings(hamr_ctx *c, const std::vector<int> &ring, const std::vector<int> &rside, int len, int ni_max)
{
    c->nring = (int)rside.size();
    c->ring_len = len;
    c->ni_max = ni_max;
    if (!c->nring) return 0;
    if (dalloc(c, &c->ring, ring.size()) || dalloc(c, &c->rside, rside.size())) return -1;
    if (h2d(c, c->ring, ring.data(), ring.size() * sizeof(int))) return -1;
    if (h2d(c, c->rside, rside.data(), rside.size() * sizeof(int))) return -1;
    return 0;
}

static void pole(hamr_ctx *c, double dt_val, const double *dt_ptr)
{
    launch_pole(c->P, c->ring, c->rside, c->nring, c->ring_len, c->ni_max, dt_val, dt_ptr, c->E, c->prec, c->s);
}

// Ghost fill of all blocks.  Uniform grid: x1, x2, x3 stages of k_bound.
// Multi-level set: outflow faces, exchange stage 1, polar faces, exchange
// stages 2 and 3, and the final outflow pass over the radial ghosts that the
// later stages left (see hamr_amr.cpp).
// Ghost fill of every block.  x3_rows_done: the update and face-field kernels
// of the batched step already wrote the x3 ghosts of the rows inside each
// block's x1 and x2 range from its push3 neighbours (only the remaining rows
// are filled along x3: on a multi-level set stage 3 leaves out its last
// npush3 entries; the result is the full fill's, cell for cell).
static void bound_all(hamr_ctx *c, double *p, double *ps, bool x3_rows_done = false)
{
    const int T = kTiles[c->tune[TS_EXCHANGE].tile], G = c->tune[TS_EXCHANGE].grid;
    if (!c->amr) {
        for (int dir = 1; dir <= 3; dir++)
            launch_bound(c->P, dir, all_launch(c, TS_BOUND1 + dir - 1), p, ps, c->pflag, 0, c->s, x3_rows_done ? 1 : 0);
        return;
    }
    launch_bound(c->P, 1, all_launch(c, TS_BOUND1), p, ps, c->pflag, 0, c->s);
    launch_amr_exchange(c->P, 1, c->ex[0], c->nex[0], c->pro[0], c->npro[0], p, ps, c->pflag, T, G, c->s);
    launch_bound(c->P, 2, all_launch(c, TS_BOUND2), p, ps, c->pflag, 0, c->s);
    launch_amr_exchange(c->P, 2, c->ex[1], c->nex[1], c->pro[1], c->npro[1], p, ps, c->pflag, T, G, c->s);
    launch_amr_exchange(c->P, 3, c->ex[2], c->nex[2] - (x3_rows_done ? c->npush3 : 0), c->pro[2], c->npro[2], p, ps,
                        c->pflag, T, G, c->s);
    launch_bound(c->P, 1, all_launch(c, TS_BOUND1), p, ps, c->pflag, 1, c->s);
}

// Flux correction at level jumps of the corrector's fluxes (flux_rec1-3 with
// calc_corr = 1 after the fluxes of every block, step_ch.c:1142-1196); dt from
// the device when dt_ptr is set.  Nothing on a single-level set.
static void reflux_all(hamr_ctx *c, double dt_val, const double *dt_ptr)
{
    for (int d = 0; d < 3; d++)
        launch_amr_reflux(c->P, c->rfx[d], c->nrfx[d], d + 1, c->F[d], dt_val, dt_ptr, c->prec, c->s);
}

static void emf_correct_all(hamr_ctx *c, double dt_val, const double *dt_ptr)
{
    launch_amr_emf(c->P, c->emfx, c->nemfx, c->E, dt_val, dt_ptr, c->prec, c->s);
}

static void emf_all(hamr_ctx *c, const double *p)
{
    launch_emf(c->P, all_launch(c, TS_ECENT), all_launch(c, TS_ECORNER), p, c->F[0], c->F[1], c->F[2], c->Ec, c->E,
               c->prec, c->s);
}

static int max_plane(const hamr_ctx *c)
{
    int m = 0;
    for (const BlockDev &B : c->blocks) m = std::max(m, B.bs[1] * B.bs[2]);
    return m;
}

static int min_bs3(const hamr_ctx *c)
{
    int m = 1 << 30;
    for (const BlockDev &B : c->blocks) m = std::min(m, B.bs[2]);
    return m;
}

// The conserved update of every block in the form the TS_CONS tuning chose.
static void cons_all(hamr_ctx *c, const Launch &L, int full, double Dt_val, const double *dt_ptr, double dt_mult,
                     const double *pi, const double *pb, const double *psf)
{
    if (c->tune[TS_CONS].march)
        launch_cons_march(c->P, nblk_of(c), max_plane(c), min_bs3(c), full, Dt_val, dt_ptr, dt_mult, pi, pb, c->U0,
                          c->Un, psf, c->F[0], c->F[1], c->F[2], c->prec, c->s);
    else
        launch_update_part(c->P, 0, L, full, Dt_val, dt_ptr, dt_mult, pi, pb, nullptr, c->U0, c->Un, psf, c->F[0],
                           c->F[1], c->F[2], c->pflag, c->failcnt, c->prec, c->s);
}

// The update's conserved update and 2D inversion: k_update (one kernel), or
// in TS_CONS's second form k_cons (or its x1-marching form) then k_invert.
static bool fused_update(const hamr_ctx *c) { return !c->tune[TS_CONS].march && !c->tune[TS_CONS].alt; }
static void update_front(hamr_ctx *c, const Launch &L, int full, double Dt_val, const double *dt_ptr,
                         double dt_mult, const double *pi, const double *pb, double *pf, const double *psf)
{
    if (fused_update(c)) {
        launch_update_part(c->P, 3, L, full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, c->U0, c->Un, psf, c->F[0],
                           c->F[1], c->F[2], c->pflag, c->failcnt, c->prec, c->s);
        return;
    }
    cons_all(c, L, full, Dt_val, dt_ptr, dt_mult, pi, pb, psf);
    launch_update_part(c->P, 1, all_launch(c, TS_INVERT), full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, c->U0, c->Un,
                       psf, c->F[0], c->F[1], c->F[2], c->pflag, c->failcnt, c->prec, c->s);
}

static void update_all(hamr_ctx *c, int full, double Dt_val, const double *dt_ptr, double dt_mult, const double *pi,
                       const double *pb, double *pf, const double *psf)
{
    update_front(c, all_launch(c, TS_CONS), full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, psf);
    launch_update_part(c->P, 2, all_launch(c, TS_FALLBACK), full, Dt_val, dt_ptr, dt_mult, pi, pb, pf, c->U0, c->Un,
                       psf, c->F[0], c->F[1], c->F[2], c->pflag, c->failcnt, c->prec, c->s);
}

extern "C" {

const char *hamr_last_error(const hamr_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

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

int hamr_create(const hamr_desc *desc, hamr_ctx **out)
{
    if (!desc || !out) return -1;
    *out = nullptr;
    hamr_ctx *c = new hamr_ctx();
    GridDesc &d = c->d;
    c->device = desc->device;
    *out = c;
    if (desc_to_grid(desc, d))
        return fail(c, "hamr_create: each block needs >= 6 cells and >= 1 block per direction");
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
    c->push3tab = uniform_push_table(c->blocks);
    c->order = zorder(d);
    if ((int)c->order.size() != (int)c->blocks.size()) return fail(c, "hamr_create: block numbering failed");
    const Layout &L = c->L;
    c->cells = (int64_t)L.nblk * L.ksize;
    c->cols = (int64_t)L.ngeo * L.kcol;
    for (int g = 0; g < L.ngeo; g++) {
        const int c1 = g / d.nb[1], c2 = g % d.nb[1];
        c->geos.push_back(AmrGeo{0, c1 * d.bs[0], c2 * d.bs[1], L.D[0], L.D[1], (int)L.kcol, (int64_t)g * L.kcol});
    }
    if (init_device(c)) return -1;
    // pole rings: for each c1, the blocks of c2 = 0 (side 0) and c2 = NB2-1 (side 1), in c3 order
    if (!d.minkowski) {
        std::vector<int> ring, rside;
        for (int side = 0; side < 2; side++)
            for (int c1 = 0; c1 < d.nb[0]; c1++) {
                const int c2 = side ? d.nb[1] - 1 : 0;
                for (int c3 = 0; c3 < d.nb[2]; c3++) ring.push_back((c1 * d.nb[1] + c2) * d.nb[2] + c3);
                rside.push_back(side);
            }
        if (upload_rings(c, ring, rside, d.nb[2], d.bs[0] + 1)) return -1;
    }
    c->err.clear();
    return 0;
}

int hamr_create_amr(const hamr_desc *desc, const hamr_amr_desc *a, hamr_ctx **out)
{
    if (!desc || !a || !out) return -1;
    *out = nullptr;
    hamr_ctx *c = new hamr_ctx();
    c->device = desc->device;
    *out = c;
    if (desc_to_grid(desc, c->d)) return fail(c, "hamr_create_amr: bad base description");
    if (c->d.minkowski || !(c->d.rout > 0.) || !(c->d.gam > 1.))
        return fail(c, "hamr_create_amr: the multi-level grid is the torus grid (minkowski = 0, rout > 0, gam > 1)");
    AmrDesc ad;
    ad.nlev = a->nlev;
    for (int l = 0; l < kAmrMaxLevels; l++)
        for (int k = 0; k < 3; k++)
            for (int e = 0; e < 2; e++) ad.box[l][k][e] = a->box[l][k][e];
    ad.seed = a->seed;
    ad.p_big = a->p_big;
    ad.p_mid = a->p_mid;
    AmrGrid g;
    std::string e = build_amr(c->d, ad, g);
    if (!e.empty()) return fail(c, "hamr_create_amr: " + e);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    c->amr = true;
    c->L = make_layout(c->d);
    c->blocks = g.blocks;
    c->push3tab = g.push3tab;
    c->geos = g.geos;
    c->cells = g.cells;
    c->cols = g.cols;
    c->order.resize(c->blocks.size());
    for (size_t n = 0; n < c->blocks.size(); n++) c->order[n] = (int)n;
    if (init_device(c)) return -1;
    if (upload_rings(c, g.ring, g.rside, g.ring_len, g.ni_max)) return -1;
    c->npush3 = g.npush3;
    for (int s = 0; s < 3; s++) {
        c->nex[s] = (int64_t)g.ex[s].size();
        c->npro[s] = (int64_t)g.pro[s].size();
        if (c->nex[s]) {
            if (dalloc(c, &c->ex[s], g.ex[s].size())) return -1;
            if (h2d(c, c->ex[s], g.ex[s].data(), g.ex[s].size() * sizeof(AmrEntry))) return -1;
        }
        if (c->npro[s]) {
            if (dalloc(c, &c->pro[s], g.pro[s].size())) return -1;
            if (h2d(c, c->pro[s], g.pro[s].data(), g.pro[s].size() * sizeof(AmrProlong))) return -1;
        }
        c->nrfx[s] = (int64_t)g.rfx[s].size();
        if (c->nrfx[s]) {
            if (dalloc(c, &c->rfx[s], g.rfx[s].size())) return -1;
            if (h2d(c, c->rfx[s], g.rfx[s].data(), g.rfx[s].size() * sizeof(AmrReflux))) return -1;
        }
        if (s == 0 && !g.emf.empty()) {
            c->nemfx = (int64_t)g.emf.size();
            if (dalloc(c, &c->emfx, g.emf.size())) return -1;
            if (h2d(c, c->emfx, g.emf.data(), g.emf.size() * sizeof(AmrEmf))) return -1;
        }
    }
    g.ex[0].clear(); g.ex[1].clear(); g.ex[2].clear();
    g.pro[0].clear(); g.pro[1].clear(); g.pro[2].clear();
    c->grid = std::move(g);
    c->err.clear();
    return 0;
}

int hamr_amr_info(const hamr_ctx *c, long long *info)
{
    if (!c || !info) return -1;
    info[0] = nblk_of(c);
    info[1] = (long long)c->geos.size();
    info[2] = c->cells;
    long long zones = 0;
    for (const BlockDev &B : c->blocks) zones += (long long)B.bs[0] * B.bs[1] * B.bs[2];
    info[3] = zones;
    for (int s = 0; s < 3; s++) {
        info[4 + s] = c->nex[s];
        info[7 + s] = c->npro[s];
    }
    info[10] = c->nring;
    info[11] = c->amr ? 1 : 0;
    for (int d = 0; d < 3; d++) info[12 + d] = c->nrfx[d];
    info[15] = c->nemfx;
    info[16] = c->npush3;
    return 0;
}

int hamr_block_layout(const hamr_ctx *c, int n, int *row)
{
    if (!c || !row || !valid_block(c, n)) return -1;
    const BlockDev &B = c->blocks[n];
    row[0] = B.level;
    for (int k = 0; k < 3; k++) {
        row[1 + k] = B.o[k];
        row[4 + k] = B.bs[k];
    }
    row[7] = B.geo;
    row[8] = B.pole;
    for (int k = 0; k < 6; k++) row[9 + k] = B.kind[k];
    return 0;
}

int hamr_destroy(hamr_ctx *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->s) (void)hipStreamSynchronize(c->s);
    drop_graphs(c);
    for (void *v : c->allocs) (void)hipFree(v);
    for (void *v : c->las_allocs) (void)hipFree(v);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->io) (void)hipFree(c->io);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return 0;
}

int hamr_layout(const hamr_ctx *c, int *D, long long *ncell, long long *ksize, int *nblk, int *ngeo)
{
    if (!c) return -1;
    const BlockDev &B = c->blocks[0];
    if (D)
        for (int k = 0; k < 3; k++) D[k] = B.D[k];
    if (ncell) *ncell = ncell_of(B);
    if (ksize) *ksize = B.ks;
    if (nblk) *nblk = nblk_of(c);
    if (ngeo) *ngeo = (int)c->geos.size();
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
        if (z.size() != t.size()) return -2;
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
    std::string e = init_torus(d, blocks, td, gcov.data(), gcon.data(), gdet.data(), tp.data(), ts.data());
    if (!e.empty()) return -2;
    for (size_t s = 0; s < (size_t)L.nblk * NPR; s++) std::memcpy(p + s * N, tp.data() + s * K, N * sizeof(double));
    for (size_t s = 0; s < (size_t)L.nblk * 3; s++) std::memcpy(ps + s * N, ts.data() + s * K, N * sizeof(double));
    return 0;
}

int hamr_desc_amr(const hamr_desc *desc, const hamr_amr_desc *a, long long *info, int *rows, int *ex, int *pro)
{
    GridDesc d;
    if (desc_to_grid(desc, d) || !a || !info) return -1;
    AmrDesc ad;
    ad.nlev = a->nlev;
    for (int l = 0; l < kAmrMaxLevels; l++)
        for (int k = 0; k < 3; k++)
            for (int e = 0; e < 2; e++) ad.box[l][k][e] = a->box[l][k][e];
    ad.seed = a->seed;
    ad.p_big = a->p_big;
    ad.p_mid = a->p_mid;
    AmrGrid g;
    if (!build_amr(d, ad, g).empty()) return -2;
    long long zones = 0;
    for (const BlockDev &B : g.blocks) zones += (long long)B.bs[0] * B.bs[1] * B.bs[2];
    info[0] = (long long)g.blocks.size();
    info[1] = (long long)g.geos.size();
    info[2] = g.cells;
    info[3] = zones;
    for (int s = 0; s < 3; s++) {
        info[4 + s] = (long long)g.ex[s].size();
        info[7 + s] = (long long)g.pro[s].size();
    }
    info[10] = (long long)g.rside.size();
    info[11] = 1;
    for (int d = 0; d < 3; d++) info[12 + d] = (long long)g.rfx[d].size();
    info[15] = (long long)g.emf.size();
    info[16] = g.npush3;
    if (rows)
        for (size_t n = 0; n < g.blocks.size(); n++) {
            const BlockDev &B = g.blocks[n];
            int *r = rows + n * 15;
            r[0] = B.level;
            for (int k = 0; k < 3; k++) {
                r[1 + k] = B.o[k];
                r[4 + k] = B.bs[k];
            }
            r[7] = B.geo;
            r[8] = B.pole;
            for (int k = 0; k < 6; k++) r[9 + k] = B.kind[k];
        }
    if (ex)
        for (int s = 0; s < 3; s++)
            for (const AmrEntry &e : g.ex[s]) {
                std::memcpy(ex, &e, sizeof e);
                ex += 4;
            }
    if (pro)
        for (int s = 0; s < 3; s++)
            for (const AmrProlong &e : g.pro[s]) {
                std::memcpy(pro, &e, sizeof e);
                pro += 16;
            }
    return 0;
}

static int amr_from_desc(const hamr_desc *desc, const hamr_amr_desc *a, GridDesc &d, AmrGrid &g)
{
    if (desc_to_grid(desc, d) || !a) return -1;
    AmrDesc ad;
    ad.nlev = a->nlev;
    for (int l = 0; l < kAmrMaxLevels; l++)
        for (int k = 0; k < 3; k++)
            for (int e = 0; e < 2; e++) ad.box[l][k][e] = a->box[l][k][e];
    ad.seed = a->seed;
    ad.p_big = a->p_big;
    ad.p_mid = a->p_mid;
    if (!build_amr(d, ad, g).empty()) return -2;
    return 0;
}

int hamr_desc_amr_geometry(const hamr_desc *desc, const hamr_amr_desc *a, int geo, double *gcov, double *gcon,
                           double *gdet, double *conn, double *rcen)
{
    GridDesc d;
    AmrGrid g;
    int r = amr_from_desc(desc, a, d, g);
    if (r) return r;
    if (geo < 0 || geo >= (int)g.geos.size()) return -1;
    const AmrGeo &G = g.geos[geo];
    const size_t KC = (size_t)G.kcol;
    std::memset(gcov, 0, NPOS * 10 * KC * sizeof(double));
    std::memset(gcon, 0, NPOS * 10 * KC * sizeof(double));
    std::memset(gdet, 0, NPOS * KC * sizeof(double));
    std::memset(conn, 0, 40 * KC * sizeof(double));
    std::memset(rcen, 0, KC * sizeof(double));
    fill_geometry_level(d, G.level, G.o1, G.o2, G.D1, G.D2, G.kcol, gcov, gcon, gdet, conn, rcen);
    return 0;
}

int hamr_desc_amr_torus(const hamr_desc *desc, const hamr_amr_desc *a, const hamr_torus_desc *t, double *p,
                        double *ps)
{
    GridDesc d;
    AmrGrid g;
    int r = amr_from_desc(desc, a, d, g);
    if (r || !t) return r ? r : -1;
    const size_t KC = (size_t)g.cols, K = (size_t)g.cells;
    std::vector<double> gcov(NPOS * 10 * KC, 0.), gcon(NPOS * 10 * KC, 0.), gdet(NPOS * KC, 0.), conn(40 * KC, 0.),
        rcen(KC, 0.);
    for (const AmrGeo &G : g.geos)
        fill_geometry_level(d, G.level, G.o1, G.o2, G.D1, G.D2, G.kcol, gcov.data() + G.moff * NPOS * 10,
                            gcon.data() + G.moff * NPOS * 10, gdet.data() + G.moff * NPOS, conn.data() + G.moff * 40,
                            rcen.data() + G.moff);
    std::vector<double> tp(NPR * K, 0.), ts(3 * K, 0.);
    TorusDesc td{t->rin, t->rmax, t->kappa, t->beta_min, t->aphi_cut, t->perturb, t->seed};
    if (!init_torus(d, g.blocks, td, gcov.data(), gcon.data(), gdet.data(), tp.data(), ts.data()).empty()) return -3;
    // unpadded [block][component][cell]
    size_t pp = 0, sp = 0;
    for (const BlockDev &B : g.blocks) {
        const size_t N = (size_t)B.D[0] * B.D[1] * B.D[2];
        for (int k = 0; k < NPR; k++, pp += N) std::memcpy(p + pp, tp.data() + NPR * B.off + (size_t)k * B.ks, N * sizeof(double));
        for (int k = 0; k < 3; k++, sp += N) std::memcpy(ps + sp, ts.data() + 3 * B.off + (size_t)k * B.ks, N * sizeof(double));
    }
    return 0;
}

int hamr_metric_sizes(const hamr_ctx *c, long long *ncol, long long *kcol)
{
    if (!c) return -1;
    const AmrGeo &G = c->geos[0];
    if (ncol) *ncol = (long long)G.D1 * G.D2;
    if (kcol) *kcol = G.kcol;
    return 0;
}

int hamr_geometry_info(const hamr_ctx *c, int geo, long long *info)
{
    if (!c || !info || geo < 0 || geo >= (int)c->geos.size()) return -1;
    const AmrGeo &G = c->geos[geo];
    info[0] = G.level;
    info[1] = G.o1;
    info[2] = G.o2;
    info[3] = G.D1;
    info[4] = G.D2;
    info[5] = G.kcol;
    return 0;
}

int hamr_upload_metric(hamr_ctx *c, int geo, const double *gcov, const double *gcon, const double *gdet,
                       const double *conn, const double *rcen)
{
    if (!c || geo < 0 || geo >= (int)c->geos.size()) return fail(c, "hamr_upload_metric: bad geometry index");
    const AmrGeo &G = c->geos[geo];
    const size_t KC = (size_t)G.kcol, mo = (size_t)G.moff;
    if (h2d(c, c->gcov + mo * NPOS * 10, gcov, NPOS * 10 * KC * sizeof(double))) return -1;
    if (h2d(c, c->gcon + mo * NPOS * 10, gcon, NPOS * 10 * KC * sizeof(double))) return -1;
    if (h2d(c, c->gdet + mo * NPOS, gdet, NPOS * KC * sizeof(double))) return -1;
    launch_geom32(c->gcov + mo * NPOS * 10, c->gcon + mo * NPOS * 10, c->gdet + mo * NPOS, c->g32 + mo * NPOS * G32N,
                  (int64_t)KC, c->s);
    if (check_launch(c, "hamr_upload_metric")) return -1;
    if (h2d(c, c->conn + mo * 40, conn, 40 * KC * sizeof(double))) return -1;
    if (h2d(c, c->rcen + mo, rcen, KC * sizeof(double))) return -1;
    // the HLLC solver's face tetrads follow the metric (set_Mud after set_grid)
    std::vector<double> tet(96 * KC);
    fill_tetrads(gcov, gcon, (int64_t)G.D1 * G.D2, (int64_t)KC, tet.data());
    if (h2d(c, c->tet + mo * 96, tet.data(), 96 * KC * sizeof(double))) return -1;
    return 0;
}

int hamr_download_tetrads(hamr_ctx *c, int geo, double *tet)
{
    if (!c || !tet || geo < 0 || geo >= (int)c->geos.size()) return fail(c, "hamr_download_tetrads: bad arguments");
    const AmrGeo &G = c->geos[geo];
    const size_t KC = (size_t)G.kcol, mo = (size_t)G.moff;
    return d2h(c, tet, c->tet + mo * 96, 96 * KC * sizeof(double));
}

int hamr_face_tetrad(const double *gcov10, const double *gcon10, int dir, double *mud, double *mudi)
{
    if (!gcov10 || !gcon10 || !mud || !mudi || dir < 1 || dir > 3) return -1;
    face_tetrad(gcov10, gcon10, dir, mud, mudi);
    return 0;
}

int hamr_set_riemann(hamr_ctx *c, int solver)
{
    if (!c) return -1;
    if (solver != HAMR_RIEMANN_HLL && solver != HAMR_RIEMANN_HLLC) return fail(c, "hamr_set_riemann: unknown solver");
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (solver != c->P.riemann) drop_graphs(c);
    c->P.riemann = solver;
    return 0;
}

int hamr_get_riemann(const hamr_ctx *c, int *solver)
{
    if (!c || !solver) return -1;
    *solver = c->P.riemann;
    return 0;
}

int hamr_download_metric(hamr_ctx *c, int geo, double *gcov, double *gcon, double *gdet, double *conn, double *rcen)
{
    if (!c || geo < 0 || geo >= (int)c->geos.size()) return fail(c, "hamr_download_metric: bad geometry index");
    const AmrGeo &G = c->geos[geo];
    const size_t KC = (size_t)G.kcol, mo = (size_t)G.moff;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, gcov, c->gcov + mo * NPOS * 10, NPOS * 10 * KC * sizeof(double))) return -1;
    if (d2h(c, gcon, c->gcon + mo * NPOS * 10, NPOS * 10 * KC * sizeof(double))) return -1;
    if (d2h(c, gdet, c->gdet + mo * NPOS, NPOS * KC * sizeof(double))) return -1;
    if (d2h(c, conn, c->conn + mo * 40, 40 * KC * sizeof(double))) return -1;
    if (d2h(c, rcen, c->rcen + mo, KC * sizeof(double))) return -1;
    return 0;
}

int hamr_set_grid(hamr_ctx *c)
{
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    for (int geo = 0; geo < (int)c->geos.size(); geo++) {
        const AmrGeo &G = c->geos[geo];
        const size_t KC = (size_t)G.kcol;
        std::vector<double> gcov(NPOS * 10 * KC, 0.), gcon(NPOS * 10 * KC, 0.), gdet(NPOS * KC, 0.), conn(40 * KC, 0.),
            rcen(KC, 0.);
        fill_geometry_level(c->d, G.level, G.o1, G.o2, G.D1, G.D2, G.kcol, gcov.data(), gcon.data(), gdet.data(),
                            conn.data(), rcen.data());
        int r = hamr_upload_metric(c, geo, gcov.data(), gcon.data(), gdet.data(), conn.data(), rcen.data());
        if (r) return r;
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_init_torus(hamr_ctx *c, const hamr_torus_desc *t)
{
    if (!c || !t) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    const size_t K = (size_t)c->cells, KC = (size_t)c->cols;
    std::vector<double> gcov(NPOS * 10 * KC), gcon(NPOS * 10 * KC), gdet(NPOS * KC);
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, gcov.data(), c->gcov, gcov.size() * sizeof(double))) return -1;
    if (d2h(c, gcon.data(), c->gcon, gcon.size() * sizeof(double))) return -1;
    if (d2h(c, gdet.data(), c->gdet, gdet.size() * sizeof(double))) return -1;
    std::vector<double> p(NPR * K, 0.), ps(3 * K, 0.);
    TorusDesc td{t->rin, t->rmax, t->kappa, t->beta_min, t->aphi_cut, t->perturb, t->seed};
    std::string e = init_torus(c->d, c->blocks, td, gcov.data(), gcon.data(), gdet.data(), p.data(), ps.data());
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

// Copy between the unpadded [block][component][cell] layout (ext) and the
// device slabs, one 2D copy per block.
// Whole-array copy between the caller's buffer (blocks' unpadded slabs one
// after another) and the device slabs.  Host buffers pass through a pinned
// staging buffer: the asynchronous 2D copies never touch pageable memory.
static int copy_array(hamr_ctx *c, int which, double *ext, int to_device, hipMemcpyKind kind)
{
    double *d;
    int nc;
    if (array_info(c, which, &d, &nc)) return -1;
    const bool fl = is_float_array(c, which);
    const size_t es = fl ? sizeof(float) : sizeof(double);
    const bool host = kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToHost;
    if (fl && !host) return fail(c, "float array in the mixed step: host copies only");
    size_t total = 0;
    for (const BlockDev &B : c->blocks) total += (size_t)nc * ncell_of(B);
    HIPCHK(c, hipStreamSynchronize(c->s));
    char *buf = (char *)ext;
    if (host) {
        if (stage_reserve(c, total * es)) return -1;
        buf = (char *)c->stage;
        if (to_device) {
            if (fl)
                for (size_t e = 0; e < total; e++) ((float *)buf)[e] = (float)ext[e];
            else
                std::memcpy(buf, ext, total * es);
        }
    }
    size_t pos = 0;
    for (const BlockDev &B : c->blocks) {
        const size_t N = (size_t)ncell_of(B), K = (size_t)B.ks;
        char *dev = (char *)d + (size_t)nc * B.off * es;
        char *hst = buf + pos * es;
        if (to_device) HIPCHK(c, hipMemcpy2DAsync(dev, K * es, hst, N * es, N * es, nc, kind, c->s));
        else HIPCHK(c, hipMemcpy2DAsync(hst, N * es, dev, K * es, N * es, nc, kind, c->s));
        pos += (size_t)nc * N;
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (host && !to_device) {
        if (fl)
            for (size_t e = 0; e < total; e++) ext[e] = (double)((const float *)buf)[e];
        else
            std::memcpy(ext, buf, total * es);
    }
    return 0;
}

int hamr_set_array(hamr_ctx *c, int which, const double *src)
{
    if (!c || !src) return -1;
    return copy_array(c, which, const_cast<double *>(src), 1, hipMemcpyHostToDevice);
}

int hamr_get_array(hamr_ctx *c, int which, double *dst)
{
    if (!c || !dst) return -1;
    return copy_array(c, which, dst, 0, hipMemcpyDeviceToHost);
}

int hamr_set_array_dev(hamr_ctx *c, int which, const double *dsrc)
{
    if (!c || !dsrc) return -1;
    if (is_float_array(c, which)) return fail(c, "hamr_set_array_dev: float array in the mixed step");
    return copy_array(c, which, const_cast<double *>(dsrc), 1, hipMemcpyDeviceToDevice);
}

int hamr_get_array_dev(hamr_ctx *c, int which, double *ddst)
{
    if (!c || !ddst) return -1;
    if (is_float_array(c, which)) return fail(c, "hamr_get_array_dev: float array in the mixed step");
    return copy_array(c, which, ddst, 0, hipMemcpyDeviceToDevice);
}

int hamr_set_precision(hamr_ctx *c, int mode)
{
    if (!c) return -1;
    if (mode != HAMR_PREC_FP64 && mode != HAMR_PREC_MIXED) return fail(c, "hamr_set_precision: unknown mode");
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (mode != c->prec) drop_graphs(c);
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
    const BlockDev &B = c->blocks[n];
    const size_t K = (size_t)B.ks, N = (size_t)ncell_of(B);
    std::vector<double> tp(NPR * K, 0.), ts(3 * K, 0.);
    for (size_t cell = 0; cell < N; cell++) {
        for (int k = 0; k < NPR; k++) tp[k * K + cell] = p_aos[cell * NPR + k];
        if (ps_aos)
            for (int k = 0; k < 3; k++) ts[k * K + cell] = ps_aos[cell * 4 + 1 + k];
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (h2d(c, c->p + NPR * B.off, tp.data(), NPR * K * sizeof(double))) return -1;
    if (ps_aos) if (h2d(c, c->ps + 3 * B.off, ts.data(), 3 * K * sizeof(double))) return -1;
    return 0;
}

int hamr_download(hamr_ctx *c, int n, double *p_aos, double *ps_aos)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_download: bad block");
    const BlockDev &B = c->blocks[n];
    const size_t K = (size_t)B.ks, N = (size_t)ncell_of(B);
    std::vector<double> tp(NPR * K), ts(3 * K);
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (d2h(c, tp.data(), c->p + NPR * B.off, NPR * K * sizeof(double))) return -1;
    if (d2h(c, ts.data(), c->ps + 3 * B.off, 3 * K * sizeof(double))) return -1;
    for (size_t cell = 0; cell < N; cell++) {
        if (p_aos)
            for (int k = 0; k < NPR; k++) p_aos[cell * NPR + k] = tp[k * K + cell];
        if (ps_aos)
            for (int k = 0; k < 3; k++) ps_aos[cell * 4 + 1 + k] = ts[k * K + cell];
    }
    return 0;
}

// ---- restart records (rdump files) ----
// Device buffer of at least `bytes` for one block's record (grown on demand,
// kept until hamr_destroy).
static int io_reserve(hamr_ctx *c, size_t bytes)
{
    if (c->io_bytes >= bytes) return 0;
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (c->io) HIPCHK(c, hipFree(c->io));
    c->io = nullptr;
    c->io_bytes = 0;
    HIPCHK(c, hipMalloc((void **)&c->io, bytes));
    c->io_bytes = bytes;
    return 0;
}

static bool red_ok(const int *red) { return red[0] >= 1 && red[1] >= 1 && red[2] >= 1 && red[0] <= 64 && red[1] <= 64 && red[2] <= 64; }

// Doubles of a record whose block is red[d] times finer than B along d.
static int64_t rdump_count(const BlockDev &B, const int *red)
{
    int64_t n = RREC;
    for (int k = 0; k < 3; k++) n *= (int64_t)B.bs[k] * red[k] + 2 * NG;
    return n;
}

// The reader's parameters: the record's refinement and, on the torus grid
// (theta range 1 - 2/N2 of the full wedge, restart.c:249-257 with
// coord.c:397), the two grids' theta fractions.
static RdumpRead rdump_desc(const hamr_ctx *c, const int *red)
{
    RdumpRead R;
    for (int k = 0; k < 3; k++) R.red[k] = red[k];
    R.spherical = c->d.minkowski ? 0 : 1;
    const int N2 = c->d.nb[1] * c->d.bs[1];
    R.fractheta = grid_coords(c->d).fractheta;
    R.fractheta_old = 1.0 - 2.0 / ((double)N2 * red[1]);
    return R;
}

static int64_t unpack_cells(const BlockDev &B) { return (int64_t)(B.bs[0] + 1) * (B.bs[1] + 1) * (B.bs[2] + 1); }
static int64_t unpack_rows(const BlockDev &B) { return (int64_t)(B.bs[0] + 1) * (B.bs[1] + 1); }
static int64_t unpack_e3(const BlockDev &B) { return B.bs[2] + 1; }

int hamr_rdump_block_size(const hamr_ctx *c, int n, int red1, int red2, int red3, long long *count)
{
    const int red[3] = {red1, red2, red3};
    if (!c || !valid_block(c, n) || !red_ok(red) || !count) return -1;
    *count = rdump_count(c->blocks[n], red);
    return 0;
}

int hamr_rdump_offsets(const hamr_ctx *c, long long *off)
{
    if (!c || !off) return -1;
    for (size_t n = 0; n < c->roff.size(); n++) off[n] = c->roff[n];
    return 0;
}

int hamr_rdump_block_write(hamr_ctx *c, int n, double *rec)
{
    if (!c || !valid_block(c, n) || !rec) return fail(c, "hamr_rdump_block_write: bad arguments");
    const BlockDev &B = c->blocks[n];
    const size_t bytes = (size_t)(RREC * ncell_of(B)) * sizeof(double);
    if (io_reserve(c, bytes)) return -1;
    launch_rdump_pack(c->P, c->io_blist, c->io_roff, c->roff.data(), n, 1, ncell_of(B), c->p, c->ps, c->io, c->s);
    if (check_launch(c, "hamr_rdump_block_write")) return -1;
    return d2h(c, rec, c->io, bytes);
}

int hamr_rdump_block_read(hamr_ctx *c, int n, const double *rec, int red1, int red2, int red3)
{
    const int red[3] = {red1, red2, red3};
    if (!c || !valid_block(c, n) || !rec) return fail(c, "hamr_rdump_block_read: bad arguments");
    if (!red_ok(red)) return fail(c, "hamr_rdump_block_read: refinement factors must be in [1, 64]");
    const BlockDev &B = c->blocks[n];
    const size_t bytes = (size_t)rdump_count(B, red) * sizeof(double);
    if (io_reserve(c, bytes)) return -1;
    if (h2d(c, c->io, rec, bytes)) return -1;
    launch_rdump_unpack(c->P, c->io_blist, c->io_roff, c->roff.data(), n, 1, unpack_cells(B), rdump_desc(c, red),
                        c->io, c->p, c->ps, B.bs[2] + 1, unpack_rows(B), c->s);
    if (check_launch(c, "hamr_rdump_block_read")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

static int64_t max_over_blocks(const hamr_ctx *c, int64_t (*f)(const BlockDev &))
{
    int64_t m = 0;
    for (const BlockDev &B : c->blocks) m = std::max(m, f(B));
    return m;
}

int hamr_rdump_write_dev(hamr_ctx *c, double *drec)
{
    if (!c || !drec) return fail(c, "hamr_rdump_write_dev: bad arguments");
    launch_rdump_pack(c->P, c->io_blist, c->io_roff, c->roff.data(), 0, nblk_of(c), max_over_blocks(c, ncell_of),
                      c->p, c->ps, drec, c->s);
    if (check_launch(c, "hamr_rdump_write_dev")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_rdump_read_dev(hamr_ctx *c, const double *drec)
{
    if (!c || !drec) return fail(c, "hamr_rdump_read_dev: bad arguments");
    const int red[3] = {1, 1, 1};
    launch_rdump_unpack(c->P, c->io_blist, c->io_roff, c->roff.data(), 0, nblk_of(c),
                        max_over_blocks(c, unpack_cells), rdump_desc(c, red), drec, c->p, c->ps,
                        (int)max_over_blocks(c, unpack_e3), max_over_blocks(c, unpack_rows), c->s);
    if (check_launch(c, "hamr_rdump_read_dev")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

static int64_t interior_cells(const BlockDev &B) { return (int64_t)B.bs[0] * B.bs[1] * B.bs[2]; }

int hamr_dump_block_size(const hamr_ctx *c, int n, long long *count)
{
    if (!c || !valid_block(c, n) || !count) return -1;
    *count = DREC * interior_cells(c->blocks[n]);
    return 0;
}

int hamr_dump_offsets(const hamr_ctx *c, long long *off)
{
    if (!c || !off) return -1;
    for (size_t n = 0; n < c->doff.size(); n++) off[n] = c->doff[n];
    return 0;
}

int hamr_dump_block_write(hamr_ctx *c, int n, float *rec)
{
    if (!c || !valid_block(c, n) || !rec) return fail(c, "hamr_dump_block_write: bad arguments");
    const BlockDev &B = c->blocks[n];
    const size_t bytes = (size_t)(DREC * interior_cells(B)) * sizeof(float);
    if (io_reserve(c, bytes)) return -1;
    launch_dump_pack(c->P, c->io_blist, c->io_doff, c->doff.data(), n, 1, interior_cells(B), c->p, (float *)c->io,
                     c->s);
    if (check_launch(c, "hamr_dump_block_write")) return -1;
    return d2h(c, rec, c->io, bytes);
}

int hamr_dump_write_dev(hamr_ctx *c, float *drec)
{
    if (!c || !drec) return fail(c, "hamr_dump_write_dev: bad arguments");
    launch_dump_pack(c->P, c->io_blist, c->io_doff, c->doff.data(), 0, nblk_of(c), max_over_blocks(c, interior_cells),
                     c->p, drec, c->s);
    if (check_launch(c, "hamr_dump_write_dev")) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

// ---- per-block stages ----
int hamr_fluxcalc(hamr_ctx *c, int dir, int flag, int n)
{
    if (!c || !valid_block(c, n) || dir < 1 || dir > 3) return fail(c, "hamr_fluxcalc: bad arguments");
    launch_flux(c->P, dir, flux_block_launch(c, dir, n), flag ? c->ph : c->p, flag ? c->psh : c->ps,
                c->F[dir - 1], c->dtmin, c->prec, c->s);
    return check_launch(c, "hamr_fluxcalc");
}

int hamr_flux_correct(hamr_ctx *c, int flag, double dt)
{
    if (!c) return -1;
    if (flag) reflux_all(c, dt, nullptr);
    return check_launch(c, "hamr_flux_correct");
}

int hamr_emf_correct(hamr_ctx *c, int flag, double dt)
{
    if (!c) return -1;
    if (flag) emf_correct_all(c, dt, nullptr);
    return check_launch(c, "hamr_emf_correct");
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
    launch_emf(c->P, block_launch(c, KC_ECENT, n), block_launch(c, KC_CORNER, n), flag ? c->ph : c->p, c->F[0],
               c->F[1], c->F[2], c->Ec, c->E, c->prec, c->s);
    return check_launch(c, "hamr_ct_emf");
}

int hamr_ct_bound(hamr_ctx *c, double dt)
{
    if (!c) return -1;
    pole(c, dt, nullptr);
    return check_launch(c, "hamr_ct_bound");
}

int hamr_ct_update(hamr_ctx *c, int flag, double Dt, int n)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_ct_update: bad block");
    launch_ct_update(c->P, block_launch(c, KC_CORNER, n), Dt, nullptr, 1.0, c->ps, flag ? c->ps : c->psh, c->E,
                     c->prec, c->s);
    return check_launch(c, "hamr_ct_update");
}

int hamr_update(hamr_ctx *c, int flag, int n, double Dt)
{
    if (!c || !valid_block(c, n)) return fail(c, "hamr_update: bad block");
    const Launch L = block_launch(c, KC_INTERIOR, n);
    if (flag)
        launch_update(c->P, L, 1, Dt, nullptr, 1.0, c->ph, c->ph, c->p, c->U0, c->Un, c->ps, c->F[0], c->F[1], c->F[2],
                      c->pflag, c->failcnt, c->prec, c->s);
    else
        launch_update(c->P, L, 0, Dt, nullptr, 1.0, c->p, c->p, c->ph, c->U0, c->Un, c->psh, c->F[0], c->F[1], c->F[2],
                      c->pflag, c->failcnt, c->prec, c->s);
    return check_launch(c, "hamr_update");
}

int hamr_invert(hamr_ctx *c, int flag)
{
    if (!c) return -1;
    const Launch Li = all_launch(c, TS_INVERT), Lf = all_launch(c, TS_FALLBACK);
    double *guess = flag ? c->ph : c->p, *out = flag ? c->p : c->ph;
    launch_update_part(c->P, 1, Li, flag ? 1 : 0, 0., nullptr, 1.0, guess, guess, out, c->U0, c->Un, nullptr, nullptr,
                       nullptr, nullptr, c->pflag, c->failcnt, c->prec, c->s);
    launch_update_part(c->P, 2, Lf, flag ? 1 : 0, 0., nullptr, 1.0, guess, guess, out, c->U0, c->Un, nullptr, nullptr,
                       nullptr, nullptr, c->pflag, c->failcnt, c->prec, c->s);
    return check_launch(c, "hamr_invert");
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
    const BlockDev &B = c->blocks[n];
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (pflag) if (d2h(c, pflag, c->pflag + B.off, (size_t)ncell_of(B) * sizeof(int))) return -1;
    if (failimage) if (d2h(c, failimage, c->failcnt + (size_t)n * 4, 4 * sizeof(int))) return -1;
    return 0;
}

int hamr_clear_flags(hamr_ctx *c)
{
    if (!c) return -1;
    HIPCHK(c, hipMemsetAsync(c->failcnt, 0, (size_t)nblk_of(c) * 4 * sizeof(int), c->s));
    return 0;
}

// ---- batched step ----
static void enqueue_step(hamr_ctx *c)
{
    const KParams &P = c->P;
    const int nb = nblk_of(c);
    hipStream_t s = c->s;
    StageRange step("hamr_step");
    // predictor: p/ps -> ph/psh with Dt = dt/2
    {
        StageRange r("predictor: fluxes");
        for (int dir = 1; dir <= 3; dir++)
            launch_flux(P, dir, all_launch(c, TS_FLUX1 + dir - 1), c->p, c->ps, c->F[dir - 1], nullptr, c->prec, s);
    }
    {
        StageRange r("predictor: EMFs and face field");
        emf_all(c, c->p);
        pole(c, 0., c->dt);
        launch_ct_update(P, all_launch(c, TS_CT), 0., c->dt, 0.5, c->ps, c->psh, c->E, c->prec, s);
    }
    {
        StageRange r("predictor: update");
        update_all(c, 0, 0., c->dt, 0.5, c->p, c->p, c->ph, c->psh);
    }
    {
        StageRange r("predictor: ghost fill");
        bound_all(c, c->ph, c->psh, true);
    }
    // corrector: U0 -> p/ps with Dt = dt, fluxes from ph/psh, Courant minima recorded
    {
        StageRange r("corrector: fluxes and corrections");
        for (int dir = 1; dir <= 3; dir++)
            launch_flux(P, dir, all_launch(c, TS_FLUX1 + dir - 1), c->ph, c->psh, c->F[dir - 1], c->dtmin, c->prec,
                        s);
        reflux_all(c, 0., c->dt);
    }
    {
        StageRange r("corrector: EMFs and face field");
        emf_all(c, c->ph);
        pole(c, 0., c->dt);
        emf_correct_all(c, 0., c->dt);
        launch_ct_update(P, all_launch(c, TS_CT), 0., c->dt, 1.0, c->ps, c->ps, c->E, c->prec, s);
    }
    {
        StageRange r("corrector: update");
        update_all(c, 1, 0., c->dt, 1.0, c->ph, c->ph, c->p, c->ps);
    }
    {
        StageRange r("corrector: ghost fill and dt");
        bound_all(c, c->p, c->ps, true);
        launch_dt_finalize(nb, c->dtmin, c->dt, s, c->bdt);
    }
}

int hamr_estimate_dt(hamr_ctx *c, double *dt)
{
    if (!c || !dt) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    const double big[2] = {1e300, 1e300};
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (h2d(c, c->dt, big, sizeof big)) return -1;
    launch_dt_reset(nblk_of(c), c->dtmin, c->s);
    for (int dir = 1; dir <= 3; dir++)
        launch_flux(c->P, dir, all_launch(c, TS_FLUX1 + dir - 1), c->p, c->ps, c->F[dir - 1], c->dtmin, c->prec, c->s);
    launch_dt_finalize(nblk_of(c), c->dtmin, c->dt, c->s, c->bdt);
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
    launch_dt_reset(nblk_of(c), c->dtmin, c->s);
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
    launch_dt_reset(nblk_of(c), c->dtmin, c->s);
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

// ---- local adaptive time stepping (multi-level sets) ----
//
// A cycle of 2 maxtl sub-iterations s (GPU_step_ch, step_ch.c:916-979, with
// PRESTEP = PRESTEP2 = 0 and AMR_SWITCHTIMELEVEL = AMR_MAXTIMELEVEL = maxtl):
// a block of time level tl takes its predictor (Dt = tl dt / 2, p -> ph) when
// s % tl == tl - 1 and its corrector (Dt = tl dt, ph -> p) when s % 2tl ==
// 2tl - 1 (advance_GPU, step_ch.c:997-1262).  Each stage runs twice per
// sub-iteration, over the work lists of the blocks in each phase; fluxes and
// EMFs are corrected across time levels and level jumps (k_las_corr); the
// ghost fill exchanges between blocks that both step and writes both array
// pairs (k_amr_copy / k_amr_prolong); the Courant minima of the last
// sub-iteration, where every block takes its corrector, give the next dt
// (min over blocks of the harmonic, clamped to 1.3 dt).

}  // extern "C"

template <class T> static int las_alloc(hamr_ctx *c, T **ptr, size_t n)
{
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, n * sizeof(T) + 256);
    if (e != hipSuccess) return fail(c, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->las_allocs.push_back(v);
    e = hipMemsetAsync(v, 0, n * sizeof(T) + 256, c->s);
    if (e != hipSuccess) return fail(c, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    *ptr = (T *)v;
    return 0;
}

extern "C" {

static void las_release(hamr_ctx *c)
{
    (void)hipStreamSynchronize(c->s);
    for (void *v : c->las_allocs) (void)hipFree(v);
    c->las_allocs.clear();
    c->las_items = nullptr;
    c->las_off.clear();
    for (int d = 0; d < 3; d++) {
        c->las_f[d] = nullptr;
        c->acc_f[d] = nullptr;
        c->nlas_f[d] = 0;
    }
    c->las_es = c->las_ej = nullptr;
    c->acc_es = c->acc_ej = nullptr;
    c->nlas_es = c->nlas_ej = 0;
    c->las_ex.clear();
}

// Bit log2(tl) of every time level that steps (predictor or corrector) in sub-iteration s.
static int stepping_levels(int s)
{
    int m = 0;
    for (int i = 0; (1 << i) <= kMaxTimeLevel; i++)
        if (las_phase(1 << i, s) != LAS_IDLE) m |= 1 << i;
    return m;
}

// the kernel classes a cycle launches over filtered lists
constexpr int kLasClasses[] = {KC_FLUX1, KC_FLUX2, KC_FLUX3, KC_ECENT, KC_CORNER, KC_INTERIOR, KC_BOUND1, KC_BOUND2};

static int log2i(int v)
{
    int i = 0;
    while ((1 << (i + 1)) <= v) i++;
    return i;
}

// Time levels tl (validated by build_las_tables) into the block table, the
// filtered work lists and the correction tables.
static int set_timelevels(hamr_ctx *c, const std::vector<int> &tl)
{
    LasTables t;
    const std::string e = build_las_tables(c->grid, tl, t);
    if (!e.empty()) return fail(c, "hamr_set_timelevels: " + e);
    las_release(c);
    drop_graphs(c);
    const int nb = nblk_of(c);
    c->tl = tl;
    c->tlmax = 1;
    for (int n = 0; n < nb; n++) {
        c->blocks[n].tl = tl[n];
        c->tlmax = std::max(c->tlmax, tl[n]);
    }
    if (h2d(c, c->blk, c->blocks.data(), (size_t)nb * sizeof(BlockDev))) return -1;
    // one list per class for each set of time levels in one phase; the lists of
    // sub-iteration s and phase ph are those of the levels in phase ph at s
    std::map<int, std::vector<int64_t>> built;
    std::vector<int2> items;
    const int nsub = 2 * c->tlmax;
    c->las_off.assign((size_t)nsub * 2 * KC_N * 2, 0);
    for (int sub = 0; sub < nsub; sub++)
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++) {
            int mask = 0;
            for (int i = 0; (1 << i) <= kMaxTimeLevel; i++)
                if (las_phase(1 << i, sub) == ph) mask |= 1 << i;
            auto it = built.find(mask);
            if (it == built.end()) {
                std::vector<int64_t> v((size_t)KC_N * 2, 0);
                for (int cls : kLasClasses) {
                    const ClassList &L = c->cl[cls];
                    v[(size_t)cls * 2] = (int64_t)items.size();
                    for (int n : c->order) {
                        if (!((mask >> log2i(tl[n])) & 1)) continue;
                        for (int k = 0; k < L.count[kDefaultTile][n]; k++) items.push_back(make_int2(n, k));
                    }
                    v[(size_t)cls * 2 + 1] = (int64_t)items.size() - v[(size_t)cls * 2];
                }
                it = built.emplace(mask, std::move(v)).first;
            }
            std::copy(it->second.begin(), it->second.end(),
                      c->las_off.begin() + ((size_t)sub * 2 + ph - 1) * KC_N * 2);
        }
    if (!items.empty()) {
        if (las_alloc(c, &c->las_items, items.size())) return -1;
        if (h2d(c, c->las_items, items.data(), items.size() * sizeof(int2))) return -1;
    }
    // exchange entries per set of stepping levels (a sub-iteration exchanges
    // only between blocks that both step)
    {
        std::vector<AmrEntry> ex[3];
        std::vector<AmrProlong> pro[3];
        for (int k = 0; k < 3; k++) {
            ex[k].resize((size_t)c->nex[k]);
            pro[k].resize((size_t)c->npro[k]);
            if (c->nex[k] && d2h(c, ex[k].data(), c->ex[k], ex[k].size() * sizeof(AmrEntry))) return -1;
            if (c->npro[k] && d2h(c, pro[k].data(), c->pro[k], pro[k].size() * sizeof(AmrProlong))) return -1;
        }
        for (int sub = 0; sub < nsub; sub++) {
            const int m = stepping_levels(sub);
            if (c->las_ex.count(m)) continue;
            auto on = [&](int b) { return (m >> log2i(tl[b])) & 1; };
            hamr_ctx::LasExchange X;
            for (int k = 0; k < 3; k++) {
                std::vector<AmrEntry> fe;
                std::vector<AmrProlong> fp;
                // stage 3's last npush3 entries: the update writes them when both blocks step
                const size_t nk = ex[k].size() - (k == 2 ? (size_t)c->npush3 : 0);
                for (size_t q = 0; q < nk; q++) {
                    const AmrEntry &e = ex[k][q];
                    if (on(e.dblk) && on(e.sblk)) fe.push_back(e);
                }
                for (const AmrProlong &e : pro[k])
                    if (on(e.dblk) && on(e.cblk)) fp.push_back(e);
                X.nex[k] = (int64_t)fe.size();
                X.npro[k] = (int64_t)fp.size();
                if (!fe.empty() && (las_alloc(c, &X.ex[k], fe.size()) ||
                                    h2d(c, X.ex[k], fe.data(), fe.size() * sizeof(AmrEntry))))
                    return -1;
                if (!fp.empty() && (las_alloc(c, &X.pro[k], fp.size()) ||
                                    h2d(c, X.pro[k], fp.data(), fp.size() * sizeof(AmrProlong))))
                    return -1;
            }
            c->las_ex[m] = X;
        }
    }
    auto table = [&](const std::vector<LasCorr> &v, LasCorr **dst, int64_t *n, double **acc, int ncomp) {
        *n = (int64_t)v.size();
        if (v.empty()) return 0;
        if (las_alloc(c, dst, v.size()) || las_alloc(c, acc, v.size() * ncomp)) return -1;
        return h2d(c, *dst, v.data(), v.size() * sizeof(LasCorr));
    };
    for (int d = 0; d < 3; d++)
        if (table(t.flux[d], &c->las_f[d], &c->nlas_f[d], &c->acc_f[d], NPR)) return -1;
    if (table(t.emf_same, &c->las_es, &c->nlas_es, &c->acc_es, 1) ||
        table(t.emf_jump, &c->las_ej, &c->nlas_ej, &c->acc_ej, 1))
        return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

// Launch of class cls over the blocks in phase ph at sub-iteration s.
static Launch las_launch(const hamr_ctx *c, int cls, int s, int ph)
{
    const int sub = s % (2 * c->tlmax);
    const int64_t *o = &c->las_off[(((size_t)sub * 2 + ph - 1) * KC_N + cls) * 2];
    Launch x;
    x.threads = kTiles[kDefaultTile];
    x.wl.items = c->las_items ? c->las_items + o[0] : nullptr;
    x.wl.n = (int)o[1];
    x.wl.swz = 0;
    x.grid = x.wl.n;
    x.lds_cols = c->cl[cls].lds_cols[kDefaultTile];
    x.line = 0;
    return x;
}

// Arrays of the blocks in phase ph: predictor p/ps -> ph/psh, corrector ph/psh -> p/ps.
static double *las_in(hamr_ctx *c, int ph) { return ph == LAS_HALF ? c->p : c->ph; }
static double *las_ins(hamr_ctx *c, int ph) { return ph == LAS_HALF ? c->ps : c->psh; }
static double *las_out(hamr_ctx *c, int ph) { return ph == LAS_HALF ? c->ph : c->p; }
static double *las_outs(hamr_ctx *c, int ph) { return ph == LAS_HALF ? c->psh : c->ps; }

// Ghost fill at the end of sub-iteration s: physical faces of the stepping
// blocks on their updated arrays, exchanges between blocks that both step.
static void las_bound(hamr_ctx *c, int s)
{
    const KParams &P = c->P;
    const int T = kTiles[c->tune[TS_EXCHANGE].tile], G = c->tune[TS_EXCHANGE].grid;
    const hamr_ctx::LasExchange &X = c->las_ex.at(stepping_levels(s % (2 * c->tlmax)));
    auto phys = [&](int cls, int dir, int final_pass) {
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++) {
            const Launch L = las_launch(c, cls, s, ph);
            if (L.wl.n) launch_bound(P, dir, L, las_out(c, ph), las_outs(c, ph), c->pflag, final_pass, c->s);
        }
    };
    phys(KC_BOUND1, 1, 0);
    launch_amr_exchange(P, 1, X.ex[0], X.nex[0], X.pro[0], X.npro[0], c->p, c->ps, c->pflag, T, G, c->s, c->ph, c->psh);
    phys(KC_BOUND2, 2, 0);
    launch_amr_exchange(P, 2, X.ex[1], X.nex[1], X.pro[1], X.npro[1], c->p, c->ps, c->pflag, T, G, c->s, c->ph, c->psh);
    launch_amr_exchange(P, 3, X.ex[2], X.nex[2], X.pro[2], X.npro[2], c->p, c->ps, c->pflag, T, G, c->s, c->ph, c->psh);
    phys(KC_BOUND1, 1, 1);
}

static void enqueue_las_cycle(hamr_ctx *c, int maxtl)
{
    KParams &P = c->P;
    hipStream_t st = c->s;
    const int mixed = c->prec;
    // the half-step arrays start as p/ps: their ghosts next to blocks that do
    // not step with a block are the values of the last exchange, which the
    // reference's receive writes into both array pairs
    StageRange cycle("hamr_step_las: cycle");
    launch_copy(c->p, c->ph, (int64_t)(NPR * c->cells * sizeof(double)), st);
    launch_copy(c->ps, c->psh, (int64_t)(3 * c->cells * sizeof(double)), st);
    for (int s = 0; s < 2 * maxtl; s++) {
        StageRange sub("hamr_step_las: sub-iteration");
        P.las_s = s;
        P.las_want = LAS_IDLE;
        const bool last = s == 2 * maxtl - 1;
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++)
            for (int dir = 1; dir <= 3; dir++) {
                const Launch L = las_launch(c, KC_FLUX1 + dir - 1, s, ph);
                if (L.wl.n)
                    launch_flux(P, dir, L, las_in(c, ph), las_ins(c, ph), c->F[dir - 1], last ? c->dtmin : nullptr,
                                mixed, st);
            }
        for (int d = 0; d < 3; d++)
            for (int stage = 0; stage < 2; stage++)
                launch_las_corr(P, c->las_f[d], c->nlas_f[d], d + 1, stage, c->F[d], c->acc_f[d], c->dt, mixed, st);
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++) {
            const Launch Lc = las_launch(c, KC_ECENT, s, ph), Le = las_launch(c, KC_CORNER, s, ph);
            if (Lc.wl.n && Le.wl.n)
                launch_emf(P, Lc, Le, las_in(c, ph), c->F[0], c->F[1], c->F[2], c->Ec, c->E, mixed, st);
        }
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++) {
            P.las_want = ph;
            pole(c, 0., c->dt);
        }
        P.las_want = LAS_IDLE;
        for (int stage = 0; stage < 2; stage++)
            launch_las_corr(P, c->las_es, c->nlas_es, 0, stage, c->E, c->acc_es, c->dt, mixed, st);
        for (int stage = 0; stage < 2; stage++)
            launch_las_corr(P, c->las_ej, c->nlas_ej, 0, stage, c->E, c->acc_ej, c->dt, mixed, st);
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++) {
            const Launch L = las_launch(c, KC_CORNER, s, ph);
            if (L.wl.n)
                launch_ct_update(P, L, 0., c->dt, ph == LAS_HALF ? 0.5 : 1.0, c->ps, las_outs(c, ph), c->E, mixed, st);
        }
        for (int ph = LAS_HALF; ph <= LAS_FULL; ph++) {
            const Launch L = las_launch(c, KC_INTERIOR, s, ph);
            if (!L.wl.n) continue;
            const int full = ph == LAS_FULL;
            for (int part : {3, 2})
                launch_update_part(P, part, L, full, 0., c->dt, full ? 1.0 : 0.5, las_in(c, ph), las_in(c, ph),
                                   las_out(c, ph), c->U0, c->Un, las_outs(c, ph), c->F[0], c->F[1], c->F[2], c->pflag,
                                   c->failcnt, mixed, st);
        }
        las_bound(c, s);
        if (last) launch_dt_finalize(nblk_of(c), c->dtmin, c->dt, st, c->bdt);
    }
    P.las_s = -1;
    P.las_want = LAS_IDLE;
}

int hamr_las_rule(int nblk, const double *bdt, double dt, int maxtl, int nring, int ring_len, const int *ring,
                  int npair, const int *pairs, int *tl)
{
    if (nblk < 0 || (nblk && (!bdt || !tl)) || !(dt > 0.) || maxtl < 1 || maxtl > kMaxTimeLevel ||
        (maxtl & (maxtl - 1)) || nring < 0 || ring_len < 0 || npair < 0 || (nring && ring_len && !ring) ||
        (npair && !pairs))
        return -1;
    std::vector<int> r(ring, ring + (size_t)nring * ring_len);
    std::vector<std::pair<int, int>> pr;
    for (int k = 0; k < npair; k++) pr.emplace_back(pairs[2 * k], pairs[2 * k + 1]);
    for (int x : r)
        if (x < 0 || x >= nblk) return -1;
    for (const auto &q : pr)
        if (q.first < 0 || q.first >= nblk || q.second < 0 || q.second >= nblk) return -1;
    las_rule(nblk, bdt, dt, maxtl, r, ring_len, pr, tl);
    return 0;
}

int hamr_set_timelevels(hamr_ctx *c, const int *tl)
{
    if (!c || !tl) return -1;
    if (!c->amr) return fail(c, "hamr_set_timelevels: local adaptive time stepping runs on multi-level sets");
    HIPCHK(c, hipSetDevice(c->device));
    return set_timelevels(c, std::vector<int>(tl, tl + nblk_of(c)));
}

int hamr_get_timelevels(const hamr_ctx *c, int *tl)
{
    if (!c || !tl) return -1;
    for (int n = 0; n < nblk_of(c); n++) tl[n] = c->tl.empty() ? 1 : c->tl[n];
    return 0;
}

int hamr_block_dts(hamr_ctx *c, double *bdt)
{
    if (!c || !bdt) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return d2h(c, bdt, c->bdt, (size_t)nblk_of(c) * sizeof(double));
}

int hamr_timelevels_from_dt(hamr_ctx *c, int maxtl, int *tl_out)
{
    if (!c) return -1;
    if (!c->amr) return fail(c, "hamr_timelevels_from_dt: local adaptive time stepping runs on multi-level sets");
    if (maxtl < 1 || maxtl > kMaxTimeLevel || (maxtl & (maxtl - 1)))
        return fail(c, "hamr_timelevels_from_dt: maxtl is a power of 2 from 1 to 16");
    const int nb = nblk_of(c);
    std::vector<double> bdt((size_t)nb);
    double dt = 0.;
    if (hamr_block_dts(c, bdt.data())) return -1;
    if (d2h(c, &dt, c->dt, sizeof dt)) return -1;
    if (!(dt > 0.) || dt > 1e299) return fail(c, "hamr_timelevels_from_dt: no step has set dt yet");
    std::vector<int> tl((size_t)nb);
    las_rule(nb, bdt.data(), dt, maxtl, c->grid.ring, c->grid.ring_len, jump_pairs(c->grid), tl.data());
    if (set_timelevels(c, tl)) return -1;
    if (tl_out) std::copy(tl.begin(), tl.end(), tl_out);
    return 0;
}

int hamr_step_las_graph(hamr_ctx *c, int maxtl, int ncycles, double *dt_inout)
{
    if (!c) return -1;
    if (!c->amr || c->tl.empty())
        return fail(c, "hamr_step_las_graph: set the time levels of a multi-level set first (hamr_set_timelevels)");
    if (maxtl < c->tlmax || maxtl > kMaxTimeLevel || (maxtl & (maxtl - 1)))
        return fail(c, "hamr_step_las_graph: maxtl is a power of 2 from the largest time level to 16");
    if (ncycles < 0) return fail(c, "hamr_step_las_graph: ncycles < 0");
    HIPCHK(c, hipSetDevice(c->device));
    if (set_dt(c, dt_inout)) return -1;
    launch_dt_reset(nblk_of(c), c->dtmin, c->s);
    if (c->las_graph && c->las_graph_maxtl != maxtl) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        (void)hipGraphExecDestroy(c->las_graph);
        c->las_graph = nullptr;
    }
    if (!c->las_graph) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        enqueue_las_cycle(c, maxtl);
        HIPCHK(c, hipStreamEndCapture(c->s, &g));
        HIPCHK(c, hipGraphInstantiate(&c->las_graph, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        c->las_graph_maxtl = maxtl;
    }
    for (int k = 0; k < ncycles; k++) HIPCHK(c, hipGraphLaunch(c->las_graph, c->s));
    if (dt_inout) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (d2h(c, dt_inout, c->dt, sizeof(double))) return -1;
    }
    return 0;
}

int hamr_las_info(const hamr_ctx *c, long long *info)
{
    if (!c || !info) return -1;
    for (int d = 0; d < 3; d++) info[d] = c->nlas_f[d];
    info[3] = c->nlas_es;
    info[4] = c->nlas_ej;
    info[5] = c->tl.empty() ? 0 : c->tlmax;
    return 0;
}

int hamr_step_las(hamr_ctx *c, int maxtl, int ncycles, double *dt_inout)
{
    if (!c) return -1;
    if (!c->amr || c->tl.empty())
        return fail(c, "hamr_step_las: set the time levels of a multi-level set first (hamr_set_timelevels)");
    if (maxtl < c->tlmax || maxtl > kMaxTimeLevel || (maxtl & (maxtl - 1)))
        return fail(c, "hamr_step_las: maxtl is a power of 2 from the largest time level to 16");
    if (ncycles < 0) return fail(c, "hamr_step_las: ncycles < 0");
    HIPCHK(c, hipSetDevice(c->device));
    if (set_dt(c, dt_inout)) return -1;
    launch_dt_reset(nblk_of(c), c->dtmin, c->s);
    for (int k = 0; k < ncycles; k++) enqueue_las_cycle(c, maxtl);
    if (check_launch(c, "hamr_step_las")) return -1;
    if (dt_inout) {
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (d2h(c, dt_inout, c->dt, sizeof(double))) return -1;
    }
    return 0;
}

// One launch of tune slot `slot` with launch configuration L (predictor-step
// operands).  p and ps are read only, except by the ghost-fill kernels, which
// rewrite ghost cells of a filled state with the same values.
static void run_slot(hamr_ctx *c, int slot, const Launch &L, int T_ex, int G_ex)
{
    const KParams &P = c->P;
    hipStream_t s = c->s;
    switch (slot) {
    case TS_FLUX1:
    case TS_FLUX2:
    case TS_FLUX3: launch_flux(P, slot - TS_FLUX1 + 1, L, c->p, c->ps, c->F[slot - TS_FLUX1], nullptr, c->prec, s); break;
    case TS_ECENT: launch_emf_centre(P, L, c->p, c->Ec, c->prec, s); break;
    case TS_ECORNER: launch_emf_corner(P, L, c->Ec, c->F[0], c->F[1], c->F[2], c->E, c->prec, s); break;
    case TS_CT: launch_ct_update(P, L, 0., c->dt, 0.5, c->ps, c->psh, c->E, c->prec, s); break;
    case TS_CONS: update_front(c, L, 0, 0., c->dt, 0.5, c->p, c->p, c->ph, c->psh); break;
    case TS_INVERT:
    case TS_FALLBACK:
        launch_update_part(P, slot - TS_CONS, L, 0, 0., c->dt, 0.5, c->p, c->p, c->ph, c->U0, c->Un, c->psh, c->F[0],
                           c->F[1], c->F[2], c->pflag, c->failcnt, c->prec, s);
        break;
    case TS_BOUND1:
    case TS_BOUND2:
    case TS_BOUND3: launch_bound(P, slot - TS_BOUND1 + 1, L, c->p, c->ps, c->pflag, 0, s); break;
    default:
        for (int st = 0; st < 3; st++)
            launch_amr_exchange(P, st + 1, c->ex[st], c->nex[st], c->pro[st], c->npro[st], c->p, c->ps, c->pflag, T_ex, G_ex, s);
        break;
    }
}

static int time_launches(hamr_ctx *c, int slot, const Launch &L, int T_ex, int G_ex, int nrep, float *ms)
{
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    run_slot(c, slot, L, T_ex, G_ex);  // warm-up
    HIPCHK(c, hipEventRecord(e0, c->s));
    for (int r = 0; r < nrep; r++) run_slot(c, slot, L, T_ex, G_ex);
    HIPCHK(c, hipEventRecord(e1, c->s));
    HIPCHK(c, hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / nrep;
    return check_launch(c, "autotune");
}

// Timing and tuning launches run the stages on the current state: the update
// kernels overwrite the flags and failure counters (and ph, U0, Un), the
// face-field update psh.  These keep the flags and counters of the last step
// and leave p/ps and ph/psh with consistent ghosts afterwards.
struct FlagCopy {
    std::vector<int> fc, fl;
};
static int save_flags(hamr_ctx *c, FlagCopy &k)
{
    HIPCHK(c, hipStreamSynchronize(c->s));
    k.fc.resize((size_t)nblk_of(c) * 4);
    k.fl.resize((size_t)c->cells);
    if (d2h(c, k.fc.data(), c->failcnt, k.fc.size() * sizeof(int))) return -1;
    return d2h(c, k.fl.data(), c->pflag, k.fl.size() * sizeof(int));
}
static int restore_flags(hamr_ctx *c, const FlagCopy &k, const char *what)
{
    if (h2d(c, c->failcnt, k.fc.data(), k.fc.size() * sizeof(int))) return -1;
    if (h2d(c, c->pflag, k.fl.data(), k.fl.size() * sizeof(int))) return -1;
    bound_all(c, c->p, c->ps);
    bound_all(c, c->ph, c->psh);
    if (check_launch(c, what)) return -1;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int hamr_autotune(hamr_ctx *c, int nrep)
{
    if (!c || nrep < 1) return fail(c, "hamr_autotune: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    FlagCopy keepf;
    if (save_flags(c, keepf)) return -1;
    static const int grids[4] = {0, 1024, 2048, 4096};
    for (int slot = 0; slot < TS_N; slot++) {
        if (slot == TS_EXCHANGE && !c->amr) continue;
        if (slot != TS_EXCHANGE && c->cl[kSlotClass[slot]].n[kDefaultTile] == 0) continue;
        TuneCfg best;
        // second forms: the fluxes line-staged, the update k_cons + k_invert
        const int nform = (slot <= TS_FLUX3 || slot == TS_CONS) ? 2 : 1;
        for (int form = 0; form < nform; form++)
        for (int ti = 0; ti < 3; ti++)
            for (int gi = 0; gi < 4; gi++)
                for (int swz = 0; swz < 2; swz++) {
                    if (slot == TS_EXCHANGE && swz) continue;
                    if (slot <= TS_FLUX3 && form == 1 && kTiles[ti] != kLineT) continue;
                    TuneCfg cand;
                    cand.tile = ti;
                    cand.grid = grids[gi];
                    cand.swz = swz;
                    cand.alt = form;
                    const TuneCfg keep = c->tune[slot];
                    c->tune[slot] = cand;
                    Launch L;
                    if (slot != TS_EXCHANGE) L = all_launch(c, slot);
                    float ms;
                    const int r = time_launches(c, slot, L, kTiles[ti], grids[gi], nrep, &ms);
                    c->tune[slot] = keep;
                    if (r) return -1;
                    cand.ms = ms;
                    if (best.ms < 0.f || ms < best.ms) best = cand;
                }
        if (slot == TS_CONS) {  // the x1-marching form of the update
            TuneCfg cand;
            cand.march = 1;
            const TuneCfg keep = c->tune[slot];
            c->tune[slot] = cand;
            float ms;
            const int r = time_launches(c, slot, all_launch(c, slot), 0, 0, nrep, &ms);
            c->tune[slot] = keep;
            if (r) return -1;
            cand.ms = ms;
            if (ms < best.ms) best = cand;
        }
        c->tune[slot] = best;
    }
    if (restore_flags(c, keepf, "autotune")) return -1;
    drop_graphs(c);
    return 0;
}

int hamr_time_kernel(hamr_ctx *c, int slot, int nrep, double *avg_ms)
{
    if (!c || !avg_ms || nrep < 1 || slot < 0 || slot >= TS_N) return fail(c, "hamr_time_kernel: bad arguments");
    if (slot == TS_EXCHANGE && !c->amr) {
        *avg_ms = 0.;
        return 0;
    }
    HIPCHK(c, hipSetDevice(c->device));
    FlagCopy keepf;
    if (save_flags(c, keepf)) return -1;
    Launch L{};
    if (slot != TS_EXCHANGE) L = all_launch(c, slot);
    float ms = 0.f;
    if (time_launches(c, slot, L, kTiles[c->tune[TS_EXCHANGE].tile], c->tune[TS_EXCHANGE].grid, nrep, &ms)) return -1;
    if (restore_flags(c, keepf, "hamr_time_kernel")) return -1;
    *avg_ms = ms;
    return 0;
}

int hamr_tune_get(const hamr_ctx *c, int slot, int *cfg, double *ms)
{
    if (!c || slot < 0 || slot >= TS_N || !cfg) return -1;
    const TuneCfg &t = c->tune[slot];
    cfg[0] = t.march ? 128 : kTiles[t.tile];
    cfg[1] = t.march ? -1 : t.grid;
    cfg[2] = t.march ? 0 : t.swz | (t.alt << 1);
    if (ms) *ms = t.ms;
    return 0;
}

int hamr_tune_set(hamr_ctx *c, int slot, const int *cfg)
{
    if (!c || slot < 0 || slot >= TS_N || !cfg) return -1;
    int ti = -1;
    for (int k = 0; k < 3; k++)
        if (kTiles[k] == cfg[0]) ti = k;
    const bool march = slot == TS_CONS && cfg[1] == -1;
    const int maxbits = (slot <= TS_FLUX3 || slot == TS_CONS) ? 3 : 1;
    if (!march && (ti < 0 || cfg[1] < 0 || cfg[2] < 0 || cfg[2] > maxbits))
        return fail(c, "hamr_tune_set: tile 64/128/256, grid >= 0 (-1 for the update's x1-marching form), "
                       "cfg[2] bit 0 XCD order, bit 1 the second form (fluxes: line-staged tiles; "
                       "update: k_cons + k_invert)");
    c->tune[slot].tile = march ? kDefaultTile : ti;
    c->tune[slot].grid = march ? 0 : cfg[1];
    c->tune[slot].swz = march ? 0 : (cfg[2] & 1);
    c->tune[slot].alt = march ? 0 : (cfg[2] >> 1) & 1;
    c->tune[slot].march = march ? 1 : 0;
    c->tune[slot].ms = -1.f;
    HIPCHK(c, hipStreamSynchronize(c->s));
    drop_graphs(c);
    return 0;
}

int hamr_time_stage(hamr_ctx *c, int stage, int nrep, double *avg_ms)
{
    if (!c || !avg_ms || nrep < 1 || stage < 0 || stage > 7) return fail(c, "hamr_time_stage: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->s));
    for (int r = 0; r < nrep; r++) {
        switch (stage) {
        case 0:
        case 1:
        case 2:
            launch_flux(c->P, stage + 1, all_launch(c, TS_FLUX1 + stage), c->p, c->ps, c->F[stage], nullptr, c->prec, c->s);
            break;
        case 3: emf_all(c, c->p); break;
        case 4: pole(c, 0., c->dt); break;
        case 5: launch_ct_update(c->P, all_launch(c, TS_CT), 0., c->dt, 0.5, c->ps, c->psh, c->E, c->prec, c->s); break;
        case 6: update_all(c, 0, 0., c->dt, 0.5, c->p, c->p, c->ph, c->psh); break;
        default: bound_all(c, c->ph, c->psh, true); break;
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
    bound_all(c, c->p, c->ps);
    for (int dir = 1; dir <= 3; dir++)
        launch_flux(c->P, dir, all_launch(c, TS_FLUX1 + dir - 1), c->p, c->ps, c->F[dir - 1], nullptr, c->prec, c->s);
    emf_all(c, c->p);
    pole(c, Dt, nullptr);
    launch_ct_update(c->P, all_launch(c, TS_CT), Dt, nullptr, 1.0, c->ps, c->psh, c->E, c->prec, c->s);
    launch_cons(c->P, all_launch(c, TS_CONS), Dt, c->p, c->U0, c->Un, c->psh, c->F[0], c->F[1], c->F[2], c->prec, c->s);
}

int hamr_stage_forward(hamr_ctx *c, double Dt)
{
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    enqueue_stage_forward(c, Dt);
    return check_launch(c, "hamr_stage_forward");
}

// The adjoint's work arrays: weights, lambda of P, ps, the updated face field,
// the corner and centre EMFs and the three fluxes (57 doubles per cell of the
// padded slab).
struct AdjBufs {
    double *w, *lP, *lps, *lpsf, *lE, *lEc, *lF[3];
};
constexpr int64_t kAdjDoublesPerCell = NPR + NPR + 3 + 3 + 3 + 3 + 3 * NPR;

// w, grad_p, grad_ps: pointers of `kind` (host or device), unpadded layout.
static int stage_adjoint(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps, hipMemcpyKind in,
                         hipMemcpyKind out, const AdjBufs &A)
{
    const BlockDev &B = c->blocks[0];
    const size_t K = (size_t)B.ks, N = (size_t)ncell_of(B);
    HIPCHK(c, hipStreamSynchronize(c->s));
    // host buffers pass through the pinned staging buffer (w in, then the gradients out)
    const bool hin = in == hipMemcpyHostToDevice, hout = out == hipMemcpyDeviceToHost;
    if ((hin || hout) && stage_reserve(c, (NPR + 3) * N * sizeof(double))) return -1;
    double *sw = (double *)c->stage;
    if (hin) std::memcpy(sw, w, NPR * N * sizeof(double));
    HIPCHK(c, hipMemcpy2DAsync(A.w, K * sizeof(double), hin ? sw : w, N * sizeof(double), N * sizeof(double), NPR,
                               in, c->s));
    if (hin) HIPCHK(c, hipStreamSynchronize(c->s));
    // the fluxes of the stage (their mass-flux signs select the EMF upwinding)
    bound_all(c, c->p, c->ps);
    for (int dir = 1; dir <= 3; dir++)
        launch_flux(c->P, dir, flux_block_launch(c, dir, 0), c->p, c->ps, c->F[dir - 1], nullptr,
                    HAMR_PREC_FP64, c->s);
    launch_adjoint(c->P, B, Dt, c->p, c->ps, c->F[0], c->F[1], c->F[2], A.w, A.lP, A.lps, A.lpsf, A.lE, A.lEc,
                   A.lF[0], A.lF[1], A.lF[2], c->s);
    if (check_launch(c, "hamr_stage_adjoint")) return -1;
    double *sp = hout ? (double *)c->stage : gp, *sps = hout ? (double *)c->stage + NPR * N : gps;
    if (gp)
        HIPCHK(c, hipMemcpy2DAsync(sp, N * sizeof(double), A.lP, K * sizeof(double), N * sizeof(double), NPR, out, c->s));
    if (gps)
        HIPCHK(c, hipMemcpy2DAsync(sps, N * sizeof(double), A.lps, K * sizeof(double), N * sizeof(double), 3, out, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (hout) {
        if (gp) std::memcpy(gp, sp, NPR * N * sizeof(double));
        if (gps) std::memcpy(gps, sps, 3 * N * sizeof(double));
    }
    return 0;
}

static int adjoint_ok(hamr_ctx *c, const double *w, const char *what)
{
    if (!c) return -1;
    if (!w) return fail(c, std::string(what) + ": no weights");
    if (!c->d.minkowski || c->amr || nblk_of(c) != 1 || c->prec != HAMR_PREC_FP64 || c->P.riemann != HAMR_RIEMANN_HLL)
        return fail(c, std::string(what) + ": needs a flat periodic box of one block in FP64 with HLL fluxes");
    HIPCHK(c, hipSetDevice(c->device));
    return 0;
}

// The work arrays of the allocating entry points: the context's own, allocated
// by the first adjoint call.
static int adjoint_bufs(hamr_ctx *c, AdjBufs &A)
{
    const size_t K = (size_t)c->blocks[0].ks;
    if (!c->aw) {
        if (dalloc(c, &c->aw, NPR * K) || dalloc(c, &c->alP, NPR * K) || dalloc(c, &c->alps, 3 * K) ||
            dalloc(c, &c->alpsf, 3 * K) || dalloc(c, &c->alE, 3 * K) || dalloc(c, &c->alEc, 3 * K) ||
            dalloc(c, &c->alF[0], NPR * K) || dalloc(c, &c->alF[1], NPR * K) || dalloc(c, &c->alF[2], NPR * K))
            return -1;
    }
    A = AdjBufs{c->aw, c->alP, c->alps, c->alpsf, c->alE, c->alEc, {c->alF[0], c->alF[1], c->alF[2]}};
    return 0;
}

int hamr_stage_adjoint(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps)
{
    AdjBufs A;
    if (adjoint_ok(c, w, "hamr_stage_adjoint") || adjoint_bufs(c, A)) return -1;
    return stage_adjoint(c, Dt, w, gp, gps, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, A);
}

int hamr_stage_adjoint_dev(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps)
{
    AdjBufs A;
    if (adjoint_ok(c, w, "hamr_stage_adjoint_dev") || adjoint_bufs(c, A)) return -1;
    return stage_adjoint(c, Dt, w, gp, gps, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice, A);
}

// The work arrays carved from this device's `scratch` (c->scratch, resolved at
// hamr_create): no allocation.  Every accumulator is cleared by a fill kernel
// inside launch_adjoint; the rest is written before it is read.
int hamr_stage_adjoint_dev_noalloc(hamr_ctx *c, double Dt, const double *w, double *gp, double *gps)
{
    if (adjoint_ok(c, w, "hamr_stage_adjoint_dev_noalloc")) return -1;
    const int64_t K = c->blocks[0].ks;
    const int64_t need = kAdjDoublesPerCell * K * (int64_t)sizeof(double);
    if (need > (int64_t)kScratchBytes)
        return fail(c, "hamr_stage_adjoint_dev_noalloc: the box needs " + std::to_string(need) +
                           " bytes of work arrays, above the " + std::to_string(kScratchBytes) +
                           "-byte scratch (HAMR_NOALLOC_SCRATCH_BYTES)");
    if (!c->scratch) return fail(c, "hamr_stage_adjoint_dev_noalloc: no scratch address on this device");
    double *q = c->scratch;
    auto take = [&](int64_t ncomp) {
        double *r = q;
        q += ncomp * K;
        return r;
    };
    AdjBufs A;
    A.w = take(NPR);
    A.lP = take(NPR);
    A.lps = take(3);
    A.lpsf = take(3);
    A.lE = take(3);
    A.lEc = take(3);
    for (int d = 0; d < 3; d++) A.lF[d] = take(NPR);
    return stage_adjoint(c, Dt, w, gp, gps, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice, A);
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
