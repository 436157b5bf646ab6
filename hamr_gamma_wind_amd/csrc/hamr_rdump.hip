// hamr_rdump.hip -- a block's restart record (the rdump files of restart.c)
// packed from and unpacked into the device state, and its dump record (the
// new_dump files of dump.c).
//
// Record of one block (restart.c:52-67, raw doubles, no header): the block's
// ghost-inclusive cells in index_3D order (z fastest), RREC = NPR + 4 values per
// cell: p[0..NPR-1], the slot of face-field component 0 (which the reader never
// uses; written as 0), then B^1 gdet(FACE1), B^2 gdet(FACE2), B^3 gdet(FACE3)
// with gdet of the cell's own (i, j) column.
#include "hamr_kernels_dev.h"

namespace hamr {

constexpr int RPACK_T = 256;

// One workgroup packs RPACK_T consecutive cells of one block: each thread
// gathers its cell's 9 primitives and 3 face-field components (coalesced slab
// reads) into LDS in record order, then the workgroup stores the RPACK_T*RREC
// doubles as one contiguous run, so the store side is coalesced too.
__global__ __launch_bounds__(RPACK_T) void k_rdump_pack(const KParams P, const int *blist, const int64_t *roff,
                                                        int first, int64_t base, const double *p, const double *ps,
                                                        double *out)
{
    __shared__ double s[RPACK_T * RREC];
    const int item = first + (int)blockIdx.y;
    const BlockDev &B = P.blk[blist[item]];
    const int64_t N = (int64_t)B.D[0] * B.D[1] * B.D[2];
    const int64_t c0 = (int64_t)blockIdx.x * RPACK_T;
    if (c0 >= N) return;  // the whole workgroup leaves together
    const int nv = (int)min((int64_t)RPACK_T, N - c0);
    const int t = threadIdx.x;
    if (t < nv) {
        const int64_t cell = c0 + t, col = cell / B.D[2];
        const double *pc = p + NPR * B.off + cell;
        const double *sc = ps + 3 * B.off + cell;
        double *r = s + t * RREC;
#pragma unroll
        for (int k = 0; k < NPR; k++) r[k] = pc[(int64_t)k * B.ks];
        r[NPR] = 0.;
        r[NPR + 1] = sc[0] * P.gdet[mg1(B, FACE1) + col];
        r[NPR + 2] = sc[B.ks] * P.gdet[mg1(B, FACE2) + col];
        r[NPR + 3] = sc[2 * (int64_t)B.ks] * P.gdet[mg1(B, FACE3) + col];
    }
    __syncthreads();
    double *o = out + (roff[item] - base) + c0 * RREC;
    for (int e = t; e < nv * RREC; e += RPACK_T) o[e] = s[e];
}

// Unpack (restart.c:233-269): one thread per cell (i, j, z) of [NG, NG+BS]^3,
// the interior and the first upper ghost row of each direction, the range the
// reference assigns.  The record is that of a block red times finer along each
// direction (red = 1: the block itself).  A cell's primitives are the mean of
// the fine cells it covers; the fine cells past the record's last upper ghost
// row (x = NG + BS red) are not in the range, so an upper-row cell sums fewer
// terms with the same 1/(red1 red2 red3) weight, as the reference does.  Each
// face-field component is the mean over the fine faces that tile the coarse
// face, divided by this grid's gdet of that face; B^2 carries the ratio of the
// two grids' theta extents (spherical grids).  Sums run in the reference's
// visiting order of the fine cells (x1 slowest, x3 fastest), starting from 0.
__global__ void k_rdump_unpack(const KParams P, const int *blist, const int64_t *roff, int first, int64_t base,
                               RdumpRead R, const double *rec, double *p, double *ps)
{
    const int item = first + (int)blockIdx.y;
    const BlockDev &B = P.blk[blist[item]];
    const int e2 = B.bs[1] + 1, e3 = B.bs[2] + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (B.bs[0] + 1) * e2 * e3) return;
    int a, b, c;
    region_index(t, e2, e3, a, b, c);
    const int r1 = R.red[0], r2 = R.red[1], r3 = R.red[2];
    const int D2 = B.bs[1] * r2 + 2 * NG, D3 = B.bs[2] * r3 + 2 * NG;
    const int f1 = NG + a * r1, f2 = NG + b * r2, f3 = NG + c * r3;
    const int n1 = a < B.bs[0] ? r1 : 1, n2 = b < B.bs[1] ? r2 : 1, n3 = c < B.bs[2] ? r3 : 1;
    const double *r0 = rec + (roff[item] - base);
    auto at = [&](int x, int y, int w) { return r0 + (((int64_t)x * D2 + y) * D3 + w) * RREC; };

    const double rf = 1.0 / (double)(r1 * r2 * r3);
    double acc[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) acc[k] = 0.;
    for (int x = 0; x < n1; x++)
        for (int y = 0; y < n2; y++)
            for (int w = 0; w < n3; w++) {
                const double *v = at(f1 + x, f2 + y, f3 + w);
#pragma unroll
                for (int k = 0; k < NPR; k++) acc[k] += v[k] * rf;
            }

    const int i = NG + a, j = NG + b, z = NG + c;
    const int64_t col = (int64_t)i * B.D[1] + j;
    const double g1 = P.gdet[mg1(B, FACE1) + col], g2 = P.gdet[mg1(B, FACE2) + col];
    const double g3 = P.gdet[mg1(B, FACE3) + col];
    double s1 = 0., s2 = 0., s3 = 0.;
    const double rf1 = 1.0 / (double)(r2 * r3), rf2 = 1.0 / (double)(r1 * r3), rf3 = 1.0 / (double)(r1 * r2);
    for (int y = 0; y < n2; y++)
        for (int w = 0; w < n3; w++) s1 += at(f1, f2 + y, f3 + w)[NPR + 1] * rf1 / g1;
    for (int x = 0; x < n1; x++)
        for (int w = 0; w < n3; w++) {
            const double v = at(f1 + x, f2, f3 + w)[NPR + 2] * rf2 / g2;
            s2 += R.spherical ? v * R.fractheta / R.fractheta_old : v;
        }
    for (int x = 0; x < n1; x++)
        for (int y = 0; y < n2; y++) s3 += at(f1 + x, f2 + y, f3)[NPR + 3] * rf3 / g3;

    const int64_t cell = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
    double *pc = p + NPR * B.off + cell;
#pragma unroll
    for (int k = 0; k < NPR; k++) pc[(int64_t)k * B.ks] = acc[k];
    double *sc = ps + 3 * B.off + cell;
    sc[0] = s1;
    sc[B.ks] = s2;
    sc[2 * (int64_t)B.ks] = s3;
}

// The unpack of a record of the block's own size (red = 1), staged: a
// workgroup takes whole z-rows (i, j) of the destination range, copies their
// records (contiguous runs of (BS3+1) RREC doubles) into LDS with coalesced
// loads, then each thread forms its cell as k_rdump_unpack does (the same
// operations: sums of one term from 0, weights 1).  Needs BS3 + 1 <= RPACK_T.
__global__ __launch_bounds__(RPACK_T) void k_rdump_unpack1(const KParams P, const int *blist, const int64_t *roff,
                                                           int first, int64_t base, RdumpRead R, const double *rec,
                                                           double *p, double *ps)
{
    __shared__ double s[RPACK_T * RREC];
    const int item = first + (int)blockIdx.y;
    const BlockDev &B = P.blk[blist[item]];
    const int e2 = B.bs[1] + 1, e3 = B.bs[2] + 1;
    const int rows = RPACK_T / e3, nrows = (B.bs[0] + 1) * e2;
    const int row0 = (int)blockIdx.x * rows;
    if (row0 >= nrows) return;  // the whole workgroup leaves together
    const int nr = min(rows, nrows - row0);
    const int run = e3 * RREC;
    const double *r0 = rec + (roff[item] - base);
    for (int e = threadIdx.x; e < nr * run; e += RPACK_T) {
        const int rr = e / run, off = e - rr * run, row = row0 + rr;
        const int a = row / e2, b = row - a * e2;
        s[e] = r0[((((int64_t)(NG + a) * B.D[1] + NG + b) * B.D[2]) + NG) * RREC + off];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= nr * e3) return;
    const int rr = t / e3, c = t - rr * e3, row = row0 + rr;
    const int a = row / e2, b = row - a * e2;
    const double *v = s + t * RREC;
    const double rf = 1.0 / (double)(1 * 1 * 1);
    const int i = NG + a, j = NG + b, z = NG + c;
    const int64_t col = (int64_t)i * B.D[1] + j;
    const double g1 = P.gdet[mg1(B, FACE1) + col], g2 = P.gdet[mg1(B, FACE2) + col];
    const double g3 = P.gdet[mg1(B, FACE3) + col];
    const int64_t cell = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
    double *pc = p + NPR * B.off + cell;
#pragma unroll
    for (int k = 0; k < NPR; k++) {
        double acc = 0.;
        acc += v[k] * rf;
        pc[(int64_t)k * B.ks] = acc;
    }
    double s1 = 0., s2 = 0., s3 = 0.;
    s1 += v[NPR + 1] * rf / g1;
    const double v2 = v[NPR + 2] * rf / g2;
    s2 += R.spherical ? v2 * R.fractheta / R.fractheta_old : v2;
    s3 += v[NPR + 3] * rf / g3;
    double *sc = ps + 3 * B.off + cell;
    sc[0] = s1;
    sc[B.ks] = s2;
    sc[2 * (int64_t)B.ks] = s3;
}

// Dump record of a block (dump_block, dump.c:276-356, ideal gas, no optional
// physics): the interior cells in index_3D order, DREC = 9 floats each: rho,
// u, u^0..u^3 (ucon_calc at the cell centre, phys.c:309) and the cell-centred
// B^1..B^3.  256 cells per workgroup, staged through LDS as the restart record.
__global__ __launch_bounds__(RPACK_T) void k_dump_pack(const KParams P, const int *blist, const int64_t *doff,
                                                       int first, int64_t base, const double *p, float *out)
{
    __shared__ float s[RPACK_T * DREC];
    const int item = first + (int)blockIdx.y;
    const BlockDev &B = P.blk[blist[item]];
    const int64_t N = (int64_t)B.bs[0] * B.bs[1] * B.bs[2];
    const int64_t c0 = (int64_t)blockIdx.x * RPACK_T;
    if (c0 >= N) return;  // the whole workgroup leaves together
    const int nv = (int)min((int64_t)RPACK_T, N - c0);
    const int t = threadIdx.x;
    if (t < nv) {
        int a, b, c;
        region_index((int)(c0 + t), B.bs[1], B.bs[2], a, b, c);
        const int i = a + NG, j = b + NG, z = c + NG;
        const int64_t cell = ((int64_t)i * B.D[1] + j) * B.D[2] + z;
        const double *pc = p + NPR * B.off + cell;
        double pr[NPR], ucon[4];
#pragma unroll
        for (int k = 0; k < NPR; k++) pr[k] = pc[(int64_t)k * B.ks];
        const Geom<double> G = load_geom<double>(P, B, CENT, (int64_t)i * B.D[1] + j);
        ucon_calc(pr, G, ucon);
        float *r = s + t * DREC;
        r[0] = (float)pr[RHO];
        r[1] = (float)pr[UU];
#pragma unroll
        for (int k = 0; k < 4; k++) r[2 + k] = (float)ucon[k];
        r[6] = (float)pr[B1];
        r[7] = (float)pr[B2];
        r[8] = (float)pr[B3];
    }
    __syncthreads();
    float *o = out + (doff[item] - base) + c0 * DREC;
    for (int e = t; e < nv * DREC; e += RPACK_T) o[e] = s[e];
}

void launch_dump_pack(const KParams &P, const int *blist, const int64_t *doff, const int64_t *doff_host, int first,
                      int count, int64_t max_cells, const double *p, float *out, hipStream_t s)
{
    const unsigned gx = (unsigned)((max_cells + RPACK_T - 1) / RPACK_T);
    for (int f = first; f < first + count; f += 65535) {
        const int cnt = min(65535, first + count - f);
        hipLaunchKernelGGL(k_dump_pack, dim3(gx, (unsigned)cnt), dim3(RPACK_T), 0, s, P, blist, doff, f,
                           doff_host[first], p, out);
    }
}

// Items [first, first + count) of (blist, roff); a grid's y extent is capped
// at 65535, so long lists go in several launches.
void launch_rdump_pack(const KParams &P, const int *blist, const int64_t *roff, const int64_t *roff_host, int first,
                       int count, int64_t max_cells, const double *p, const double *ps, double *out, hipStream_t s)
{
    const unsigned gx = (unsigned)((max_cells + RPACK_T - 1) / RPACK_T);
    for (int f = first; f < first + count; f += 65535) {
        const int cnt = min(65535, first + count - f);
        hipLaunchKernelGGL(k_rdump_pack, dim3(gx, (unsigned)cnt), dim3(RPACK_T), 0, s, P, blist, roff, f,
                           roff_host[first], p, ps, out);
    }
}

void launch_rdump_unpack(const KParams &P, const int *blist, const int64_t *roff, const int64_t *roff_host, int first,
                         int count, int64_t max_cells, const RdumpRead &R, const double *rec, double *p, double *ps,
                         int max_e3, int64_t max_rows, hipStream_t s)
{
    if (R.red[0] == 1 && R.red[1] == 1 && R.red[2] == 1 && max_e3 <= RPACK_T) {
        // the staged form: max_rows z-rows per block at most, RPACK_T / e3 rows per workgroup
        const unsigned gx = (unsigned)((max_rows + (RPACK_T / max_e3) - 1) / (RPACK_T / max_e3));
        for (int f = first; f < first + count; f += 65535) {
            const int cnt = min(65535, first + count - f);
            hipLaunchKernelGGL(k_rdump_unpack1, dim3(gx, (unsigned)cnt), dim3(RPACK_T), 0, s, P, blist, roff, f,
                               roff_host[first], R, rec, p, ps);
        }
        return;
    }
    const int T = 256;
    const unsigned gx = (unsigned)((max_cells + T - 1) / T);
    for (int f = first; f < first + count; f += 65535) {
        const int cnt = min(65535, first + count - f);
        hipLaunchKernelGGL(k_rdump_unpack, dim3(gx, (unsigned)cnt), dim3(T), 0, s, P, blist, roff, f,
                           roff_host[first], R, rec, p, ps);
    }
}

}  // namespace hamr
