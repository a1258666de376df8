/*
 * fused.hip -- the fused paths of one disc: one kernel per RK stage over
 * (phi, z) tile columns marching in r (the work item is in fused_tile.h),
 * the r-boundary faces, the ghost-cell pass, Setdt from a stored state,
 * conversions, totals, and the host launchers behind fused.h.
 */
#include "fused.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

/* This file is compiled twice: FUSED_PREC=64 (FP64 state, -ffp-contract=off
 * like the plain path) and FUSED_PREC=32 (FP32 state; contraction into FMAs
 * within each expression and the fast FP32 divide/sqrt are allowed there,
 * its bar is a relative L1 error, not bit parity).  Each build defines the
 * entry points of its own namespace, cyl::f64 or cyl::f32. */
#if FUSED_PREC == 32
#define FUSED_NS f32
#define FUSED_FT float
#elif FUSED_PREC == 64
#define FUSED_NS f64
#define FUSED_FT double
#else
#error "build with -DFUSED_PREC=32 or -DFUSED_PREC=64"
#endif

#include "fused_tile.h"

namespace cyl {
namespace FUSED_NS {

/* ---- single-disc kernels: one work item per block --------------------- */
template <typename T, int LIM, bool HLL, bool ALT>
__global__ __launch_bounds__((tile_threads<T, 0>()), (tile_waves<T, 0>())) void k_fused(FArgs<T> a)
{
    const int b = remap_block(blockIdx.x, gridDim.x);
    if (b >= a.nblocks) {                         /* surplus block: neutral Setdt partial */
        if (a.want_dt && threadIdx.x == 0) { a.dtpart[2 * blockIdx.x] = 0.0; a.dtpart[2 * blockIdx.x + 1] = 9.3e18; }
        return;
    }
    fused_tile<T, LIM, HLL, 0, ALT>(a, b % a.ntiles, b / a.ntiles, blockIdx.x);
}

/* ---- the split stage (FP32 state): r-march, then the per-cell update ---- */
/* waves per SIMD the r-march is built for: FP32 three 256-thread blocks
 * per CU (51 KB of LDS each, <= 168 VGPRs), FP64 two waves (256 VGPRs) */
template <typename T> constexpr int march_waves() { return sizeof(T) == 4 ? 3 : 2; }

/* FP32: the z-side ghosts in packed planes (FusedState::zg) */
template <typename T, int LIM, bool HLL>
__global__ __launch_bounds__((tile_threads<T, 0>()), (march_waves<T>())) void k_march(FArgs<T> a)
{
    const int b = remap_block(blockIdx.x, gridDim.x);
    if (b >= a.nblocks) return;
    march_tile<T, LIM, HLL, 0, sizeof(T) == 4>(a, b % a.ntiles, b / a.ntiles);
}

/* rows j a wavefront of the update advances, their loads issued together
 * and the k and (i, k) table rows loaded once for them: two in FP32 (three
 * waves per SIMD at 158 VGPRs, 6 % faster than one row at four waves); one
 * in FP64, whose update already streams at the device-copy rate and would
 * drop to two waves per SIMD */
template <typename T> constexpr int update_rows() { return sizeof(T) == 4 ? 2 : 1; }

template <typename T, bool ALT>
__global__ __launch_bounds__(256) void k_update(FArgs<T> a, int nch)
{
    update_cell<T, ALT, false, update_rows<T>()>(a, nch, blockIdx.x, blockIdx.x);
}

/* blocks of the split stage's update: four wavefronts, each 64 k of
 * update_rows rows */
template <typename T>
static int update_blocks(const DevGeom &g, int &nch)
{
    nch = (g.NZ - 2 + 63) / 64;
    constexpr int RJ = update_rows<T>();
    const long long waves = (long long)(g.NX - 2) * ((g.NY - 2 + RJ - 1) / RJ) * nch;
    return (int)((waves + 3) / 4);
}

template <typename T, int LIM, bool HLL>
__global__ __launch_bounds__(256) void k_edge(FArgs<T> a, T *edge)
{
    edge_cell<T, LIM, HLL>(a, edge, blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T, bool ZGL>
__global__ __launch_bounds__(256) void k_halo(FArgs<T> a)
{
    halo_range<T, ZGL>(a, blockIdx.x * (size_t)blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

/* ---- RK2 only: interior plane k = 1 gets U3 = 0 after the update (the
 * fused kernel does it for its cells; nothing else to do) ---------------- */

/* ---- Setdt from the stored state (first step, or after an upload) ------- */
template <typename T>
__global__ __launch_bounds__(256) void k_setdt_state(FArgs<T> a)
{
    const DevGeom &g = a.g;
    const size_t ni = (size_t)(g.NX - 12), nj = (size_t)(g.NY - 2), nk = (size_t)(g.NZ - 2);
    const size_t M = ni * nj * nk;
    double best = 0.0;
    long long bix = LLONG_MAX;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < M; q += (size_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(q / (nj * nk)), j = 1 + (int)((q / nk) % nj), k = 1 + (int)(q % nk);
        const size_t cc = C3(i, j, k), cs = SI(i, j, k);
        T u[9], v[9], ub[NOFF];
#pragma unroll
        for (int m = 0; m < 9; m++) u[m] = a.Un[m * a.NS + cs];
        ub_load(a, i, k, ub);
        Prec<T>::utov(a.c, u, ub, g.oneox[i], v);
        const double V0 = v[0], V1 = v[1], V2 = v[2], V3 = v[3], V4 = v[4], V5 = v[5], V6 = v[6], V7 = v[7];
        const double v2 = V1 * V1 + V2 * V2 + V3 * V3;
        const double b2 = V4 * V4 + V5 * V5 + V6 * V6;
        const double t1 = sqrt(v2 + (a.c.gm * V7 + b2) / V0) / g.minlength[(size_t)i * g.NZ + k];
        const double tc = (2.0 * ((V0 * V0 * cooling(a.c, V0, V7)) / (1.5 * V7)));
        const double tt = rmax(t1, tc);
        if ((tt > best || (tt == best && (long long)cc < bix)) && tt > 0.0) { best = tt; bix = (long long)cc; }
    }
    __shared__ double sv[4];
    __shared__ long long sx[4];
    block_dt_max<4>(best, bix, sv, sx);
    if (threadIdx.x == 0) { a.dtpart[2 * blockIdx.x] = best; a.dtpart[2 * blockIdx.x + 1] = (double)bix; }
}

/* the Setdt partials (up to one per block of the last stage's update
 * kernel: a hundred thousand on the split stage) -> dt and its cell, in
 * two launches where there are many: block b reduces the partials of
 * slice b (of `per` each) into partial b of out (same encoding: value,
 * index, an index >= 9e18 for none), then one block reduces those into
 * (dt-rate, cell, or -1 for none).  Four partials per thread and round,
 * loaded before any is compared, so that the loads are in flight
 * together. */
constexpr int kDtSlices = 256;

__global__ __launch_bounds__(1024) void k_dt_final(const double *__restrict__ part, int nb, int per, bool last,
                                                   double *__restrict__ out)
{
    __shared__ double sv[16];
    __shared__ long long sx[16];
    const int t = threadIdx.x;
    const int lo = blockIdx.x * per, hi = min(nb, lo + per);
    double best = 0.0;
    long long bix = LLONG_MAX;
    for (int b0 = lo + t; b0 < hi; b0 += 4 * 1024) {
        double2 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int b = b0 + u * 1024;
            q[u] = b < hi ? ((const double2 *)part)[b] : make_double2(0.0, 9.3e18);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double v = q[u].x;
            const long long ix = q[u].y >= 9.0e18 ? LLONG_MAX : (long long)q[u].y;
            if (dt_better(v, ix, best, bix) && v > 0.0) { best = v; bix = ix; }
        }
    }
    block_dt_max<16>(best, bix, sv, sx);
    if (t == 0) {
        out[2 * blockIdx.x] = best;
        out[2 * blockIdx.x + 1] = last ? (bix == LLONG_MAX ? -1.0 : (double)bix) : (double)bix;
    }
}

/* ---- conversions and diagnostics ---------------------------------------- */
/* reference layout [9][NX][NY][NZ] FP64 <-> padded state layout (SI) */
__device__ __forceinline__ size_t pad_index(int nz, int NZP, size_t r)
{
    const size_t NZ = nz, row = r / NZ;
    return row * (size_t)NZP + (r - row * NZ) + KOFF;
}

/* element (m, cell r of the reference layout) of a state buffer: the padded
 * body, or with the batch layout (zg) for the z-side ghosts the packed
 * planes after it */
__device__ __forceinline__ size_t state_index(const DevGeom &g, int NZP, size_t NS, bool zg, size_t m, size_t r)
{
    const size_t NZ = g.NZ, row = r / NZ, k = r - row * NZ;
    if (zg && (k == 0 || k == NZ - 1)) return 9 * NS + ((k != 0 ? (size_t)g.NX * g.NY : 0) + row) * 9 + m;
    return m * NS + row * (size_t)NZP + k + KOFF;
}

/* offset of component m at cell r of the reference layout (FP32 state:
 * the phi-mean of the initial state for the components with a slot; FP64:
 * none) */
template <typename T>
__device__ __forceinline__ double ub_of(const DevGeom &g, const T *ikg, size_t m, size_t r)
{
    if constexpr (sizeof(T) == 4) {
        const int sl = oslot((int)m);
        if (sl >= 0) {
            const size_t NYZ = (size_t)g.NY * g.NZ, i = r / NYZ, k = r % g.NZ;
            return (double)ikg[(i * g.NZ + k) * 8 + 4 + sl];
        }
    }
    return 0.0;
}

template <typename T>
__global__ void k_to_state(DevGeom g, int NZP, size_t NS, bool zg, const T *__restrict__ ikg, const double *__restrict__ src,
                           T *__restrict__ dst, size_t N)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < 9 * N; q += (size_t)gridDim.x * blockDim.x) {
        const size_t m = q / N, r = q - m * N;
        dst[state_index(g, NZP, NS, zg, m, r)] = T(src[q] - ub_of(g, ikg, m, r));
    }
}

/* offsets ikg[i][k][4 + slot] = phi-mean over the interior j of
 * Uinitial[m][i][j][k] for m = 0, 2, 5, 7, rounded to T (FP32 state only) */
template <typename T>
__global__ void k_ubar(DevGeom g, const double *__restrict__ Uinit, T *__restrict__ ikg, size_t N)
{
    const size_t nik = (size_t)g.NX * g.NZ;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < NOFF * nik; q += (size_t)gridDim.x * blockDim.x) {
        const size_t sl = q / nik, ik = q - sl * nik, i = ik / g.NZ, k = ik - i * g.NZ;
        const size_t m = sl == 0 ? 0 : (sl == 1 ? 2 : (sl == 2 ? 5 : 7));
        double s = 0.0;
        for (int j = 1; j <= g.NY - 2; j++) s += Uinit[m * N + (i * g.NY + j) * g.NZ + k];
        ikg[ik * 8 + 4 + sl] = T(s / (double)(g.NY - 2));
    }
}

template <typename T>
__global__ void k_rhofloor(int NZ, int NZP, const double *__restrict__ Uinit0, T *__restrict__ rf, size_t N, double rhocr)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < N; q += (size_t)gridDim.x * blockDim.x)
        rf[pad_index(NZ, NZP, q)] = T(Uinit0[q] * rhocr);
}

template <typename T>
__global__ void k_export(DevGeom g, int NZP, size_t NS, bool zg, Consts c, const T *__restrict__ ikg,
                         const T *__restrict__ U, double *__restrict__ U64, double *__restrict__ V64, size_t N)
{
    const size_t NYZ = (size_t)g.NY * g.NZ;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < N; q += (size_t)gridDim.x * blockDim.x) {
        double u[9], v[9];
#pragma unroll
        for (int m = 0; m < 9; m++) {
            u[m] = (double)U[state_index(g, NZP, NS, zg, m, q)] + ub_of(g, ikg, m, q);
            U64[m * N + q] = u[m];
        }
        utov<double>(c, u, g.oneox[q / NYZ], v);
#pragma unroll
        for (int m = 0; m < 9; m++) V64[m * N + q] = v[m];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_diag_t(DevGeom g, int NZP, bool zg, const T *__restrict__ ikg, const T *__restrict__ U,
                                                size_t N, double *__restrict__ red)
{
    /* N is the component stride of the padded state arrays */
    const size_t ni = g.NX - 2, nj = g.NY - 2, nk = g.NZ - 2, M = ni * nj * nk;
    const size_t si = (size_t)g.NY * NZP, sj = NZP;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < M; q += (size_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(q / (nj * nk)), j = 1 + (int)((q / nk) % nj), k = 1 + (int)(q % nk);
        const size_t cc = (((size_t)i) * g.NY + (size_t)j) * (size_t)NZP + (size_t)k + KOFF, ik = (size_t)i * g.NZ + k;
        const double dv = g.dvol[ik];
        /* full values: the stored deviation + the offset of its (i, k) */
        auto ob = [&](int ii, int kk, int m) -> double {
            if constexpr (sizeof(T) == 4) {
                if (oslot(m) >= 0) return (double)ikg[((size_t)ii * g.NZ + kk) * 8 + 4 + oslot(m)];
            }
            return 0.0;
        };
        const double b4 = (double)U[4 * N + cc] + ob(i, k, 4), b5 = (double)U[5 * N + cc] + ob(i, k, 5);
        const double b6 = (double)U[6 * N + cc] + ob(i, k, 6);
        const double brp = 0.5 * (b4 + ((double)U[4 * N + cc + si] + ob(i + 1, k, 4)));
        const double brm = 0.5 * (b4 + ((double)U[4 * N + cc - si] + ob(i - 1, k, 4)));
        const double bpp = 0.5 * (b5 + ((double)U[5 * N + cc + sj] + ob(i, k, 5)));
        const double bpm = 0.5 * (b5 + ((double)U[5 * N + cc - sj] + ob(i, k, 5)));
        /* the z neighbours k = 0 and NZ-1 are ghosts: packed planes with zg */
        const size_t zp = zg && k == g.NZ - 2 ? 9 * N + (((size_t)g.NX + i) * g.NY + j) * 9 + 6 : 6 * N + cc + 1;
        const size_t zm = zg && k == 1 ? 9 * N + ((size_t)i * g.NY + j) * 9 + 6 : 6 * N + cc - 1;
        const double bzp = 0.5 * (b6 + ((double)U[zp] + ob(i, k + 1, 6))), bzm = 0.5 * (b6 + ((double)U[zm] + ob(i, k - 1, 6)));
        const double db = (g.g2[ik] * brp - g.g1[ik] * brm) + (g.g3[ik] * bpp - g.g3[ik] * bpm) + (g.g5[ik] * bzp - g.g5[ik] * bzm);
        acc[0] += ((double)U[cc] + ob(i, k, 0)) * dv;
        acc[1] += ((double)U[7 * N + cc] + ob(i, k, 7)) * dv;
        acc[2] += fabs(db) * dv;
        acc[3] += fabs((double)U[8 * N + cc]) * dv;
    }
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    for (int q = 0; q < 4; q++) sh[q][t] = acc[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < 4; q++) sh[q][t] += sh[q][t + off];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < 4; q++) red[4 * blockIdx.x + q] = sh[q][0];
}

__global__ __launch_bounds__(256) void k_diag_final(double *__restrict__ red, int nb)
{
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = t; b < nb; b += 256)
        for (int q = 0; q < 4; q++) acc[q] += red[4 * b + q];
    for (int q = 0; q < 4; q++) sh[q][t] = acc[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < 4; q++) sh[q][t] += sh[q][t + off];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < 4; q++) red[q] = sh[q][0];
}

/* ======================================================================== */
static unsigned nblocks(size_t n, unsigned bs, unsigned cap)
{
    size_t b = (n + bs - 1) / bs;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, cap));
}

int alloc(FusedState &fs, const DevGeom &g, const DevState &d, AllocFn alloc)
{
    using T = FT;
    fs.fp32 = sizeof(T) == 4;
    fs.esize = sizeof(T);
    const size_t N = d.N;
    /* padded state layout: rows of NZP >= NZ + KOFF cells, NZP a multiple
     * of 32, so that k = 1 of every row starts a 128-byte line */
    fs.NZ = g.NZ;
    fs.NZP = (g.NZ + KOFF + 31) / 32 * 32;
    fs.NS = (size_t)g.NX * g.NY * fs.NZP;
    const size_t NS = fs.NS;
    int rc;
    if (fs.batch) {
        fs.zg = true;                        /* a batch's discs are advanced by the batch */
    } else {
        const char *e = getenv("CYLMHD_STAGE_KERNELS");
        fs.split = !(e && e[0] == '1');
        /* the FP32 two-kernel stage reads and writes packed z-side ghosts
         * (its ghost pass then writes 36 contiguous bytes per cell, not
         * nine 64-byte lines); in the FP64 r-march the extra address
         * selection spills registers */
        fs.zg = fs.split && fs.fp32;
    }
    fs.nstate = 9 * NS + (fs.zg ? 18 * (size_t)g.NX * g.NY : 0);
    if ((rc = alloc(fs.nstate * fs.esize, &fs.Un)) || (rc = alloc(fs.nstate * fs.esize, &fs.A)) ||
        (rc = alloc(fs.nstate * fs.esize, &fs.B)) || (rc = alloc((9 * NS + 27 * (size_t)g.NX * g.NY) * fs.esize, &fs.shadow)) ||
        (rc = alloc(NS * fs.esize, &fs.rhofloor)) || (rc = alloc(2 * 9 * (size_t)g.NY * g.NZ * fs.esize, &fs.edge)))
        return rc;
    void *p;
    if ((rc = alloc(9 * N * sizeof(double), &p))) return rc;
    fs.u64_scratch = (double *)p;
    /* the divergence of the two-kernel stage (a batch's discs may run it) */
    if ((fs.split || fs.batch) && (rc = alloc(9 * NS * fs.esize, &fs.div))) return rc;
    const int TJ = Tile<T, 0>::TJ;
    const int TK = Tile<T, 0>::TK;
    fs.tiles_j = (g.NY - 2 + TJ - 1) / TJ;
    fs.tiles_k = (g.NZ - 2 + TK - 1) / TK;
    const int cols = fs.tiles_j * fs.tiles_k;
    /* Segments of the r-march: the grid is dealt to the CUs in rounds of
     * (CUs x resident blocks per CU); pick the count that wastes least in
     * the last round, charging each segment its prologue (~2.5 planes), with
     * segments of at least 8 planes. */
    int ncu = 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || ncu < 1) ncu = 256;
    int per_cu = 0;                  /* resident stage blocks per CU (registers and LDS) */
    if (fs.split)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_march<T, 1, false>, tile_threads<T, 0>(), 0);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused<T, 1, false, false>, tile_threads<T, 0>(), 0);
    const double slots = (double)std::max(per_cu, 1) * ncu;
    const int planes = g.NX - 2;
    int nseg = 1;
    double bestc = 1e300;
    for (int ns = 1; ns <= std::max(1, std::min(32, planes / 8)); ns++) {
        const int sl = (planes + ns - 1) / ns;
        const int nsr = (planes + sl - 1) / sl;
        const double rounds = (double)cols * nsr / slots;
        const double cost = std::ceil(rounds) / rounds * (sl + 2.5) / sl * (rounds < 1.0 ? 1.0 / rounds : 1.0);
        if (cost < bestc - 1e-9) { bestc = cost; nseg = ns; }
    }
    fs.seglen = (planes + nseg - 1) / nseg;
    fs.nseg = (planes + fs.seglen - 1) / fs.seglen;
    /* The model above assumes the blocks run in lockstep rounds; they do
     * not quite, and neighbouring counts differ by a few per cent.  The
     * first stage-1 launch times the model's count and these candidates
     * and keeps the fastest (the segmentation does not change the results:
     * a segment's prologue recomputes the carried r-states and flux with
     * the same arithmetic). */
    fs.ncand = 0;
    fs.cand[fs.ncand++] = fs.nseg;
    int maxseg = fs.nseg;
    for (int ns : {4, 5, 6, 8, 12, 16, 24}) {
        if (fs.ncand >= 8 || planes / ns < 8) continue;
        const int sl = (planes + ns - 1) / ns, nsr = (planes + sl - 1) / sl;
        bool dup = false;
        for (int q = 0; q < fs.ncand; q++) dup = dup || fs.cand[q] == nsr;
        if (!dup) { fs.cand[fs.ncand++] = nsr; maxseg = std::max(maxseg, nsr); }
    }
    fs.tuned = fs.ncand < 2;
    fs.dt_blocks = std::max((cols * maxseg + 7) / 8 * 8, 1024);
    if (fs.split) {
        int nch;
        fs.dt_blocks = std::max(fs.dt_blocks, update_blocks<T>(g, nch));
    }
    fs.dt_nb = (cols * fs.nseg + 7) / 8 * 8;
    if ((rc = alloc(2 * (size_t)fs.dt_blocks * sizeof(double), &p))) return rc;
    fs.dtpart = (double *)p;
    if ((rc = alloc(2 * (size_t)kDtSlices * sizeof(double), &p))) return rc;
    fs.dtpart2 = (double *)p;
    const size_t nik = (size_t)g.NX * g.NZ;
    if ((rc = alloc((size_t)g.NX * IT * sizeof(double), &p))) return rc;
    fs.itab = (double *)p;
    if ((rc = alloc((size_t)g.NZ * KT * sizeof(double), &p))) return rc;
    fs.ktab = (double *)p;
    if ((rc = alloc(4 * nik * sizeof(double), &p))) return rc;
    fs.ikd = (double *)p;
    /* g-factors and (FP32 state) the offsets, zero until the initial state
     * is known: a state uploaded before cylmhd_set_initial is stored as it is */
    if ((rc = alloc(ikg_w<T>() * nik * sizeof(T), &fs.ikg))) return rc;
    hipLaunchKernelGGL(k_pack<T>, dim3(nblocks(nik, 256, 4096)), dim3(256), 0, 0, g, fs.itab, (T *)fs.ikg, fs.ikd, fs.ktab);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return CYLMHD_EHIP;
    return 0;
}

template <typename T>
static FArgs<T> make_args(FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c)
{
    FArgs<T> a{};
    a.g = g; a.c = c;
    a.N = d.N;
    a.NS = fs.NS; a.NZP = fs.NZP;
    a.rhofloor = (const T *)fs.rhofloor;
    a.Uinit = d.Uinit; a.Vinit_b = d.Vinit_z;
    a.edge = (const T *)fs.edge;
    a.counts = d.counts;
    a.dtpart = fs.dtpart;
    a.seglen = fs.seglen; a.tiles_j = fs.tiles_j; a.tiles_k = fs.tiles_k; a.ntiles = fs.tiles_j * fs.tiles_k;
    a.Un = (const T *)fs.Un;
    a.shadow = (T *)fs.shadow;
    a.itab = fs.itab; a.ikd = fs.ikd; a.ktab = fs.ktab; a.ikg = (const T *)fs.ikg;
    a.div = (T *)fs.div;
    return a;
}

int set_initial_tables(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, double rhocr)
{
    hipLaunchKernelGGL(k_rhofloor<FT>, dim3(nblocks(d.N, 256, 8192)), dim3(256), 0, s, fs.NZ, fs.NZP, d.Uinit, (FT *)fs.rhofloor, d.N, rhocr);
    if (fs.fp32) {
        const size_t n = NOFF * (size_t)g.NX * g.NZ;
        hipLaunchKernelGGL(k_ubar<FT>, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, g, d.Uinit, (FT *)fs.ikg, d.N);
    }
    return hipGetLastError() == hipSuccess ? 0 : CYLMHD_EHIP;
}

int load_state(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const double *U64)
{
    (void)g;
    const size_t n = 9 * d.N;
    hipLaunchKernelGGL(k_to_state<FT>, dim3(nblocks(n, 256, 16384)), dim3(256), 0, s, g, fs.NZP, fs.NS, fs.zg, (const FT *)fs.ikg, U64, (FT *)fs.Un, d.N);
    fs.dt_valid = false;
    return hipGetLastError() == hipSuccess ? 0 : CYLMHD_EHIP;
}

int export_state(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c, double *V64)
{
    hipLaunchKernelGGL(k_export<FT>, dim3(nblocks(d.N, 256, 16384)), dim3(256), 0, s, g, fs.NZP, fs.NS, fs.zg, c, (const FT *)fs.ikg, (const FT *)fs.Un, fs.u64_scratch, V64, d.N);
    return hipGetLastError() == hipSuccess ? 0 : CYLMHD_EHIP;
}

template <typename T>
static void setdt_t(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c)
{
    int nb = fs.dt_blocks;
    if (!fs.dt_valid) {
        FArgs<T> a = make_args<T>(fs, g, d, c);
        const size_t M = (size_t)(g.NX - 12) * (g.NY - 2) * (g.NZ - 2);
        nb = (int)nblocks(M, 256, (unsigned)std::min(fs.dt_blocks, 4096));
        hipLaunchKernelGGL(k_setdt_state<T>, dim3(nb), dim3(256), 0, s, a);
    } else {
        nb = fs.dt_nb;
    }
    if (nb > 4 * 1024) {
        const int per = (nb + kDtSlices - 1) / kDtSlices;
        const int ns = (nb + per - 1) / per;
        hipLaunchKernelGGL(k_dt_final, dim3(ns), dim3(1024), 0, s, fs.dtpart, nb, per, false, fs.dtpart2);
        hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1024), 0, s, fs.dtpart2, ns, ns, true, d.red);
    } else {
        hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1024), 0, s, fs.dtpart, nb, nb, true, d.red);
    }
}

int setdt(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c)
{
    setdt_t<FT>(s, fs, g, d, c);
    return hipGetLastError() == hipSuccess ? 0 : CYLMHD_EHIP;
}

/* Time the stage kernel of a first stage (its inputs are not overwritten:
 * Un -> A) at each candidate segment count and keep the fastest.  Runs once
 * per context, inside the first step; the extra launches' floor counts are
 * cleared afterwards. */
template <typename T, int LIM, bool HLL, bool ALT>
static void tune_segments(hipStream_t s, FusedState &fs, FArgs<T> a)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) { fs.tuned = true; return; }
    if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); fs.tuned = true; return; }
    const int planes = a.g.NX - 2;
    float best = 1e30f;
    int pick = fs.nseg;
    for (int q = 0; q < fs.ncand; q++) {
        const int ns = fs.cand[q];
        a.seglen = (planes + ns - 1) / ns;
        a.nblocks = fs.tiles_j * fs.tiles_k * ((planes + a.seglen - 1) / a.seglen);
        const int nb = (a.nblocks + 7) / 8 * 8;
        auto launch = [&]() {
            if (fs.split) hipLaunchKernelGGL((k_march<T, LIM, HLL>), dim3(nb), dim3(tile_threads<T, 0>()), 0, s, a);
            else hipLaunchKernelGGL((k_fused<T, LIM, HLL, ALT>), dim3(nb), dim3(tile_threads<T, 0>()), 0, s, a);
        };
        launch();
        hipEventRecord(e0, s);
        for (int r = 0; r < 2; r++) launch();
        hipEventRecord(e1, s);
        float ms = 0.0f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
            best = ms;
            pick = ns;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    fs.seglen = (planes + pick - 1) / pick;
    fs.nseg = (planes + fs.seglen - 1) / fs.seglen;
    fs.tuned = true;
    hipMemsetAsync(a.counts, 0, 3 * sizeof(int), s);
}

template <typename T, int LIM, bool HLL, bool ALT>
static void stage_t(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,
                    const StageArgs &sa, hipEvent_t ev0, hipEvent_t evm, hipEvent_t ev1)
{
    FArgs<T> a = make_args<T>(fs, g, d, c);
    const bool last = sa.rk3 ? sa.stage == 3 : sa.stage == 2;
    a.kind = sa.rk3 ? sa.stage : 3 + sa.stage;
    a.last = last;
    a.rk3 = sa.rk3;
    a.want_dt = last;
    a.counts = d.counts + 3 * (sa.stage - 1);
    /* buffers: RK3 Un -> A -> B -> Un ; RK2 Un -> A -> Un */
    const T *in = sa.stage == 1 ? (const T *)fs.Un : (sa.stage == 2 ? (const T *)fs.A : (const T *)fs.B);
    T *out = last ? (T *)fs.Un : (sa.stage == 1 ? (T *)fs.A : (T *)fs.B);
    a.Uin = in;
    a.Uout = out;
    a.nblocks = fs.tiles_j * fs.tiles_k * fs.nseg;
    const int ne = (g.NY - 2) * (g.NZ - 2);
    if (LIM == 2) a.Vinit_b = d.Vinit_y;          /* vanLeer: YVlrInitializeBoundary */
    hipLaunchKernelGGL((k_edge<T, LIM, HLL>), dim3((ne + 255) / 256), dim3(256), 0, s, a, (T *)fs.edge);
    if (!fs.tuned && sa.stage == 1) {
        tune_segments<T, LIM, HLL, ALT>(s, fs, a);
        a.seglen = fs.seglen;
        a.nblocks = fs.tiles_j * fs.tiles_k * fs.nseg;
    }
    const int nb = (a.nblocks + 7) / 8 * 8;
    if (ev0) hipEventRecord(ev0, s);
    if (fs.split) {
        int nch;
        const int nu = update_blocks<T>(g, nch);
        hipLaunchKernelGGL((k_march<T, LIM, HLL>), dim3(nb), dim3(tile_threads<T, 0>()), 0, s, a);
        if (evm) hipEventRecord(evm, s);
        hipLaunchKernelGGL((k_update<T, ALT>), dim3(nu), dim3(256), 0, s, a, nch);
        if (last) fs.dt_nb = nu;
    } else {
        hipLaunchKernelGGL((k_fused<T, LIM, HLL, ALT>), dim3(nb), dim3(tile_threads<T, 0>()), 0, s, a);
        if (last) fs.dt_nb = nb;
    }
    if (ev1) hipEventRecord(ev1, s);
    const size_t nghost = 2 * (size_t)g.NY * g.NZ + (size_t)(g.NX - 2) * 2 * g.NZ + (size_t)(g.NX - 2) * (g.NY - 2) * 2;
    if (fs.zg) hipLaunchKernelGGL((k_halo<T, true>), dim3(nblocks(nghost, 256, 4096)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_halo<T, false>), dim3(nblocks(nghost, 256, 4096)), dim3(256), 0, s, a);
    if (last) fs.dt_valid = true;
}

int stage(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,
          const StageArgs &a, hipEvent_t ev0, hipEvent_t evm, hipEvent_t ev1)
{
    if (fs.batch) return CYLMHD_EINVAL;   /* a batch's disc: the batch advances it */
    /* one kernel per limiter x solver; the default source (SpiralGravityCooling)
     * runs a kernel compiled for it alone */
#define CYL_STAGE(L, H, A) stage_t<FT, L, H, A>(s, fs, g, d, c, a, ev0, evm, ev1)
#define CYL_STAGE_LIM(H, A)                                  \
    switch (a.limiter) {                                      \
    case 0: CYL_STAGE(0, H, A); break;                        \
    case 1: CYL_STAGE(1, H, A); break;                        \
    case 2: CYL_STAGE(2, H, A); break;                        \
    default: CYL_STAGE(3, H, A); break;                       \
    }
    if (c.source == 0) {
        if (a.hll) { CYL_STAGE_LIM(true, false) } else { CYL_STAGE_LIM(false, false) }
    } else {
        if (a.hll) { CYL_STAGE_LIM(true, true) } else { CYL_STAGE_LIM(false, true) }
    }
#undef CYL_STAGE_LIM
#undef CYL_STAGE
    return hipGetLastError() == hipSuccess ? 0 : CYLMHD_EHIP;
}

int diag(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d)
{
    const size_t M = (size_t)(g.NX - 2) * (g.NY - 2) * (g.NZ - 2);
    const int nb = (int)nblocks(M, 256, (unsigned)(d.nred / 4));
    hipLaunchKernelGGL(k_diag_t<FT>, dim3(nb), dim3(256), 0, s, g, fs.NZP, fs.zg, (const FT *)fs.ikg, (const FT *)fs.Un, fs.NS, d.red);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, s, d.red, nb);
    return hipGetLastError() == hipSuccess ? 0 : CYLMHD_EHIP;
}

}  // namespace FUSED_NS
}  // namespace cyl
