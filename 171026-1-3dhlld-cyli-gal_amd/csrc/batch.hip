/*
 * batch.hip -- many independent discs advanced together (FP32 state).
 *
 * Every disc keeps its own context (tables, state, Uinitial, counters), so
 * set-up, dumps and restarts go through the single-disc ABI.  The batch adds:
 *   - a work list of (disc, tile, r-segment) items over all discs, processed
 *     by ONE launch per RK stage and tile shape whose blocks loop over it
 *     (XCD-aware order within each round of gridDim items): either the
 *     whole stage per item (one kernel), or, as on a single disc, the
 *     r-march per item (limiter states, faces, each cell's flux divergence)
 *     followed by ONE per-cell update launch over every disc's cells (two
 *     kernels); the autotuner picks the form;
 *   - per-disc Setdt on the device: the last stage leaves one partial per
 *     work item (one kernel) or per update block (two kernels); k_dt_batch
 *     reduces each disc's range, forms dt = cfl/max,
 *     ch = minminlength * max, psi damping and the spiral amplitude at the
 *     disc's own time, and writes them into that disc's stage arguments;
 *     a disc whose dt falls below dtmin (or is not finite) stops, as main.c
 *     would stop it (main.c:331-334), while the others go on;
 *   - so a whole step needs nothing from the host: it is captured once into
 *     a hipGraph and replayed;
 *   - an autotuner over the stage form (one or two kernels), the tile (its
 *     phi x z extents, waves per tile and register budget), the r-segment
 *     length and the grid of the stage launch, first for all discs together
 *     and then per class of disc size (discs of one phi x z size share a
 *     tile; one stage launch per tile in use), timed with HIP events on the
 *     batch stream.
 */
#define FUSED_NS b32
#define FUSED_FT float
#include "fused_tile.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "batch.h"

namespace cyl {
namespace b32 {

/* per-disc step state, device resident */
struct BDisc {
    double t, dt, ch;
    int alive, steps;
    double minminlength, cfl, dtmin, cr, epssp0, tausp, domega;
    const double *logxs;
    double *sincos;
    float *edge;
    int dt_lo, dt_hi;          /* the disc's Setdt partials: its work items or update blocks */
};

__device__ __forceinline__ int find_disc(const int *__restrict__ off, int n, int b)
{
    int lo = 0, hi = n;                       /* off[lo] <= b < off[hi] */
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (off[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* the work items [lo, hi) of one tile shape, the whole stage per item */
template <int LIM, bool HLL, int S, bool ALT>
__global__ __launch_bounds__((tile_threads<float, S>()), (tile_waves<float, S>())) void k_fused_batch(
    const FArgs<float> *__restrict__ descs, const int4 *__restrict__ items, int lo, int hi, const BDisc *__restrict__ bd)
{
    const int b = remap_block(blockIdx.x, gridDim.x);
    for (int it = lo + b; it < hi; it += gridDim.x) {
        const int4 w = items[it];
        if (!bd[w.x].alive) continue;
        fused_tile<float, LIM, HLL, S, ALT, true>(descs[w.x], w.y, w.z, it);
    }
}

/* the r-march of the work items [lo, hi) of one tile shape (three waves
 * per SIMD) */
template <int LIM, bool HLL, int S>
__global__ __launch_bounds__((tile_threads<float, S>()), 3) void k_march_batch(
    const FArgs<float> *__restrict__ descs, const int4 *__restrict__ items, int lo, int hi, const BDisc *__restrict__ bd)
{
    const int b = remap_block(blockIdx.x, gridDim.x);
    for (int it = lo + b; it < hi; it += gridDim.x) {
        const int4 w = items[it];
        if (!bd[w.x].alive) continue;
        march_tile<float, LIM, HLL, S, true>(descs[w.x], w.y, w.z);
        /* the next item's table rows go where this one's last divergence
         * reads them */
        lds_barrier();
    }
}

/* the per-cell update of every disc, each plane's cells dealt flat (k
 * fastest, 64 to a wavefront, whatever the disc's NZ; four wavefronts to a
 * block): disc d's blocks are off[d] .. off[d+1]-1, block b of the batch
 * writes Setdt partial b */
template <bool ALT>
__global__ __launch_bounds__(256) void k_update_batch(const FArgs<float> *__restrict__ descs, const int *__restrict__ off,
                                                      int n, const BDisc *__restrict__ bd)
{
    const int d = find_disc(off, n, blockIdx.x);
    if (!bd[d].alive) return;
    const FArgs<float> &a = descs[d];
    update_cell<float, ALT, true>(a, 0, blockIdx.x - off[d], blockIdx.x);
}

/* update blocks of one disc: each plane's interior cells in wavefronts of
 * 64, four wavefronts to a block */
inline int update_blocks_of(int NX, int NY, int NZ)
{
    const long long waves = (long long)(NX - 2) * (((long long)(NY - 2) * (NZ - 2) + 63) / 64);
    return (int)((waves + 3) / 4);
}

/* tile shapes compiled for a limiter x solver: all four for minmod and MC
 * with HLLD, shapes 0 and 1 for the others */
constexpr int kShapes = 4;
constexpr bool shape_built(int lim, bool hll, int s) { return s < 2 || ((lim == 0 || lim == 1) && !hll); }
constexpr int shape_tj(int s) { return s == 1 || s == 2 ? 6 : (s == 3 ? 10 : 14); }
constexpr int shape_tk(int s) { return s == 1 ? 32 : 16; }

template <int LIM, bool HLL>
__global__ __launch_bounds__(256) void k_edge_batch(const FArgs<float> *__restrict__ descs, const int *__restrict__ off,
                                                    int n, const BDisc *__restrict__ bd)
{
    const int d = find_disc(off, n, blockIdx.x);
    if (!bd[d].alive) return;
    edge_cell<float, LIM, HLL>(descs[d], bd[d].edge, (blockIdx.x - off[d]) * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_halo_batch(const FArgs<float> *__restrict__ descs, const int *__restrict__ off,
                                                    int n, const BDisc *__restrict__ bd)
{
    const int d = find_disc(off, n, blockIdx.x);
    if (!bd[d].alive) return;
    const size_t nb = off[d + 1] - off[d];
    halo_range<float, true>(descs[d], (size_t)(blockIdx.x - off[d]) * 256 + threadIdx.x, nb * 256);
}

/* Setdt partials of a stored state (before the first step): the blocks of
 * disc d are its partial slots dt_lo..dt_hi-1 */
__global__ __launch_bounds__(256) void k_setdt_batch(const FArgs<float> *__restrict__ descs, const BDisc *__restrict__ bd,
                                                     const int *__restrict__ off, int n, double *__restrict__ dtpart)
{
    const int d = find_disc(off, n, blockIdx.x);
    const FArgs<float> &a = descs[d];
    const DevGeom &g = a.g;
    const size_t ni = (size_t)(g.NX - 12), nj = (size_t)(g.NY - 2), nk = (size_t)(g.NZ - 2);
    const size_t M = ni * nj * nk;
    const size_t nb = off[d + 1] - off[d];
    double best = 0.0;
    long long bix = LLONG_MAX;
    for (size_t q = (blockIdx.x - off[d]) * (size_t)256 + threadIdx.x; q < M; q += nb * 256) {
        const int i = 1 + (int)(q / (nj * nk)), j = 1 + (int)((q / nk) % nj), k = 1 + (int)(q % nk);
        const size_t cc = C3(i, j, k), cs = SI(i, j, k);
        float u[9], v[9], ub[NOFF];
#pragma unroll
        for (int m = 0; m < 9; m++) u[m] = a.Un[m * a.NS + cs];
        ub_load(a, i, k, ub);
        Prec<float>::utov(a.c, u, ub, g.oneox[i], v);
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
    if (threadIdx.x == 0) {
        const int slot = bd[d].dt_lo + (int)(blockIdx.x - off[d]);
        dtpart[2 * slot] = best;
        dtpart[2 * slot + 1] = (double)bix;
    }
}

/* per-disc Setdt (setdt.h:24-94 reduced from the partials), TVDRK's psi
 * damping (tvdrungekutta.h:22) and the spiral amplitude at the disc's time
 * (spiralgravity.h:41-42), written into the disc's stage arguments */
__global__ __launch_bounds__(256) void k_dt_batch(FArgs<float> *__restrict__ descs, int nst, int n, BDisc *__restrict__ bd,
                                                  const double *__restrict__ dtpart)
{
    const int d = blockIdx.x;
    BDisc &D = bd[d];
    if (!D.alive) return;
    const int t = threadIdx.x;
    double best = 0.0;
    long long bix = LLONG_MAX;
    for (int q = D.dt_lo + t; q < D.dt_hi; q += 256) {
        const double v = dtpart[2 * q];
        const long long ix = dtpart[2 * q + 1] >= 9.0e18 ? LLONG_MAX : (long long)dtpart[2 * q + 1];
        if ((v > best || (v == best && ix < bix)) && v > 0.0) { best = v; bix = ix; }
    }
    __shared__ double sv[4];
    __shared__ long long sx[4];
    block_dt_max<4>(best, bix, sv, sx);
    if (t == 0) {
        const double dtmax = best;
        const double dt = D.cfl / dtmax, ch = D.minminlength * dtmax;
        if (!(dt >= D.dtmin) || isinf(dt)) {
            D.alive = 0;
            return;
        }
        D.dt = dt;
        D.ch = ch;
        const double psicoef = exp(-dt * ch * (1.0 / D.cr) * 0.5);
        const double sw = (D.t - D.tausp) / fabs(D.t - D.tausp);
        const double epssp = 0.5 * D.epssp0 * (D.t / D.tausp * (1.0 - sw) + (1.0 + sw));
        for (int s = 0; s < nst; s++) {
            FArgs<float> &a = descs[(size_t)s * n + d];
            a.c.dt = dt;
            a.c.ch = ch;
            a.c.psicoef = psicoef;
            a.g.epssp = epssp;
        }
    }
}

/* sin/cos of the arm phase at each disc's time (spiralgravity.h:63-65) */
__global__ __launch_bounds__(256) void k_sincos_batch(const FArgs<float> *__restrict__ descs, const int *__restrict__ off,
                                                      int n, const BDisc *__restrict__ bd)
{
    const int d = find_disc(off, n, blockIdx.x);
    const BDisc &D = bd[d];
    if (!D.alive) return;
    const DevGeom &g = descs[d].g;
    const size_t tot = (size_t)g.NX * g.NY, nb = off[d + 1] - off[d];
    for (size_t q = (blockIdx.x - off[d]) * (size_t)256 + threadIdx.x; q < tot; q += nb * 256) {
        const int i = (int)(q / g.NY), j = (int)(q % g.NY);
        const double phase = g.armnum * (g.y[j] - D.domega * D.t + g.cotpitch * D.logxs[i]);
        double sp, cp;
        sincos(phase, &sp, &cp);
        D.sincos[q] = sp;
        D.sincos[tot + q] = cp;
    }
}

__global__ void k_time_batch(BDisc *__restrict__ bd, int n)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n && bd[d].alive) {
        bd[d].t += bd[d].dt;
        bd[d].steps += 1;
    }
}

}  // namespace b32
}  // namespace cyl

using namespace cyl;
using namespace cyl::b32;

namespace {

#define BCHK(call)                                                                   \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            char m_[400];                                                            \
            snprintf(m_, sizeof m_, "%s: %s", #call, hipGetErrorString(e_));         \
            set_errmsg(m_);                                                          \
            return CYLMHD_EHIP;                                                      \
        }                                                                            \
    } while (0)

unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

struct cylmhd_batch {
    int n = 0;
    std::vector<cylmhd_ctx *> disc;
    std::vector<DiscView> view;
    hipStream_t stream = nullptr;
    int lim = 0;               /* CYLMHD_LIMITER_*, shared by every disc */
    bool alt = false;          /* some disc runs a source other than the default */
    bool hll = false, rk3 = true;
    int nst = 3;
    /* tuned launch shape: a tile shape per disc, the r-segment length and
     * the stage launch's grid; the work list holds the items of one shape
     * contiguously, groups[s] = [first, end) of shape s */
    std::vector<int> dshape;
    int seglen = 0, grid = 0;  /* grid 0: the blocks resident at once */
    int cus = 256;             /* compute units of the device */
    int groups[kShapes][2] = {};
    /* device side */
    FArgs<float> *d_descs = nullptr;
    int4 *d_items = nullptr;
    BDisc *d_bd = nullptr;
    int *d_eoff = nullptr, *d_hoff = nullptr, *d_soff = nullptr, *d_uoff = nullptr, *d_ioff = nullptr;
    double *d_dtpart = nullptr;
    int nitems = 0, eblocks = 0, hblocks = 0, sblocks = 0, ublocks = 0;
    bool split = true;         /* the stage as two kernels (r-march, update) */
    std::vector<void *> allocs;
    float *d_backup = nullptr;
    size_t backup_elems = 0;
    bool built = false, primed = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

int balloc(cylmhd_batch *b, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes + 256);
    if (e != hipSuccess) {
        set_errmsg("hipMalloc failed for the batch work list");
        return CYLMHD_ENOMEM;
    }
    b->allocs.push_back(*p);
    return CYLMHD_OK;
}

void free_graph(cylmhd_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_lists(cylmhd_batch *b)
{
    free_graph(b);
    for (void *p : b->allocs) hipFree(p);
    b->allocs.clear();
    b->d_descs = nullptr; b->d_items = nullptr; b->d_bd = nullptr; b->d_dtpart = nullptr;
    b->d_eoff = b->d_hoff = b->d_soff = b->d_uoff = b->d_ioff = nullptr;
    b->built = false;
}

/* tiles and r-segments of one disc for the chosen shape */
void disc_tiling(const DiscView &v, int shape, int seglen, int &tiles_j, int &tiles_k, int &seg, int &nseg)
{
    const int TJ = shape_tj(shape), TK = shape_tk(shape);
    tiles_j = (v.dg.NY - 2 + TJ - 1) / TJ;
    tiles_k = (v.dg.NZ - 2 + TK - 1) / TK;
    const int planes = v.dg.NX - 2;
    seg = seglen > 0 ? std::min(seglen, planes) : planes;
    nseg = (planes + seg - 1) / seg;
}

/* the per-disc step state (time, last dt, steps taken, alive) on the host */
int snapshot(cylmhd_batch *b, std::vector<BDisc> &out)
{
    out.resize(b->n);
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipMemcpy(out.data(), b->d_bd, sizeof(BDisc) * b->n, hipMemcpyDeviceToHost));
    return CYLMHD_OK;
}

/* work list, stage arguments and per-disc state for the current shape;
 * keep (optional): the step state each disc resumes from (else t = 0,
 * no steps, alive) */
int build(cylmhd_batch *b, const std::vector<BDisc> *keep)
{
    const bool keep_time = keep != nullptr;
    const std::vector<BDisc> &old = keep_time ? *keep : std::vector<BDisc>();
    free_lists(b);
    const int n = b->n, nst = b->nst;
    std::vector<int4> items;
    std::vector<FArgs<float>> descs((size_t)nst * n);
    std::vector<BDisc> bd(n);
    std::vector<int> eoff(n + 1, 0), hoff(n + 1, 0), soff(n + 1, 0), uoff(n + 1, 0), ioff(n + 1, 0);
    std::vector<int> ilo(n), ihi(n);
    /* the items of each shape together (discs of a shape in disc order) */
    for (int s = 0; s < kShapes; s++) {
        b->groups[s][0] = (int)items.size();
        for (int d = 0; d < n; d++) {
            if (b->dshape[d] != s) continue;
            int tj, tk, seg, nseg;
            disc_tiling(b->view[d], s, b->seglen, tj, tk, seg, nseg);
            ilo[d] = (int)items.size();
            for (int sgm = 0; sgm < nseg; sgm++)
                for (int tile = 0; tile < tj * tk; tile++) items.push_back(make_int4(d, tile, sgm, 0));
            ihi[d] = (int)items.size();
        }
        b->groups[s][1] = (int)items.size();
    }
    for (int d = 0; d < n; d++) {
        const DiscView &v = b->view[d];
        int tj, tk, seg, nseg;
        disc_tiling(v, b->dshape[d], b->seglen, tj, tk, seg, nseg);
        /* block -> disc maps: of the update, disc d's blocks are uoff[d] ..
         * uoff[d+1]-1; of k_setdt_batch, one block per Setdt partial of the
         * form in use: per update block (uoff) or per work item (ioff) */
        uoff[d + 1] = uoff[d] + update_blocks_of(v.dg.NX, v.dg.NY, v.dg.NZ);
        ioff[d + 1] = ioff[d] + (ihi[d] - ilo[d]);
        const Consts c = disc_consts(b->disc[d], 0.0, 0.0);
        for (int s = 0; s < nst; s++) {
            FArgs<float> a{};
            a.g = v.dg;
            a.c = c;
            a.N = v.ds.N;
            a.NS = v.fs->NS; a.NZP = v.fs->NZP;
            a.rhofloor = (const float *)v.fs->rhofloor;
            a.Uinit = v.ds.Uinit;
            a.Vinit_b = b->lim == CYLMHD_LIMITER_VANLEER ? v.ds.Vinit_y : v.ds.Vinit_z;
            a.edge = (const float *)v.fs->edge;
            a.dtpart = nullptr;            /* set below */
            a.seglen = seg; a.tiles_j = tj; a.tiles_k = tk; a.ntiles = tj * tk;
            a.nblocks = tj * tk * nseg;
            a.Un = (const float *)v.fs->Un;
            a.shadow = (float *)v.fs->shadow;
            a.div = (float *)v.fs->div;
            a.itab = v.fs->itab; a.ikd = v.fs->ikd; a.ktab = v.fs->ktab; a.ikg = (const float *)v.fs->ikg;
            const int stage = s + 1;
            const bool last = b->rk3 ? stage == 3 : stage == 2;
            a.kind = b->rk3 ? stage : 3 + stage;
            a.last = last;
            a.rk3 = b->rk3;
            a.want_dt = last;
            a.counts = v.ds.counts + 3 * s;
            a.Uin = stage == 1 ? (const float *)v.fs->Un : (stage == 2 ? (const float *)v.fs->A : (const float *)v.fs->B);
            a.Uout = last ? (float *)v.fs->Un : (stage == 1 ? (float *)v.fs->A : (float *)v.fs->B);
            descs[(size_t)s * n + d] = a;
        }
        BDisc &D = bd[d];
        D.t = keep_time ? old[d].t : 0.0;
        D.dt = keep_time ? old[d].dt : 0.0;
        D.ch = 0.0;
        D.alive = keep_time ? old[d].alive : 1;
        D.steps = keep_time ? old[d].steps : 0;
        D.minminlength = v.minminlength;
        D.cfl = v.p->cfl;
        D.dtmin = v.p->dtmin;
        D.cr = v.p->cr;
        D.epssp0 = v.p->epssp0;
        D.tausp = v.p->tausp;
        D.domega = v.p->omegapattern - v.p->omegasp;
        D.logxs = v.logxs;
        D.sincos = v.sincos;
        D.edge = (float *)v.fs->edge;
        eoff[d + 1] = eoff[d] + (int)cdiv((size_t)(v.dg.NY - 2) * (v.dg.NZ - 2), 256);
        const size_t nh = 2 * (size_t)v.dg.NY * v.dg.NZ + (size_t)(v.dg.NX - 2) * 2 * v.dg.NZ
                          + (size_t)(v.dg.NX - 2) * (v.dg.NY - 2) * 2;
        hoff[d + 1] = hoff[d] + (int)std::min<size_t>(cdiv(nh, 256), 512);
        soff[d + 1] = soff[d] + (int)std::min<size_t>(cdiv((size_t)v.dg.NX * v.dg.NY, 256), 256);
    }
    b->nitems = (int)items.size();
    for (int d = 0; d < n; d++) {
        bd[d].dt_lo = b->split ? uoff[d] : ilo[d];
        bd[d].dt_hi = b->split ? uoff[d + 1] : ihi[d];
    }
    b->eblocks = eoff[n]; b->hblocks = hoff[n]; b->sblocks = soff[n]; b->ublocks = uoff[n];
    void *p;
    int rc;
    if ((rc = balloc(b, &p, 2 * sizeof(double) * (size_t)std::max(b->ublocks, b->nitems)))) return rc;
    b->d_dtpart = (double *)p;
    for (auto &a : descs) a.dtpart = b->d_dtpart;
    if ((rc = balloc(b, &p, sizeof(FArgs<float>) * descs.size()))) return rc;
    b->d_descs = (FArgs<float> *)p;
    if ((rc = balloc(b, &p, sizeof(int4) * items.size()))) return rc;
    b->d_items = (int4 *)p;
    if ((rc = balloc(b, &p, sizeof(BDisc) * n))) return rc;
    b->d_bd = (BDisc *)p;
    if ((rc = balloc(b, &p, sizeof(int) * (n + 1)))) return rc;
    b->d_eoff = (int *)p;
    if ((rc = balloc(b, &p, sizeof(int) * (n + 1)))) return rc;
    b->d_hoff = (int *)p;
    if ((rc = balloc(b, &p, sizeof(int) * (n + 1)))) return rc;
    b->d_soff = (int *)p;
    if ((rc = balloc(b, &p, sizeof(int) * (n + 1)))) return rc;
    b->d_uoff = (int *)p;
    if ((rc = balloc(b, &p, sizeof(int) * (n + 1)))) return rc;
    b->d_ioff = (int *)p;
    BCHK(hipMemcpy(b->d_descs, descs.data(), sizeof(FArgs<float>) * descs.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_bd, bd.data(), sizeof(BDisc) * n, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_eoff, eoff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_hoff, hoff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_soff, soff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_uoff, uoff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->d_ioff, ioff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    b->built = true;
    b->primed = false;
    return CYLMHD_OK;
}

/* blocks of a kernel resident on the whole device at once, a multiple of 8
 * (the XCD remap needs one); 0 if the runtime cannot say */
int resident_blocks(cylmhd_batch *b, const void *fn, int threads)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess) return 0;
    return per_cu * b->cus / 8 * 8;
}

template <int LIM, bool HLL, int S, bool ALT>
void launch_shape(cylmhd_batch *b, const FArgs<float> *descs)
{
    if constexpr (shape_built(LIM, HLL, S)) {
        const int lo = b->groups[S][0], hi = b->groups[S][1];
        if (hi <= lo) return;
        const void *fn = b->split ? reinterpret_cast<const void *>(&k_march_batch<LIM, HLL, S>)
                                  : reinterpret_cast<const void *>(&k_fused_batch<LIM, HLL, S, ALT>);
        /* grid 0: as many blocks as are resident at once, so that every
         * block's run of items starts at launch */
        int want = b->grid;
        if (want <= 0) want = resident_blocks(b, fn, tile_threads<float, S>());
        if (want <= 0) want = 1024;
        const int G = std::max(8, std::min((hi - lo + 7) / 8 * 8, want));
        if (b->split)
            hipLaunchKernelGGL((k_march_batch<LIM, HLL, S>), dim3(G), dim3(tile_threads<float, S>()), 0, b->stream,
                               descs, b->d_items, lo, hi, b->d_bd);
        else
            hipLaunchKernelGGL((k_fused_batch<LIM, HLL, S, ALT>), dim3(G), dim3(tile_threads<float, S>()), 0, b->stream,
                               descs, b->d_items, lo, hi, b->d_bd);
    }
}

template <int LIM, bool HLL, bool ALT>
void launch_stage(cylmhd_batch *b, int s)
{
    const FArgs<float> *descs = b->d_descs + (size_t)s * b->n;
    hipLaunchKernelGGL((k_edge_batch<LIM, HLL>), dim3(b->eblocks), dim3(256), 0, b->stream, descs, b->d_eoff, b->n, b->d_bd);
    launch_shape<LIM, HLL, 0, ALT>(b, descs);
    launch_shape<LIM, HLL, 1, ALT>(b, descs);
    launch_shape<LIM, HLL, 2, ALT>(b, descs);
    launch_shape<LIM, HLL, 3, ALT>(b, descs);
    if (b->split)
        hipLaunchKernelGGL((k_update_batch<ALT>), dim3(b->ublocks), dim3(256), 0, b->stream, descs, b->d_uoff, b->n, b->d_bd);
    hipLaunchKernelGGL(k_halo_batch, dim3(b->hblocks), dim3(256), 0, b->stream, descs, b->d_hoff, b->n, b->d_bd);
}

/* one main.c step for every disc: Setdt, SpiralGravity, TVDRK */
void enqueue_step(cylmhd_batch *b)
{
    hipLaunchKernelGGL(k_dt_batch, dim3(b->n), dim3(256), 0, b->stream, b->d_descs, b->nst, b->n, b->d_bd, b->d_dtpart);
    hipLaunchKernelGGL(k_sincos_batch, dim3(b->sblocks), dim3(256), 0, b->stream, b->d_descs, b->d_soff, b->n, b->d_bd);
    /* one kernel per limiter x solver; discs that all run the default
     * source (SpiralGravityCooling) use a kernel compiled for it alone */
#define CYL_BSTAGE(L, H, A) launch_stage<L, H, A>(b, s)
#define CYL_BSTAGE_LIM(H, A)                                   \
    switch (b->lim) {                                           \
    case 0: CYL_BSTAGE(0, H, A); break;                         \
    case 1: CYL_BSTAGE(1, H, A); break;                         \
    case 2: CYL_BSTAGE(2, H, A); break;                         \
    default: CYL_BSTAGE(3, H, A); break;                        \
    }
    for (int s = 0; s < b->nst; s++) {
        if (b->alt) {
            if (b->hll) { CYL_BSTAGE_LIM(true, true) } else { CYL_BSTAGE_LIM(false, true) }
        } else {
            if (b->hll) { CYL_BSTAGE_LIM(true, false) } else { CYL_BSTAGE_LIM(false, false) }
        }
    }
#undef CYL_BSTAGE_LIM
#undef CYL_BSTAGE
    hipLaunchKernelGGL(k_time_batch, dim3(cdiv(b->n, 64)), dim3(64), 0, b->stream, b->d_bd, b->n);
}

/* Setdt partials of the stored states, floor counters cleared */
int prime(cylmhd_batch *b)
{
    for (int d = 0; d < b->n; d++) {
        BCHK(hipStreamSynchronize(b->view[d].stream));
        BCHK(hipMemsetAsync(b->view[d].ds.counts, 0, 9 * sizeof(int), b->stream));
    }
    hipLaunchKernelGGL(k_setdt_batch, dim3(b->split ? b->ublocks : b->nitems), dim3(256), 0, b->stream, b->d_descs, b->d_bd,
                       b->split ? b->d_uoff : b->d_ioff, b->n, b->d_dtpart);
    BCHK(hipGetLastError());
    b->primed = true;
    return CYLMHD_OK;
}

int capture(cylmhd_batch *b)
{
    free_graph(b);
    BCHK(hipStreamSynchronize(b->stream));
    BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
    if (e != hipSuccess) { set_errmsg("hipStreamEndCapture failed"); return CYLMHD_EHIP; }
    BCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return CYLMHD_OK;
}

int backup(cylmhd_batch *b, bool save)
{
    size_t off = 0;
    for (int d = 0; d < b->n; d++) {
        const size_t cnt = b->view[d].fs->nstate;
        float *un = (float *)b->view[d].fs->Un;
        if (save) BCHK(hipMemcpyAsync(b->d_backup + off, un, cnt * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
        else BCHK(hipMemcpyAsync(un, b->d_backup + off, cnt * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
        off += cnt;
    }
    return CYLMHD_OK;
}

}  // namespace

extern "C" {

int cylmhd_batch_create(const cylmhd_params *p, int n, cylmhd_batch **out)
{
    if (!p || n < 1 || !out) return CYLMHD_EINVAL;
    *out = nullptr;
    for (int d = 0; d < n; d++) {
        if (p[d].path != CYLMHD_PATH_FUSED32 || p[d].limiter != p[0].limiter || p[d].solver != p[0].solver ||
            p[d].rk != p[0].rk) {
            set_errmsg("batch: every disc must use the FP32 fused path and the same limiter, solver and RK");
            return CYLMHD_EINVAL;
        }
    }
    cylmhd_batch *b = new cylmhd_batch();
    b->n = n;
    {
        /* the stage form until the tuner picks one: two kernels, or one
         * with CYLMHD_STAGE_KERNELS=1 in the environment */
        const char *e = getenv("CYLMHD_STAGE_KERNELS");
        b->split = !(e && e[0] == '1');
    }
    b->dshape.assign(n, 0);
    b->lim = p[0].limiter;
    for (int d = 0; d < n; d++) b->alt = b->alt || p[d].source != CYLMHD_SOURCE_SPIRAL_COOLING;
    b->hll = p[0].solver == CYLMHD_SOLVER_HLL;
    b->rk3 = p[0].rk != 2;
    b->nst = b->rk3 ? 3 : 2;
    int rc = CYLMHD_OK;
    size_t tot = 0;
    for (int d = 0; d < n && !rc; d++) {
        cylmhd_ctx *c = nullptr;
        rc = create_ctx(&p[d], &c, true);           /* batch layout of the z-side ghosts */
        if (rc) break;
        b->disc.push_back(c);
        DiscView v;
        rc = disc_view(c, &v);
        b->view.push_back(v);
        tot += v.fs->nstate;
    }
    if (!rc && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) rc = CYLMHD_EHIP;
    if (!rc) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            rc = CYLMHD_EHIP;
        else
            b->cus = cus;
    }
    if (!rc && (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)) rc = CYLMHD_EHIP;
    if (!rc) {
        void *q;
        rc = balloc(b, &q, tot * sizeof(float));
        if (!rc) {
            b->d_backup = (float *)q;
            b->backup_elems = tot;
            b->allocs.clear();          /* the backup outlives work-list rebuilds */
        }
    }
    if (!rc) rc = build(b, nullptr);
    if (rc) {
        cylmhd_batch_destroy(b);
        return rc;
    }
    *out = b;
    return CYLMHD_OK;
}

int cylmhd_batch_destroy(cylmhd_batch *b)
{
    if (!b) return CYLMHD_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    free_lists(b);
    if (b->d_backup) hipFree(b->d_backup);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    for (cylmhd_ctx *c : b->disc) cylmhd_destroy(c);
    delete b;
    return CYLMHD_OK;
}

int cylmhd_batch_disc(cylmhd_batch *b, int d, cylmhd_ctx **ctx)
{
    if (!b || !ctx || d < 0 || d >= b->n) return CYLMHD_EINVAL;
    *ctx = b->disc[d];
    b->primed = false;            /* the caller may change the state */
    return CYLMHD_OK;
}

int cylmhd_batch_run(cylmhd_batch *b, int nsteps, double *ms)
{
    if (!b || nsteps < 0) return CYLMHD_EINVAL;
    int rc;
    if (!b->primed && (rc = prime(b))) return rc;
    if (!b->exec && (rc = capture(b))) return rc;
    BCHK(hipEventRecord(b->ev0, b->stream));
    for (int k = 0; k < nsteps; k++) BCHK(hipGraphLaunch(b->exec, b->stream));
    BCHK(hipEventRecord(b->ev1, b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    if (ms) {
        float f = 0.0f;
        BCHK(hipEventElapsedTime(&f, b->ev0, b->ev1));
        *ms = f;
    }
    for (cylmhd_ctx *c : b->disc) disc_invalidate_dt(c);
    return CYLMHD_OK;
}

int cylmhd_batch_status(cylmhd_batch *b, int d, double *t, double *dt, int *steps, int *alive)
{
    if (!b || d < 0 || d >= b->n) return CYLMHD_EINVAL;
    BDisc D;
    BCHK(hipMemcpy(&D, b->d_bd + d, sizeof D, hipMemcpyDeviceToHost));
    if (t) *t = D.t;
    if (dt) *dt = D.dt;
    if (steps) *steps = D.steps;
    if (alive) *alive = D.alive;
    return CYLMHD_OK;
}

int cylmhd_batch_tune(cylmhd_batch *b, int steps, int *choice)
{
    if (!b || steps < 1) return CYLMHD_EINVAL;
    static const int seglens[4] = {8, 16, 32, 0};
    static const int grids[3] = {0, 512, 1024};     /* 0: the resident blocks */
    int rc;
    /* every candidate starts from the discs' current states and step state
     * (time, dt, steps, alive: the spiral ramp and arm phase follow t, and
     * a stopped disc stays stopped and is skipped by the kernels); both are
     * put back after each candidate and at the end */
    std::vector<BDisc> keep;
    if (b->built && (rc = snapshot(b, keep))) return rc;
    const std::vector<BDisc> *kp = b->built ? &keep : nullptr;
    if ((rc = backup(b, true))) return rc;
    BCHK(hipStreamSynchronize(b->stream));
    auto time_it = [&](float *ms) -> int {
        int r;
        if ((r = build(b, kp)) || (r = prime(b))) return r;
        enqueue_step(b);                      /* warm-up */
        BCHK(hipEventRecord(b->ev0, b->stream));
        for (int k = 0; k < steps; k++) enqueue_step(b);
        BCHK(hipEventRecord(b->ev1, b->stream));
        BCHK(hipStreamSynchronize(b->stream));
        BCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
        return backup(b, false);
    };
    std::vector<int> avail;
    for (int sh = 0; sh < kShapes; sh++)
        if (sh < 2 || ((b->lim == 0 || b->lim == 1) && !b->hll)) avail.push_back(sh);
    /* 1: the stage form x one shape for every disc x segment length x grid */
    double best = 1e300;
    int bs = 0, bl = 0, bg = 0;
    bool bf = true;
    for (int form = 1; form >= 0; form--)
        for (int sh : avail)
            for (int li = 0; li < 4; li++)
                for (int gi = 0; gi < 3; gi++) {
                    b->split = form == 1;
                    b->dshape.assign(b->n, sh);
                    b->seglen = seglens[li];
                    b->grid = grids[gi];
                    float f = 0.0f;
                    if ((rc = time_it(&f))) return rc;
                    if (f < best) { best = f; bs = sh; bl = seglens[li]; bg = grids[gi]; bf = form == 1; }
                }
    b->split = bf;
    b->dshape.assign(b->n, bs);
    b->seglen = bl;
    b->grid = bg;
    /* 2: per class of disc size (the discs of one phi x z size), the other
     * shapes, largest classes first, each kept where it is faster */
    std::vector<std::pair<long long, long long>> classes;   /* (cells, key) */
    for (int d = 0; d < b->n; d++) {
        const long long key = (long long)b->view[d].dg.NY * 100000 + b->view[d].dg.NZ;
        bool found = false;
        for (auto &c : classes)
            if (c.second == key) { c.first += (long long)b->view[d].dg.NX * b->view[d].dg.NY * b->view[d].dg.NZ; found = true; }
        if (!found) classes.push_back({(long long)b->view[d].dg.NX * b->view[d].dg.NY * b->view[d].dg.NZ, key});
    }
    std::sort(classes.begin(), classes.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    if (avail.size() > 1 && classes.size() > 1) {
        for (const auto &c : classes) {
            std::vector<int> members;
            for (int d = 0; d < b->n; d++)
                if ((long long)b->view[d].dg.NY * 100000 + b->view[d].dg.NZ == c.second) members.push_back(d);
            const int cur = b->dshape[members[0]];
            int pick = cur;
            for (int sh : avail) {
                if (sh == cur) continue;
                for (int d : members) b->dshape[d] = sh;
                float f = 0.0f;
                if ((rc = time_it(&f))) return rc;
                if (f < best) { best = f; pick = sh; }
            }
            for (int d : members) b->dshape[d] = pick;
        }
    }
    BCHK(hipStreamSynchronize(b->stream));
    if ((rc = build(b, kp))) return rc;
    /* the floor counters of the timed candidates are not the discs' */
    for (int d = 0; d < b->n; d++) BCHK(hipMemsetAsync(b->view[d].ds.counts, 0, 9 * sizeof(int), b->stream));
    BCHK(hipStreamSynchronize(b->stream));
    if (choice) { choice[0] = bs; choice[1] = bl; choice[2] = bg; choice[3] = bf ? 2 : 1; }
    return CYLMHD_OK;
}

int cylmhd_batch_disc_tile(cylmhd_batch *b, int d, int tile[4])
{
    if (!b || !tile || d < 0 || d >= b->n) return CYLMHD_EINVAL;
    const int sh = b->dshape[d];
    tile[0] = sh;
    tile[1] = shape_tj(sh);
    tile[2] = shape_tk(sh);
    tile[3] = (shape_tj(sh) + 2) * shape_tk(sh);
    return CYLMHD_OK;
}

int cylmhd_batch_floor_counts(cylmhd_batch *b, int d, int counts[9])
{
    if (!b || d < 0 || d >= b->n) return CYLMHD_EINVAL;
    BCHK(hipStreamSynchronize(b->stream));
    return cylmhd_floor_counts(b->disc[d], counts);
}

}  // extern "C"
