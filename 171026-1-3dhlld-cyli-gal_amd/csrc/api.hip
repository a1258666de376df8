/*
 * api.hip -- the C ABI of include/cylmhd.h on top of the device paths.
 *
 * A context owns one HIP stream (default priority), the geometry / gravity
 * tables and the state of its path.  Host arrays cross the boundary in the
 * reference layout [9][NX][NY][NZ] FP64.  No function exits the process;
 * HIP failures come back as CYLMHD_EHIP with the HIP message kept for
 * cylmhd_error_string.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "batch.h"
#include "fused.h"
#include "paths.h"
#include "textio.h"

using namespace cyl;

namespace {
thread_local char g_errmsg[512] = "";

int hip_fail(hipError_t e, const char *what)
{
    snprintf(g_errmsg, sizeof g_errmsg, "%s: %s", what, hipGetErrorString(e));
    return CYLMHD_EHIP;
}
#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #call);   \
    } while (0)

const int kEventPairs = 96;
}  // namespace

struct cylmhd_ctx {
    cylmhd_params p;
    cyl_grid g;
    cyl_spiral_tables sp;
    hipStream_t stream = nullptr;
    DevGeom dg{};
    DevState ds{};
    FusedState fs{};
    std::vector<void *> allocs;
    double *d_sincos = nullptr, *d_logxs = nullptr;
    double *d_dzm = nullptr;        /* [NZ] cell heights dzm (the column integral of a slice) */
    double *d_slice = nullptr;      /* one slice, the largest plane */
    size_t slice_cap = 0;
    bool have_state = false;
    bool batch_layout = false;      /* a disc of a batch (z-side ghosts packed) */
    /* timing of the per-stage hot kernels */
    bool timing = false;
    hipEvent_t ev[kEventPairs][3];  /* start, between the two kernels of a split stage, end */
    bool ev_mid[kEventPairs] = {};
    int ev_used = 0;
    bool ev_ready = false;
};

namespace {

template <typename T>
int dalloc(cylmhd_ctx *c, T **ptr, size_t count)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
    if (e != hipSuccess) {
        snprintf(g_errmsg, sizeof g_errmsg, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        return CYLMHD_ENOMEM;
    }
    c->allocs.push_back(p);
    *ptr = (T *)p;
    return CYLMHD_OK;
}

template <typename T>
int dupload(cylmhd_ctx *c, const T **dst, const T *src, size_t count)
{
    T *p;
    int rc = dalloc(c, &p, count);
    if (rc) return rc;
    HIPCHK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    *dst = p;
    return CYLMHD_OK;
}

Consts consts_of(const cylmhd_params *p, double dt, double ch)
{
    Consts k{};
    k.gm = p->gm; k.gm1 = p->gm1; k.oneogm1 = p->oneogm1;
    k.dt = dt; k.ch = ch;
    k.psicoef = exp(-dt * ch * (1.0 / p->cr) * 0.5);   /* tvdrungekutta.h:22 */
    k.prcr = p->prcr; k.betacr = p->betacr; k.oneobetacr = (1.0 / p->betacr); k.rhocr = p->rhocr;
    k.omegasp = p->omegasp;
    k.te0 = p->te0;
    k.cooling = p->cooling;
    k.cool_lambda0 = p->cool_lambda0; k.cool_tmin = p->cool_tmin; k.cool_index = p->cool_index;
    k.source = p->source;
    return k;
}

Consts make_consts(const cylmhd_ctx *c, double dt, double ch) { return consts_of(&c->p, dt, ch); }

int validate(const cylmhd_params *p)
{
    if (!p) return CYLMHD_EINVAL;
    if (p->ix < 12 || p->jy < 2 || (p->jy & 1) || p->kz < 2) {
        snprintf(g_errmsg, sizeof g_errmsg, "grid %dx%dx%d: need ix >= 12, even jy >= 2, kz >= 2", p->ix, p->jy, p->kz);
        return CYLMHD_EINVAL;
    }
    if (p->rk != 2 && p->rk != 3) { snprintf(g_errmsg, sizeof g_errmsg, "rk must be 2 or 3"); return CYLMHD_EINVAL; }
    if (p->rk == 2 && p->kz < 9) {
        /* ZRightAbsorbBoundary reaches k = kzmax1-10 (zboundary.h:33-58) */
        snprintf(g_errmsg, sizeof g_errmsg, "RK2 needs kz >= 9 (z absorbing layer of 11 cells)");
        return CYLMHD_EINVAL;
    }
    if ((int)(p->rin / p->dx0) + 1 > p->ix + 2) { snprintf(g_errmsg, sizeof g_errmsg, "rin too large"); return CYLMHD_EINVAL; }
    if (p->path < CYLMHD_PATH_PLAIN64 || p->path > CYLMHD_PATH_FUSED32) return CYLMHD_EINVAL;
    if (p->limiter < CYLMHD_LIMITER_MINMOD || p->limiter > CYLMHD_LIMITER_SUPERBEE || p->solver < CYLMHD_SOLVER_HLLD ||
        p->solver > CYLMHD_SOLVER_HLL) {
        snprintf(g_errmsg, sizeof g_errmsg, "unknown limiter %d or solver %d", p->limiter, p->solver);
        return CYLMHD_EINVAL;
    }
    if (p->source < CYLMHD_SOURCE_SPIRAL_COOLING || p->source > CYLMHD_SOURCE_NONE) {
        snprintf(g_errmsg, sizeof g_errmsg, "unknown source %d", p->source);
        return CYLMHD_EINVAL;
    }
    return CYLMHD_OK;
}

void destroy_ctx(cylmhd_ctx *c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->ev_ready)
        for (int q = 0; q < kEventPairs; q++)
            for (int e = 0; e < 3; e++) hipEventDestroy(c->ev[q][e]);
    for (void *p : c->allocs) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    cyl_grid_free(&c->g);
    cyl_spiral_tables_free(&c->sp);
    delete c;
}

int build_tables(cylmhd_ctx *c)
{
    const cylmhd_params *p = &c->p;
    cyl_grid &g = c->g;
    int rc;
    if ((rc = cyl_grid_build(p, &g))) return rc;
    if ((rc = cyl_spiral_tables_build(p, &g, &c->sp))) return rc;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    std::vector<double> grx((size_t)NX * NZ), grz((size_t)NX * NZ);
    if ((rc = cyl_gravity_axi(p, &g, grx.data(), grz.data()))) return rc;
    DevGeom &d = c->dg;
    d.NX = NX; d.NY = NY; d.NZ = NZ; d.jy = p->jy; d.irin = g.irin;
    d.oneody = g.oneody;
    const size_t T = (size_t)NX * NZ;
    if ((rc = dupload(c, &d.x, g.x, NX)) || (rc = dupload(c, &d.xm, g.xm, NX)) ||
        (rc = dupload(c, &d.oneox, g.oneox, NX)) || (rc = dupload(c, &d.oneodx, g.oneodx, NX)) ||
        (rc = dupload(c, &d.oneodz, g.oneodz, NZ)) || (rc = dupload(c, &d.y, g.y, NY)) ||
        (rc = dupload(c, &d.z, g.z, NZ)) || (rc = dupload(c, &d.g1, g.g1, T)) ||
        (rc = dupload(c, &d.g2, g.g2, T)) || (rc = dupload(c, &d.g3, g.g3, T)) ||
        (rc = dupload(c, &d.g5, g.g5, T)) || (rc = dupload(c, &d.dvol, g.dvol, T)) ||
        (rc = dupload(c, &d.minlength, g.minlength, T)) ||
        (rc = dupload(c, &d.grxaxi, (const double *)grx.data(), T)) ||
        (rc = dupload(c, &d.grzaxi, (const double *)grz.data(), T)) ||
        (rc = dupload(c, &d.sp_x2s2, c->sp.x2s2, NX)) || (rc = dupload(c, &d.sp_o1x2s2, c->sp.o1x2s2, NX)) ||
        (rc = dupload(c, &d.sp_xs2o15, c->sp.xs2o15, NX)) || (rc = dupload(c, &d.sp_x2s2o15, c->sp.x2s2o15, NX)) ||
        (rc = dupload(c, &d.sp_szs, c->sp.szs, NZ)) || (rc = dupload(c, &d.sp_onepa, c->sp.onepa, NZ)) ||
        (rc = dupload(c, &d.sp_oz2s2, c->sp.oz2s2, NZ)) || (rc = dupload(c, &d.sp_q, c->sp.q, T)) ||
        (rc = dupload(c, &d.xabs, (const double *)g.xabs, 11)) || (rc = dupload(c, &d.zabs, (const double *)g.zabs, 11)) ||
        (rc = dupload(c, &d.cabs, g.cabs, (size_t)(g.irin + 1) * NZ)))
        return rc;
    const double *lx;
    if ((rc = dupload(c, &lx, c->sp.logxs, NX))) return rc;
    c->d_logxs = (double *)lx;
    const double *dz;
    if ((rc = dupload(c, &dz, g.dzm, NZ))) return rc;
    c->d_dzm = (double *)dz;
    if ((rc = dalloc(c, &c->d_sincos, 2 * (size_t)NX * NY))) return rc;
    HIPCHK(hipMemset(c->d_sincos, 0, 2 * sizeof(double) * NX * NY));
    d.sincos = c->d_sincos;
    d.cotpitch = c->sp.cotpitch;
    d.armnum = p->armnum;
    d.epssp = 0.0;
    /* absorbing i-planes: [0, irin] and [NX-11, NX-1] */
    std::vector<int> il, ol;
    for (int i = 0; i < NX; i++) {
        if (i <= g.irin || i >= NX - 11) il.push_back(i);
        else ol.push_back(i);
    }
    const int *pil, *pol;
    if ((rc = dupload(c, &pil, (const int *)il.data(), il.size()))) return rc;
    if ((rc = dupload(c, &pol, (const int *)(ol.empty() ? il.data() : ol.data()), ol.empty() ? 1 : ol.size()))) return rc;
    c->ds.absorb_i = (int *)pil; c->ds.n_absorb_i = (int)il.size();
    c->ds.nonabsorb_i = (int *)pol; c->ds.n_nonabsorb_i = (int)ol.size();
    return CYLMHD_OK;
}

int alloc_state(cylmhd_ctx *c)
{
    DevState &d = c->ds;
    const int NX = c->g.NX, NY = c->g.NY, NZ = c->g.NZ;
    d.N = (size_t)NX * NY * NZ;
    d.NF = (size_t)(NX - 1) * (NY - 1) * (NZ - 1);
    int rc;
    if ((rc = dalloc(c, &d.Uinit, 9 * d.N)) || (rc = dalloc(c, &d.Vinit_z, 9 * 2 * (size_t)NX * NY)) ||
        (rc = dalloc(c, &d.Vinit_y, 9 * 2 * (size_t)NX * NZ)) ||
        (rc = dalloc(c, &d.counts, 9)))
        return rc;
    d.nred = 8192;
    if ((rc = dalloc(c, &d.red, (size_t)d.nred))) return rc;
    {
        const size_t rphi = (size_t)(NX - 2) * (NY - 2), xz = 2 * (size_t)NX * (NZ - 2);
        c->slice_cap = rphi > xz ? rphi : xz;
        if ((rc = dalloc(c, &c->d_slice, c->slice_cap))) return rc;
    }
    HIPCHK(hipMemset(d.counts, 0, 9 * sizeof(int)));
    if (c->p.path == CYLMHD_PATH_PLAIN64) {
        if ((rc = dalloc(c, &d.U, 9 * d.N)) || (rc = dalloc(c, &d.U1, 9 * d.N)) || (rc = dalloc(c, &d.V, 9 * d.N)) ||
            (rc = dalloc(c, &d.F, 9 * d.NF)) || (rc = dalloc(c, &d.G, 9 * d.NF)) || (rc = dalloc(c, &d.H, 9 * d.NF)))
            return rc;
    } else {
        /* fused paths keep no V and no face arrays; V is only materialised
         * on demand (Setdt, download) into a scratch array */
        if ((rc = dalloc(c, &d.V, 9 * d.N))) return rc;
        c->fs.batch = c->batch_layout;
        if ((rc = fused_alloc(c->fs, c->dg, d, c->p.path == CYLMHD_PATH_FUSED32,
                              [c](size_t bytes, void **ptr) {
                                  char *q;
                                  int r = dalloc(c, &q, bytes);
                                  *ptr = q;
                                  return r;
                              })))
            return rc;
    }
    return CYLMHD_OK;
}

}  // namespace

namespace cyl {

int disc_view(cylmhd_ctx *c, DiscView *v)
{
    if (!c || !v || c->p.path != CYLMHD_PATH_FUSED32) return CYLMHD_EINVAL;
    v->p = &c->p;
    v->dg = c->dg;
    v->ds = c->ds;
    v->fs = &c->fs;
    v->sincos = c->d_sincos;
    v->logxs = c->d_logxs;
    v->minminlength = c->g.minminlength;
    v->stream = c->stream;
    return CYLMHD_OK;
}

Consts disc_consts(const cylmhd_ctx *c, double dt, double ch) { return make_consts(c, dt, ch); }

void disc_invalidate_dt(cylmhd_ctx *c)
{
    if (c) c->fs.dt_valid = false;
}

void set_errmsg(const char *msg) { snprintf(g_errmsg, sizeof g_errmsg, "%s", msg); }

}  // namespace cyl

extern "C" {

const char *cylmhd_error_string(int code)
{
    static thread_local char buf[640];
    const char *base = code == CYLMHD_OK ? "ok" : code == CYLMHD_EINVAL ? "invalid argument"
                     : code == CYLMHD_ENOMEM ? "out of memory" : code == CYLMHD_EHIP ? "HIP error"
                     : code == CYLMHD_EIO ? "I/O error" : code == CYLMHD_EUNSUP ? "unsupported" : "unknown error";
    snprintf(buf, sizeof buf, "%s%s%s", base, g_errmsg[0] ? ": " : "", g_errmsg);
    return buf;
}

int cylmhd_create(const cylmhd_params *p, cylmhd_ctx **out)
{
    return cyl::create_ctx(p, out, false);
}

}  // extern "C"

int cyl::create_ctx(const cylmhd_params *p, cylmhd_ctx **out, bool batch_layout)
{
    if (!out) return CYLMHD_EINVAL;
    *out = nullptr;
    g_errmsg[0] = 0;
    int rc = validate(p);
    if (rc) return rc;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) {
        snprintf(g_errmsg, sizeof g_errmsg, "no HIP device (%s)", hipGetErrorString(e));
        return CYLMHD_EHIP;
    }
    cylmhd_ctx *c = new cylmhd_ctx();
    c->p = *p;
    c->batch_layout = batch_layout;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        rc = hip_fail(e, "hipStreamCreate");
        destroy_ctx(c);
        return rc;
    }
    if ((rc = build_tables(c)) || (rc = alloc_state(c))) {
        destroy_ctx(c);
        return rc;
    }
    *out = c;
    return CYLMHD_OK;
}

extern "C" {

int cylmhd_destroy(cylmhd_ctx *c)
{
    destroy_ctx(c);
    return CYLMHD_OK;
}

int cylmhd_dims(const cylmhd_ctx *c, int *nx, int *ny, int *nz)
{
    if (!c) return CYLMHD_EINVAL;
    if (nx) *nx = c->g.NX;
    if (ny) *ny = c->g.NY;
    if (nz) *nz = c->g.NZ;
    return CYLMHD_OK;
}

int cylmhd_upload_state(cylmhd_ctx *c, const double *U, const double *V)
{
    if (!c || !U || !V) return CYLMHD_EINVAL;
    const size_t bytes = 9 * c->ds.N * sizeof(double);
    if (c->p.path == CYLMHD_PATH_PLAIN64) {
        HIPCHK(hipMemcpyAsync(c->ds.U, U, bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->ds.V, V, bytes, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(c->ds.V, U, bytes, hipMemcpyHostToDevice, c->stream));   /* staging */
        int rc = fused_load_state(c->stream, c->fs, c->dg, c->ds, c->ds.V);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_state = true;
    return CYLMHD_OK;
}

int cylmhd_set_initial(cylmhd_ctx *c, const double *U, const double *V)
{
    if (!c || !U || !V) return CYLMHD_EINVAL;
    const int NX = c->g.NX, NY = c->g.NY, NZ = c->g.NZ;
    const size_t N = c->ds.N, NP = (size_t)NX * NY;
    /* the copies below go through the null stream: first let the work
     * already queued on the context stream (a step reading Uinit) finish */
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->ds.Uinit, U, 9 * N * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> vz(9 * 2 * NP);
    for (int m = 0; m < 9; m++)
        for (size_t ij = 0; ij < NP; ij++) {
            vz[(size_t)m * 2 * NP + ij] = V[(size_t)m * N + ij * NZ + 0];
            vz[(size_t)m * 2 * NP + NP + ij] = V[(size_t)m * N + ij * NZ + (NZ - 1)];
        }
    HIPCHK(hipMemcpy(c->ds.Vinit_z, vz.data(), vz.size() * sizeof(double), hipMemcpyHostToDevice));
    const size_t NQ = (size_t)NX * NZ;
    std::vector<double> vy(9 * 2 * NQ);
    for (int m = 0; m < 9; m++)
        for (int i = 0; i < NX; i++)
            for (int k = 0; k < NZ; k++) {
                vy[(size_t)m * 2 * NQ + (size_t)i * NZ + k] = V[(size_t)m * N + ((size_t)i * NY + 0) * NZ + k];
                vy[(size_t)m * 2 * NQ + NQ + (size_t)i * NZ + k] = V[(size_t)m * N + ((size_t)i * NY + NY - 1) * NZ + k];
            }
    HIPCHK(hipMemcpy(c->ds.Vinit_y, vy.data(), vy.size() * sizeof(double), hipMemcpyHostToDevice));
    if (c->p.path != CYLMHD_PATH_PLAIN64) {
        int rc = fused_set_initial_tables(c->stream, c->fs, c->dg, c->ds, c->p.rhocr);
        if (rc) return rc;
    }
    return cylmhd_upload_state(c, U, V);
}

int cylmhd_download_state(cylmhd_ctx *c, double *U, double *V)
{
    if (!c || !c->have_state) return CYLMHD_EINVAL;
    const size_t bytes = 9 * c->ds.N * sizeof(double);
    if (c->p.path == CYLMHD_PATH_PLAIN64) {
        if (U) HIPCHK(hipMemcpyAsync(U, c->ds.U, bytes, hipMemcpyDeviceToHost, c->stream));
        if (V) HIPCHK(hipMemcpyAsync(V, c->ds.V, bytes, hipMemcpyDeviceToHost, c->stream));
    } else {
        /* U (FP64 copy of the path's state) -> fs.u64_scratch, V = UtoV(U) -> ds.V */
        Consts k = make_consts(c, 0.0, 0.0);
        int rc = fused_export(c->stream, c->fs, c->dg, c->ds, k, c->ds.V);
        if (rc) return rc;
        if (U) HIPCHK(hipMemcpyAsync(U, c->fs.u64_scratch, bytes, hipMemcpyDeviceToHost, c->stream));
        if (V) HIPCHK(hipMemcpyAsync(V, c->ds.V, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return CYLMHD_OK;
}

int cylmhd_setdt(cylmhd_ctx *c, double *dt, double *ch, int ijk[3])
{
    if (!c || !c->have_state) return CYLMHD_EINVAL;
    Consts k = make_consts(c, 0.0, 0.0);
    double res[2];
    if (c->p.path == CYLMHD_PATH_PLAIN64) {
        launch_setdt(c->stream, c->dg, k, c->ds.V, c->ds.red, c->ds.nred);
    } else {
        int rc = fused_setdt(c->stream, c->fs, c->dg, c->ds, k);
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(res, c->ds.red, sizeof res, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double oneocfldt = res[0];
    const long long ix = res[1] < 0 ? -1 : (long long)res[1];
    if (dt) *dt = c->p.cfl / oneocfldt;                 /* setdt.h:84 */
    if (ch) *ch = c->g.minminlength * oneocfldt;        /* setdt.h:89 */
    if (ijk) {
        const long long NY = c->g.NY, NZ = c->g.NZ;
        ijk[0] = ix < 0 ? -1 : (int)(ix / (NY * NZ));
        ijk[1] = ix < 0 ? -1 : (int)((ix / NZ) % NY);
        ijk[2] = ix < 0 ? -1 : (int)(ix % NZ);
    }
    return CYLMHD_OK;
}

int cylmhd_gravity(cylmhd_ctx *c, double t)
{
    if (!c) return CYLMHD_EINVAL;
    c->dg.epssp = cyl_spiral_eps(&c->p, t);
    launch_sincos(c->stream, c->dg, c->d_sincos, c->d_logxs, t, c->p.omegapattern, c->p.omegasp);
    HIPCHK(hipGetLastError());
    return CYLMHD_OK;
}

int cylmhd_step(cylmhd_ctx *c, double dt, double ch)
{
    if (!c || !c->have_state) return CYLMHD_EINVAL;
    if (c->batch_layout) {
        snprintf(g_errmsg, sizeof g_errmsg, "this disc belongs to a batch: cylmhd_batch_run advances it");
        return CYLMHD_EINVAL;
    }
    const Consts k = make_consts(c, dt, ch);
    const bool rk3 = c->p.rk != 2;
    const int nst = rk3 ? 3 : 2;
    HIPCHK(hipMemsetAsync(c->ds.counts, 0, 9 * sizeof(int), c->stream));
    for (int s = 1; s <= nst; s++) {
        hipEvent_t e0 = nullptr, em = nullptr, e1 = nullptr;
        if (c->timing && c->ev_used < kEventPairs) {
            e0 = c->ev[c->ev_used][0];
            em = c->ev[c->ev_used][1];
            e1 = c->ev[c->ev_used][2];
            c->ev_mid[c->ev_used] = c->p.path != CYLMHD_PATH_PLAIN64 && c->fs.split;
            c->ev_used++;
        }
        StageArgs a{s, rk3, c->p.limiter == CYLMHD_LIMITER_MC, c->p.solver == CYLMHD_SOLVER_HLL, c->p.limiter};
        if (c->p.path == CYLMHD_PATH_PLAIN64) {
            plain_stage(c->stream, c->dg, c->ds, k, a, e0, e1);
        } else {
            int rc = fused_stage(c->stream, c->fs, c->dg, c->ds, k, a, e0, em, e1);
            if (rc) return rc;
        }
    }
    HIPCHK(hipGetLastError());
    return CYLMHD_OK;
}

int cylmhd_floor_counts(cylmhd_ctx *c, int counts[9])
{
    if (!c || !counts) return CYLMHD_EINVAL;
    HIPCHK(hipMemcpyAsync(counts, c->ds.counts, 9 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CYLMHD_OK;
}

int cylmhd_diagnostics(cylmhd_ctx *c, cylmhd_diag *d)
{
    if (!c || !d || !c->have_state) return CYLMHD_EINVAL;
    double res[4];
    if (c->p.path == CYLMHD_PATH_PLAIN64) {
        launch_diag(c->stream, c->dg, c->ds.U, c->ds.red, c->ds.nred);
    } else {
        int rc = fused_diag(c->stream, c->fs, c->dg, c->ds);
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(res, c->ds.red, sizeof res, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    d->mass = res[0]; d->energy = res[1]; d->divb_l1 = res[2]; d->psi_l1 = res[3];
    return CYLMHD_OK;
}

int cylmhd_sync(cylmhd_ctx *c)
{
    if (!c) return CYLMHD_EINVAL;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipGetLastError());
    return CYLMHD_OK;
}

int cylmhd_kernel_timing(cylmhd_ctx *c, int which, int reset, double *avg_ms, int *launches)
{
    if (!c || which < 0 || which > 2) return CYLMHD_EINVAL;
    if (!c->ev_ready) {
        for (int q = 0; q < kEventPairs; q++)
            for (int e = 0; e < 3; e++) HIPCHK(hipEventCreate(&c->ev[q][e]));
        c->ev_ready = true;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    int n = 0;
    for (int q = 0; q < c->ev_used; q++) {
        if (which != 0 && !c->ev_mid[q]) continue;      /* one-kernel stage: no split to report */
        const int e0 = which == 2 ? 1 : 0, e1 = which == 1 ? 1 : 2;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[q][e0], c->ev[q][e1]));
        tot += ms;
        n++;
    }
    if (avg_ms) *avg_ms = n ? tot / n : 0.0;
    if (launches) *launches = n;
    if (reset) c->ev_used = 0;
    c->timing = true;
    return CYLMHD_OK;
}

}  // extern "C"

/* ---- slices and text writers ------------------------------------------ */

namespace {

long long slice_count(const cylmhd_ctx *c, int plane)
{
    const long long NX = c->g.NX, NY = c->g.NY, NZ = c->g.NZ;
    switch (plane) {
    case CYLMHD_PLANE_RPHI:
    case CYLMHD_PLANE_RPHI_ZMEAN:
    case CYLMHD_PLANE_RPHI_ZINT: return (NX - 2) * (NY - 2);
    case CYLMHD_PLANE_RZ:
    case CYLMHD_PLANE_ZR: return (NX - 2) * (NZ - 2);
    case CYLMHD_PLANE_XZ:
    case CYLMHD_PLANE_XZ_PHIMEAN: return 2 * NX * (NZ - 2);
    case CYLMHD_PLANE_PHI: return NY - 2;
    }
    return -1;
}

int slice_check(const cylmhd_ctx *c, int q, int plane, int a, int b)
{
    const int NX = c->g.NX, NY = c->g.NY, NZ = c->g.NZ;
    if (q < 0 || q >= CYLMHD_NQ) return CYLMHD_EINVAL;
    switch (plane) {
    case CYLMHD_PLANE_RPHI: return a >= 0 && a < NZ ? CYLMHD_OK : CYLMHD_EINVAL;
    case CYLMHD_PLANE_RPHI_ZMEAN:
    case CYLMHD_PLANE_RPHI_ZINT:
    case CYLMHD_PLANE_XZ_PHIMEAN: return CYLMHD_OK;
    case CYLMHD_PLANE_RZ:
    case CYLMHD_PLANE_ZR: return a >= 0 && a < NY ? CYLMHD_OK : CYLMHD_EINVAL;
    case CYLMHD_PLANE_XZ: return a >= 1 && a <= NY - 2 && (NY - 2) % 2 == 0 ? CYLMHD_OK : CYLMHD_EINVAL;
    case CYLMHD_PLANE_PHI: return a >= 0 && a < NX && b >= 0 && b < NZ ? CYLMHD_OK : CYLMHD_EINVAL;
    }
    return CYLMHD_EINVAL;
}

/* the FP64 primitives of the current state in ds.V (the fused paths keep
 * conserved values only and convert here, as cylmhd_download_state does) */
int primitives(cylmhd_ctx *c)
{
    if (c->p.path == CYLMHD_PATH_PLAIN64) return CYLMHD_OK;
    Consts k = make_consts(c, 0.0, 0.0);
    return fused_export(c->stream, c->fs, c->dg, c->ds, k, c->ds.V);
}

int slice_host(cylmhd_ctx *c, int q, int plane, int a, int b, double *out, size_t n)
{
    Consts k = make_consts(c, 0.0, 0.0);
    launch_slice(c->stream, c->dg, k, c->ds.V, c->d_dzm, q, plane, a, b, c->d_slice, (int)n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->d_slice, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CYLMHD_OK;
}

/* cut one slice (ds.V already current) and write it as text */
int slice_file(cylmhd_ctx *c, std::vector<double> &buf, const char *dir, const char *name, int q, int plane,
               int a)
{
    int rc = slice_check(c, q, plane, a, 0);
    if (rc) return rc;
    const size_t n = (size_t)slice_count(c, plane);
    buf.resize(n);
    if ((rc = slice_host(c, q, plane, a, 0, buf.data(), n))) return rc;
    return cyl_write_text(dir, name, buf.data(), n);
}

}  // namespace

extern "C" {

int cylmhd_slice_size(const cylmhd_ctx *c, int plane, size_t *n)
{
    if (!c || !n) return CYLMHD_EINVAL;
    const long long m = slice_count(c, plane);
    if (m < 0) return CYLMHD_EINVAL;
    *n = (size_t)m;
    return CYLMHD_OK;
}

int cylmhd_slice(cylmhd_ctx *c, int quantity, int plane, int a, int b, double *out, size_t n)
{
    if (!c || !out || !c->have_state) return CYLMHD_EINVAL;
    int rc = slice_check(c, quantity, plane, a, b);
    if (rc) return rc;
    if ((long long)n != slice_count(c, plane)) return CYLMHD_EINVAL;
    if ((rc = primitives(c))) return rc;
    return slice_host(c, quantity, plane, a, b, out, n);
}

int cylmhd_gdl_output_ctx(cylmhd_ctx *c, const char *dir, int nout)
{
    /* GDLOutputCylinder (output.h:280-384) with the planes cut on the device */
    if (!c || !dir || !c->have_state) return CYLMHD_EINVAL;
    int rc;
    if ((rc = cyl_mkdir(dir))) return rc;
    const cyl_grid &g = c->g;
    if ((rc = cyl_write_grid_text(dir, "r.dat", "phi.dat", "z.dat", g.NX, g.NY, g.NZ, g.x, g.y, g.z))) return rc;
    if ((rc = primitives(c))) return rc;
    std::vector<double> buf;
    char name[128];
    static const char *what[3] = {"density", "pressure", "benergy"};
    static const int qs[3] = {0, 7, CYLMHD_Q_BENERGY};
    for (int w = 0; w < 3; w++) {
        snprintf(name, sizeof name, "2d-%s-rphi-%d.dat", what[w], nout);
        if ((rc = slice_file(c, buf, dir, name, qs[w], CYLMHD_PLANE_RPHI, 1))) return rc;
    }
    snprintf(name, sizeof name, "2d-density-rz-%d.dat", nout);
    return slice_file(c, buf, dir, name, 0, CYLMHD_PLANE_RZ, 0);
}

int cylmhd_idl_output(cylmhd_ctx *c, const char *root, int nout, const cylmhd_idl_planes *planes)
{
    if (!c || !root || !c->have_state || nout < 0) return CYLMHD_EINVAL;
    const cyl_grid &g = c->g;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    cylmhd_idl_planes o = {1, 100, 1, {0, 0, 0, 0}, 0};
    if (planes) o = *planes;
    for (int K = 0; K < 4; K++) {
        if (o.azim_i[K] == 0) o.azim_i[K] = 1 + (K + 1) * (NX - 2) / 5;
        if (o.azim_i[K] < 1 || o.azim_i[K] > NX - 2) return CYLMHD_EINVAL;
    }
    if (o.azim_k == 0) o.azim_k = 1;
    if (o.azim_k < 1 || o.azim_k > NZ - 2) return CYLMHD_EINVAL;
    if (o.k100 > NZ - 2) o.k100 = NZ - 2;
    if (o.kplane < 1 || o.kplane > NZ - 2 || o.k100 < 1 || o.jplane < 1 || o.jplane > NY - 2 || (NY - 2) % 2)
        return CYLMHD_EINVAL;
    char dir0[4096], dir[4096];
    if (snprintf(dir0, sizeof dir0, "%s/idl0", root) >= (int)sizeof dir0 ||
        snprintf(dir, sizeof dir, "%s/idl%d", root, nout) >= (int)sizeof dir)
        return CYLMHD_EINVAL;
    int rc;
    if ((rc = cyl_mkdir(root)) || (rc = cyl_mkdir(dir0)) || (rc = cyl_mkdir(dir))) return rc;
    /* coordinates: r, phi, z of the interior cells; x of the cut through
     * the axis, -x[NX-1] .. -x[0], x[0] .. x[NX-1] (2*NX values) */
    std::vector<double> xs(2 * (size_t)NX);
    for (int i = 0; i < NX; i++) {
        xs[(size_t)(NX - 1 - i)] = -g.x[i];
        xs[(size_t)(NX + i)] = g.x[i];
    }
    for (const char *d : {(const char *)dir0, (const char *)dir}) {
        if ((rc = cyl_write_grid_text(d, "r.txt", "phi.txt", "z.txt", NX, NY, NZ, g.x, g.y, g.z)) ||
            (rc = cyl_write_text(d, "x.txt", xs.data(), xs.size())))
            return rc;
        if (nout == 0) break;
    }
    if ((rc = primitives(c))) return rc;
    struct Item {
        const char *name;
        int q, plane, a;
    };
    const int kp = o.kplane, jp = o.jplane;
    const Item items[] = {
        /* r-phi planes, fltarr(ny, nx): density.pro, beta.pro, benergy.pro, bphi0.pro, bphi100.pro,
         * br0.pro, divb0.pro, faradaydepth.pro, vdotb.pro, temperature.pro */
        {"zplanerho", 0, CYLMHD_PLANE_RPHI, kp},
        {"zplanebeta", CYLMHD_Q_BETA, CYLMHD_PLANE_RPHI, kp},
        {"zplanebenergy", CYLMHD_Q_BENERGY, CYLMHD_PLANE_RPHI, kp},
        {"z0bphi", 5, CYLMHD_PLANE_RPHI, 1},
        {"z100bphi", 5, CYLMHD_PLANE_RPHI, o.k100},
        {"zplanemeanbr", 4, CYLMHD_PLANE_RPHI_ZMEAN, 0},
        {"divb0", CYLMHD_Q_DIVB, CYLMHD_PLANE_RPHI, kp},
        {"faradaydepth-rphi", CYLMHD_Q_RHOBZ, CYLMHD_PLANE_RPHI_ZINT, 0},
        {"vdotb", CYLMHD_Q_VDOTB, CYLMHD_PLANE_RPHI, kp},
        {"temp-rphi", CYLMHD_Q_TEMP, CYLMHD_PLANE_RPHI, kp},
        /* r-z planes, fltarr(nz, nx): density-rz.pro, vphi-rz.pro, bphi-rz.pro, temp-rz.pro */
        {"rho-rz", 0, CYLMHD_PLANE_ZR, jp},
        {"vphi-rz", 2, CYLMHD_PLANE_ZR, jp},
        {"bphi-rz", 5, CYLMHD_PLANE_ZR, jp},
        {"temp-rz", CYLMHD_Q_TEMP, CYLMHD_PLANE_ZR, jp},
        /* cuts through the axis, fltarr(2*NX, nz): density-xz.pro, beta-xz.pro (and tcool-xz.pro),
         * bphi-xz.pro, divb-xz.pro, temp-xz.pro, vz-xz.pro, oneotcool-xz.pro */
        {"rho-xz", 0, CYLMHD_PLANE_XZ, jp},
        {"xzbeta", CYLMHD_Q_BETA, CYLMHD_PLANE_XZ, jp},
        {"xzmeanbphi", 5, CYLMHD_PLANE_XZ_PHIMEAN, 0},
        {"divb-xz", CYLMHD_Q_DIVB, CYLMHD_PLANE_XZ, jp},
        {"temp-xz", CYLMHD_Q_TEMP, CYLMHD_PLANE_XZ, jp},
        {"xzvz", 3, CYLMHD_PLANE_XZ, jp},
        {"oneocoolingtime-xz", CYLMHD_Q_ONEOTCOOL, CYLMHD_PLANE_XZ, jp},
    };
    std::vector<double> buf;
    char name[160];
    for (const Item &it : items) {
        snprintf(name, sizeof name, "%s-%d.dat", it.name, nout);
        if ((rc = slice_file(c, buf, dir, name, it.q, it.plane, it.a))) return rc;
    }
    /* azimuthal profiles, fltarr(ny) + readf: phiout0..3.pro, fourieramp.pro */
    static const struct { const char *name; int q; } prof[] = {
        {"potential", CYLMHD_Q_SPIRALPOT}, {"rho", 0}, {"te", CYLMHD_Q_TEMP},
        {"vperp", CYLMHD_Q_VPERP}, {"vpara", CYLMHD_Q_VPARA},
        {"velocitypara", CYLMHD_Q_VPARA}};   /* phiout1..3.pro read the same profile under this name */
    const size_t nph = (size_t)NY - 2;
    buf.resize(nph);
    for (int K = 0; K < 4; K++) {
        char adir[4096];
        if (snprintf(adir, sizeof adir, "%s/azimuth-%d", root, K) >= (int)sizeof adir) return CYLMHD_EINVAL;
        if ((rc = cyl_mkdir(adir)) || (rc = cyl_write_text(adir, "phi.dat", g.y + 1, nph))) return rc;
        for (const auto &pr : prof) {
            snprintf(name, sizeof name, "%s-%d.dat", pr.name, nout);
            if ((rc = slice_host(c, pr.q, CYLMHD_PLANE_PHI, o.azim_i[K], o.azim_k, buf.data(), nph)) ||
                (rc = cyl_write_text(adir, name, buf.data(), nph)))
                return rc;
        }
    }
    return CYLMHD_OK;
}

}  // extern "C"

/* ---- Setdt with no allocation ------------------------------------------
 * The block partials of the CFL maximum live in `scratch`, a file-scope
 * device array: each GPU the library runs on holds its own copy until the
 * process ends.  Host code never passes the array's name to a kernel (on the
 * host it only stands for the symbol): the device address of each GPU's copy
 * is looked up once per device, with that device current, and used from
 * then on.  Every call on a GPU shares that GPU's array; calls that overlap
 * on one GPU race. */
#include <mutex>

namespace {
constexpr int kNoallocBlocks = 1024;                 /* partials: (max, index) per block */
static __device__ __align__(256) double scratch[2 * kNoallocBlocks];
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

int scratch_ptr(double **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        snprintf(g_errmsg, sizeof g_errmsg, "device %d: at most %d devices", dev, kMaxDevices);
        return CYLMHD_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return CYLMHD_OK;
}
}  // namespace

extern "C" int cylmhd_setdt_noalloc(const cylmhd_params *p, const double *V, const double *minlength,
                                    double minminlength, void *stream, double *dt, double *ch, int ijk[3])
{
    if (!p || !V || !minlength) return CYLMHD_EINVAL;
    g_errmsg[0] = 0;
    if (p->ix < 12 || p->jy < 1 || p->kz < 1) {
        snprintf(g_errmsg, sizeof g_errmsg, "grid %dx%dx%d: Setdt needs ix >= 12 (it skips the outer 11 planes)",
                 p->ix, p->jy, p->kz);
        return CYLMHD_EINVAL;
    }
    const size_t NX = (size_t)p->ix + 2, NY = (size_t)p->jy + 2, NZ = (size_t)p->kz + 2;
    if (NX * NY * NZ >= (size_t)1 << 53) {
        /* the argmax index travels through the partials as a double */
        snprintf(g_errmsg, sizeof g_errmsg, "grid of %zu cells: at most 2^53 cells", NX * NY * NZ);
        return CYLMHD_EINVAL;
    }
    double *red = nullptr;
    int rc = scratch_ptr(&red);
    if (rc) return rc;
    /* the fields of the device tables that Setdt reads */
    DevGeom g{};
    g.NX = (int)NX; g.NY = (int)NY; g.NZ = (int)NZ; g.jy = p->jy;
    g.minlength = minlength;
    const Consts k = consts_of(p, 0.0, 0.0);
    hipStream_t s = (hipStream_t)stream;
    launch_setdt(s, g, k, V, red, 2 * kNoallocBlocks);   /* at most kNoallocBlocks partials */
    HIPCHK(hipGetLastError());
    double res[2];
    HIPCHK(hipMemcpyAsync(res, red, sizeof res, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double oneocfldt = res[0];
    const long long ix = res[1] < 0 ? -1 : (long long)res[1];
    if (dt) *dt = p->cfl / oneocfldt;                  /* setdt.h:84 */
    if (ch) *ch = minminlength * oneocfldt;            /* setdt.h:89 */
    if (ijk) {
        ijk[0] = ix < 0 ? -1 : (int)(ix / (long long)(NY * NZ));
        ijk[1] = ix < 0 ? -1 : (int)((ix / (long long)NZ) % (long long)NY);
        ijk[2] = ix < 0 ? -1 : (int)(ix % (long long)NZ);
    }
    return CYLMHD_OK;
}
