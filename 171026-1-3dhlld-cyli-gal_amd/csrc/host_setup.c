/*
 * host_setup.c -- host-side parameters, grid, gravity tables, initial discs
 * and the dump/restart byte format.  Plain C, no GPU.
 *
 * Expression order follows the reference line by line where a value feeds
 * the hot path (grid factors, gravity, VtoU), so the tables handed to the GPU
 * are bit-identical to the reference's globals.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "grid.h"

/* reference max/min (maxmin.h:16-26): ties and NaNs resolve to the 2nd / 1st */
static double rmax(double a, double b) { if (a > b) return a; else return b; }
static double rmin(double a, double b) { if (a > b) return b; else return a; }

/* ------------------------------------------------------------------------ */
int cylmhd_params_default(cylmhd_params *p)
{
    if (!p) return CYLMHD_EINVAL;
    memset(p, 0, sizeof *p);
    /* input.h:45-76 */
    p->ix = 240; p->jy = 64; p->kz = 200;
    p->lx = 2.0; p->lz = 0.03;
    p->x0 = 0.04; p->y0 = 0.0; p->z0 = 0.0;
    p->dx0 = 0.01;
    p->dy0 = (2.0 * M_PI) / (double)p->jy;          /* ly/(double)jy, ly = pi2 */
    p->dz0 = 0.0003;
    p->stretchratio = 1.05;
    p->dxmax = 50.0 * p->dx0;
    p->dzmax = 50.0 * p->dz0;
    p->rin = 0.1;
    p->cfl = 0.4; p->dtmin = 1.0e-9;
    p->rhocr = 1.0e-4; p->prcr = 0.015; p->betacr = 0.1;
    /* constant.h:62-79 (literals kept separate, as there) */
    p->gm = 5.0 / 3.0; p->gm1 = 2.0 / 3.0; p->oneogm1 = 1.5;
    p->cr = 0.18;
    /* te0 = mump*oneokb*v0*v0 (constant.h:194) */
    p->te0 = (0.619047619 * 1.673e-24) * (1.0 / 1.3806504e-16) * 6.558e+6 * 6.558e+6;
    /* input.h:88-96 */
    p->epssp0 = 0.05;
    p->tausp = 200.0 * 1.0e+6 * 3.1536e+7 / 4.7058e+15;
    p->scalez = 0.03; p->scaler = 0.7; p->scalephase = 0.1;
    p->armnum = 2;
    p->omegasp = 10.0e+5 / (1.0 / 4.7058e+15) * (1.0 / 3.08568e+21);
    p->omegapattern = 10.0e+5 / (1.0 / 4.7058e+15) * (1.0 / 3.08568e+21);
    p->pitch = 15.0 * M_PI / 180.0;
    /* active scheme of the shipped code */
    p->limiter = CYLMHD_LIMITER_MINMOD;
    p->solver = CYLMHD_SOLVER_HLLD;
    p->rk = 3;
    p->cooling = CYLMHD_COOL_RAYMOND;
    p->gravity = CYLMHD_GRAV_GALACTIC;
    p->gm_point = 1.0;
    p->cool_lambda0 = 1.0; p->cool_tmin = 1.0e-2; p->cool_index = 0.5;
    p->path = CYLMHD_PATH_PLAIN64;
    /* synthetic disc defaults */
    p->disc_h = 0.1; p->disc_floor = 1.0e-3; p->disc_cs2 = 0.01;
    p->disc_beta = 10.0; p->beta_lo = 0.1; p->beta_hi = 100.0; p->beta_mode = 0;
    p->pert_amp = 0.01;
    p->seed = 20171026ULL;
    return CYLMHD_OK;
}

int cylmhd_params_config(cylmhd_params *p, int config, int nr, int nphi, int nz)
{
    static const int dims[6][3] = {
        {0, 0, 0}, {64, 64, 32}, {256, 256, 128}, {512, 512, 256}, {64, 64, 32}, {768, 768, 384}};
    if (!p || config < 1 || config > 5) return CYLMHD_EINVAL;
    cylmhd_params_default(p);
    p->ix = nr > 0 ? nr : dims[config][0];
    p->jy = nphi > 0 ? nphi : dims[config][1];
    p->kz = nz > 0 ? nz : dims[config][2];
    if (p->ix < 12 || p->jy < 2 || (p->jy & 1) || p->kz < 3) return CYLMHD_EINVAL;
    /* r in [0.1, 1], phi in [0, 2pi), z in [-0.5, 0.5], uniform (no stretch) */
    p->dx0 = 0.9 / (double)p->ix;
    p->x0 = 0.1 + 0.5 * p->dx0;
    p->dy0 = (2.0 * M_PI) / (double)p->jy;
    p->y0 = 0.0;
    p->dz0 = 1.0 / (double)p->kz;
    p->z0 = -0.5 + 0.5 * p->dz0;
    p->lx = 1.0e300; p->lz = 1.0e300;
    p->dxmax = 50.0 * p->dx0; p->dzmax = 50.0 * p->dz0;
    p->gravity = CYLMHD_GRAV_POINTMASS;
    p->gm_point = 1.0;
    p->omegasp = 0.0; p->omegapattern = 0.0;   /* inertial frame, static pattern */
    p->tausp = 1.0e-30;                        /* spiral at full amplitude after t=0 */
    p->limiter = (config == 1 || config == 4) ? CYLMHD_LIMITER_MINMOD : CYLMHD_LIMITER_MC;
    p->epssp0 = config == 1 ? 0.0 : (config == 5 ? 0.10 : 0.05);
    p->armnum = 2;
    p->pitch = 15.0 * M_PI / 180.0;
    /* Discs 1-4 carry no cooling: the reference's Raymond fit is in its
     * galactic units (te0), and in these units its 2 rho^2 Lambda/(1.5 p)
     * CFL term collapses dt within ~40 steps.  The pressure floor is put at
     * 10 % of the disc's p/rho = cs^2 = 0.01: the reference's prcr = 0.015
     * lies above it and would floor every cell (and zero vr, vz) at every
     * stage. */
    p->cooling = CYLMHD_COOL_NONE;
    p->prcr = 1.0e-3;
    p->beta_mode = 0; p->disc_beta = 10.0;
    if (config == 3) { p->beta_mode = 1; p->beta_lo = 0.1; p->beta_hi = 100.0; }
    if (config == 4) { p->beta_mode = 1; p->beta_lo = 1.0; p->beta_hi = 100.0; }
    if (config == 5) {
        cyl_grid g;
        int rc;
        p->cooling = CYLMHD_COOL_POWERLAW;
        p->cool_lambda0 = 1.0; p->cool_tmin = 1.0e-2; p->cool_index = 0.5;
        /* GLM alpha = 0.1 of Mignone & Tzeferacos (2010): rate alpha*ch/dh,
         * i.e. cr = dh/alpha with dh the smallest cell length.            */
        if ((rc = cyl_grid_build(p, &g)) != 0) return rc;
        p->cr = g.minminlength / 0.1;
        cyl_grid_free(&g);
    }
    p->path = config == 1 ? CYLMHD_PATH_PLAIN64 : (config == 2 ? CYLMHD_PATH_PLAIN64 : CYLMHD_PATH_FUSED32);
    return CYLMHD_OK;
}

static double u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);

int cylmhd_params_batch(cylmhd_params *p, int n, uint64_t seed)
{
    static const int dims[5] = {32, 48, 64, 96, 128};
    if (!p || n < 1) return CYLMHD_EINVAL;
    for (int d = 0; d < n; d++) {
        /* draw 1: nr, 2: nphi, 3: nz among the dims not above nr, 4: spiral */
        const int ir = (int)(5.0 * u01(seed, (uint64_t)d, 1, 0));
        const int nr = dims[ir], nphi = dims[(int)(5.0 * u01(seed, (uint64_t)d, 2, 0))];
        const int nz = dims[(int)((double)(ir + 1) * u01(seed, (uint64_t)d, 3, 0))];
        int rc = cylmhd_params_config(&p[d], 4, nr, nphi, nz);
        if (rc) return rc;
        p[d].seed = seed + 7919ULL * (uint64_t)(d + 1);
        p[d].epssp0 = 0.1 * u01(seed, (uint64_t)d, 4, 0);
        p[d].path = CYLMHD_PATH_FUSED32;
    }
    return CYLMHD_OK;
}

/* ------------------------------------------------------------------------ */
void cyl_grid_free(cyl_grid *g)
{
    double **ptrs[] = {&g->x, &g->xm, &g->dx, &g->dxm, &g->oneox, &g->oneoxm, &g->oneodx,
                       &g->y, &g->z, &g->zm, &g->dz, &g->dzm, &g->oneodz,
                       &g->g1, &g->g2, &g->g3, &g->g5, &g->dvol, &g->minlength, &g->cabs};
    for (size_t n = 0; n < sizeof ptrs / sizeof ptrs[0]; ++n) { free(*ptrs[n]); *ptrs[n] = NULL; }
}

int cyl_grid_build(const cylmhd_params *p, cyl_grid *g)
{
    int i, j, k;
    memset(g, 0, sizeof *g);
    if (!p || p->ix < 9 || p->jy < 2 || (p->jy & 1) || p->kz < 1) return CYLMHD_EINVAL;
    const int NX = p->ix + 2, NY = p->jy + 2, NZ = p->kz + 2;
    g->NX = NX; g->NY = NY; g->NZ = NZ;
    g->irin = (int)(p->rin / p->dx0);
    if (g->irin + 1 > NX) return CYLMHD_EINVAL;
    g->x = calloc(NX, sizeof(double)); g->xm = calloc(NX, sizeof(double));
    g->dx = calloc(NX, sizeof(double)); g->dxm = calloc(NX, sizeof(double));
    g->oneox = calloc(NX, sizeof(double)); g->oneoxm = calloc(NX, sizeof(double));
    g->oneodx = calloc(NX, sizeof(double));
    g->y = calloc(NY, sizeof(double));
    g->z = calloc(NZ, sizeof(double)); g->zm = calloc(NZ, sizeof(double));
    g->dz = calloc(NZ, sizeof(double)); g->dzm = calloc(NZ, sizeof(double));
    g->oneodz = calloc(NZ, sizeof(double));
    g->g1 = calloc((size_t)NX * NZ, sizeof(double)); g->g2 = calloc((size_t)NX * NZ, sizeof(double));
    g->g3 = calloc((size_t)NX * NZ, sizeof(double)); g->g5 = calloc((size_t)NX * NZ, sizeof(double));
    g->dvol = calloc((size_t)NX * NZ, sizeof(double));
    g->minlength = calloc((size_t)NX * NZ, sizeof(double));
    g->cabs = calloc((size_t)(g->irin + 1) * NZ, sizeof(double));
    if (!g->x || !g->xm || !g->dx || !g->dxm || !g->oneox || !g->oneoxm || !g->oneodx || !g->y ||
        !g->z || !g->zm || !g->dz || !g->dzm || !g->oneodz || !g->g1 || !g->g2 || !g->g3 ||
        !g->g5 || !g->dvol || !g->minlength || !g->cabs) {
        cyl_grid_free(g);
        return CYLMHD_ENOMEM;
    }
    double *x = g->x, *xm = g->xm, *dx = g->dx, *dxm = g->dxm;
    double *z = g->z, *zm = g->zm, *dz = g->dz, *dzm = g->dzm;

    /* r: cylindricalgrid.h:36-65 */
    dx[0] = p->dx0;
    x[0] = p->x0 - p->dx0;
    for (i = 0; i < NX - 1; i++) {
        x[i + 1] = x[i] + dx[i];
        if (x[i] > p->lx) dx[i + 1] = rmin(p->stretchratio * dx[i], p->dxmax);
        else dx[i + 1] = dx[i];
        xm[i] = (x[i] + x[i + 1]) * 0.5;
        dxm[i] = (dx[i] + dx[i + 1]) * 0.5;
        g->oneox[i] = 1.0 / x[i];
        g->oneoxm[i] = 1.0 / xm[i];
        g->oneodx[i] = 1.0 / dx[i];
    }
    xm[NX - 1] = x[NX - 1] + 0.5 * dx[NX - 1];
    dxm[NX - 1] = dx[NX - 1];
    g->oneox[NX - 1] = 1.0 / x[NX - 1];
    g->oneodx[NX - 1] = 1.0 / dx[NX - 1];

    /* phi: :70-74 */
    g->dy = p->dy0;
    g->y[0] = p->y0 - g->dy;
    g->oneody = 1.0 / g->dy;
    for (j = 0; j < NY - 1; j++) g->y[j + 1] = g->y[j] + g->dy;

    /* z: :79-102 */
    dz[0] = p->dz0;
    z[0] = p->z0 - p->dz0;
    for (k = 0; k < NZ - 1; k++) {
        z[k + 1] = z[k] + dz[k];
        if (z[k] > p->lz) dz[k + 1] = rmin(p->stretchratio * dz[k], p->dzmax);
        else dz[k + 1] = dz[k];
        zm[k] = (z[k] + z[k + 1]) * 0.5;
        dzm[k] = (dz[k] + dz[k + 1]) * 0.5;
        g->oneodz[k] = 1.0 / dz[k];
    }
    zm[NZ - 1] = z[NZ - 1] + 0.5 * dz[NZ - 1];
    g->oneodz[NZ - 1] = 1.0 / dz[NZ - 1];

    /* volumes and face factors: :107-130 (depend on i,k only) */
    for (i = 1; i < NX - 1; i++) {
        for (k = 1; k < NZ - 1; k++) {
            const size_t ik = (size_t)i * NZ + k;
            double dvolume = 0.5 * (xm[i] * xm[i] - xm[i - 1] * xm[i - 1]) * g->dy * dzm[k - 1];
            double oneodvol = 1.0 / dvolume;
            g->dvol[ik] = dvolume;
            g->g1[ik] = xm[i - 1] * g->dy * dzm[k - 1] * oneodvol;
            g->g2[ik] = xm[i] * g->dy * dzm[k - 1] * oneodvol;
            g->g3[ik] = dxm[i - 1] * dzm[k - 1] * oneodvol;
            g->g5[ik] = 0.5 * (xm[i] * xm[i] - xm[i - 1] * xm[i - 1]) * g->dy * oneodvol;
        }
    }

    /* minimum cell length: :136-163 (the maximum of 1/minlength is exact in
     * any order, so the reference's racy search has one correct answer)   */
    double oneominminlength = 0.0;
    for (i = 0; i < NX; i++) {
        for (k = 0; k < NZ; k++) {
            double ml = rmin(dx[i], rmin(x[i] * g->dy, dz[k]));
            double oneominlength;
            g->minlength[(size_t)i * NZ + k] = ml;
            oneominlength = 1.0 / ml;
            if (oneominlength > oneominminlength) oneominminlength = oneominlength;
        }
    }
    g->minminlength = 1.0 / oneominminlength;

    /* absorbing profiles: :168-201 */
    for (i = NX - 1 - 10; i < NX; i++)
        g->xabs[i - (NX - 1 - 10)] =
            (0.5 * (1.0 - tanh((x[NX - 1] - x[i] - 5.0 * dx[NX - 1]) / (2.0 * dx[NX - 1]))));
    for (k = NZ - 1 - 10; k < NZ; k++)
        if (k >= 0)
            g->zabs[k - (NZ - 1 - 10)] =
                (0.5 * (1.0 - tanh((z[NZ - 1] - z[k] - 5.0 * dz[NZ - 1]) / (2.0 * dz[NZ - 1]))));
    for (i = 0; i < g->irin + 1; i++) {
        for (k = 0; k < NZ; k++) {
            double r = sqrt(x[i] * x[i] + z[k] * z[k]);
            g->cabs[(size_t)i * NZ + k] = (0.5 * (1.0 - tanh((r - p->rin + 5.0 * p->dx0) / (2.0 * p->dx0))));
        }
    }
    return CYLMHD_OK;
}

int cylmhd_grid_host(const cylmhd_params *p, double *x, double *y, double *z)
{
    cyl_grid g;
    int rc = cyl_grid_build(p, &g);
    if (rc) return rc;
    if (x) memcpy(x, g.x, sizeof(double) * g.NX);
    if (y) memcpy(y, g.y, sizeof(double) * g.NY);
    if (z) memcpy(z, g.z, sizeof(double) * g.NZ);
    cyl_grid_free(&g);
    return CYLMHD_OK;
}

/* ------------------------------------------------------------------------ */
int cyl_gravity_axi(const cylmhd_params *p, const cyl_grid *g, double *grx, double *grz)
{
    const int NX = g->NX, NZ = g->NZ;
    if (p->gravity == CYLMHD_GRAV_GALACTIC) {
        /* initialgravity.h:37-68 */
        const double vhalo0 = 131.5e+5 / 6.558e+6, dh = 1.2;
        const double a = 0.65, b = 0.026, md = 10.0;
        const double db = 0.07, mb = 3.4;
        for (int i = 0; i < NX; i++) {
            for (int k = 0; k < NZ; k++) {
                const double x = g->x[i], zz = g->z[k];
                const double x2 = x * x, z2 = zz * zz;
                grx[(size_t)i * NZ + k] =
                    ((-2.0 * vhalo0 * vhalo0 / (x2 + z2 + dh * dh)
                      - md / pow(x2 + (a + sqrt(z2 + b * b)) * (a + sqrt(z2 + b * b)), 1.5)
                      - mb / ((sqrt(x2 + z2) + db) * (sqrt(x2 + z2) + db) * sqrt(x2 + z2)))
                     * x);
                grz[(size_t)i * NZ + k] =
                    ((-2.0 * vhalo0 * vhalo0 / (x2 + z2 + dh * dh)
                      - md / pow(x2 + (a + sqrt(z2 + b * b)) * (a + sqrt(z2 + b * b)), 1.5)
                            * (1.0 + a / sqrt(z2 + b * b))
                      - mb / ((sqrt(x2 + z2) + db) * (sqrt(x2 + z2) + db) * sqrt(x2 + z2)))
                     * zz);
            }
        }
    } else if (p->gravity == CYLMHD_GRAV_POINTMASS) {
        for (int i = 0; i < NX; i++) {
            for (int k = 0; k < NZ; k++) {
                const double x = g->x[i], zz = g->z[k];
                const double r2 = x * x + zz * zz;
                const double inv = 1.0 / (r2 * sqrt(r2));
                grx[(size_t)i * NZ + k] = -p->gm_point * x * inv;
                grz[(size_t)i * NZ + k] = -p->gm_point * zz * inv;
            }
        }
    } else {
        return CYLMHD_EINVAL;
    }
    return CYLMHD_OK;
}

double cyl_spiral_eps(const cylmhd_params *p, double t)
{
    /* spiralgravity.h:41-42 (NaN at t == tausp, as there) */
    double switchfunc = (t - p->tausp) / fabs(t - p->tausp);
    return 0.5 * p->epssp0 * (t / p->tausp * (1.0 - switchfunc) + (1.0 + switchfunc));
}

void cyl_spiral_tables_free(cyl_spiral_tables *s)
{
    double **ptrs[] = {&s->x2s2, &s->o1x2s2, &s->xs2o15, &s->x2s2o15, &s->logxs,
                       &s->szs, &s->onepa, &s->oz2s2, &s->q};
    for (size_t n = 0; n < sizeof ptrs / sizeof ptrs[0]; ++n) { free(*ptrs[n]); *ptrs[n] = NULL; }
}

int cyl_spiral_tables_build(const cylmhd_params *p, const cyl_grid *g, cyl_spiral_tables *s)
{
    const int NX = g->NX, NZ = g->NZ;
    const double a1 = 0.65, b1 = 0.026;     /* spiralgravity.h:25-26 */
    memset(s, 0, sizeof *s);
    s->x2s2 = calloc(NX, sizeof(double)); s->o1x2s2 = calloc(NX, sizeof(double));
    s->xs2o15 = calloc(NX, sizeof(double)); s->x2s2o15 = calloc(NX, sizeof(double));
    s->logxs = calloc(NX, sizeof(double));
    s->szs = calloc(NZ, sizeof(double)); s->onepa = calloc(NZ, sizeof(double));
    s->oz2s2 = calloc(NZ, sizeof(double));
    s->q = calloc((size_t)NX * NZ, sizeof(double));
    if (!s->x2s2 || !s->o1x2s2 || !s->xs2o15 || !s->x2s2o15 || !s->logxs || !s->szs ||
        !s->onepa || !s->oz2s2 || !s->q) {
        cyl_spiral_tables_free(s);
        return CYLMHD_ENOMEM;
    }
    const double oneoscalephase = 1.0 / p->scalephase;
    const double oneoscaler = 1.0 / p->scaler;
    s->cotpitch = 1.0 / tan(p->pitch);      /* cot() has no libm definition */
    for (int i = 0; i < NX; i++) {
        const double x = g->x[i];
        const double x2 = x * x;
        const double xoscalephase = x * oneoscalephase;
        const double xoscaler = x * oneoscaler;
        const double xoscaler2 = xoscaler * oneoscaler;
        const double x2oscaler2 = xoscaler * xoscaler;
        const double oneoonex2oscaler2 = 1.0 / (1.0 + x2oscaler2);
        const double oneoonex2oscaler215 = 1.0 * pow(oneoonex2oscaler2, 1.5);
        s->x2s2[i] = x2oscaler2;
        s->o1x2s2[i] = oneoonex2oscaler2;
        s->xs2o15[i] = xoscaler2 * oneoonex2oscaler215;
        s->x2s2o15[i] = x2oscaler2 * oneoonex2oscaler215;
        s->logxs[i] = log(xoscalephase);
        for (int k = 0; k < NZ; k++) {
            const double z2 = g->z[k] * g->z[k];
            const double a1zk2b12 = ((a1 + sqrt(z2 + b1 * b1)) * (a1 + sqrt(z2 + b1 * b1)));
            s->q[(size_t)i * NZ + k] = 1.0 / sqrt(x2 + a1zk2b12);
        }
    }
    for (int k = 0; k < NZ; k++) {
        const double z2 = g->z[k] * g->z[k];
        s->szs[k] = (p->scalez / sqrt(z2 + p->scalez * p->scalez));
        s->onepa[k] = (1.0 + a1 / sqrt(z2 + b1 * b1));
        s->oz2s2[k] = 1.0 / (z2 + p->scalez * p->scalez);
    }
    return CYLMHD_OK;
}

/* ------------------------------------------------------------------------ */
/* counter-based RNG: splitmix64 finaliser chained over the key */
static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    uint64_t h = mix64(seed);
    h = mix64(h ^ a);
    h = mix64(h ^ (b << 21));
    h = mix64(h ^ (c << 42));
    return (double)(h >> 11) * 0x1.0p-53;
}

int cylmhd_model_disc(const cylmhd_params *p, double *V)
{
    cyl_grid g;
    int rc = cyl_grid_build(p, &g);
    if (rc) return rc;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    const size_t N = (size_t)NX * NY * NZ;
    for (int i = 0; i < NX; i++) {
        double beta = p->disc_beta;
        if (p->beta_mode == 1) {
            const double lo = log10(p->beta_lo), hi = log10(p->beta_hi);
            beta = pow(10.0, lo + (hi - lo) * u01(p->seed, 0xBE7AULL, (uint64_t)i, 0));
        }
        const double vphi0 = 1.0 / sqrt(g.x[i]);
        for (int j = 0; j < NY; j++) {
            for (int k = 0; k < NZ; k++) {
                const size_t c = ((size_t)i * NY + j) * NZ + k;
                const double zz = g.z[k];
                const double rho = exp(-zz * zz / (2.0 * p->disc_h * p->disc_h)) + p->disc_floor;
                const double pr = p->disc_cs2 * rho;
                const double u = 2.0 * u01(p->seed, (uint64_t)i + 1, (uint64_t)j + 1, (uint64_t)k + 1) - 1.0;
                V[0 * N + c] = rho;
                V[1 * N + c] = 0.0;
                V[2 * N + c] = vphi0 * (1.0 + p->pert_amp * u);
                V[3 * N + c] = 0.0;
                V[4 * N + c] = 0.0;
                V[5 * N + c] = sqrt(2.0 * pr / beta);
                V[6 * N + c] = 0.0;
                V[7 * N + c] = pr;
                V[8 * N + c] = 0.0;
            }
        }
    }
    cyl_grid_free(&g);
    return CYLMHD_OK;
}

/* Fill the ghost cells of U with the RK3 cell-boundary set (Y periodic, X
 * axis, X right free, Z periodic: firststep.h:125-157) applied to the
 * interior, and V there with UtoV (utov.h:24-68), so the initial state is
 * consistent with what every later stage leaves in the ghosts. */
static void fill_ghosts_rk3(const cylmhd_params *p, const cyl_grid *g, double *U, double *V)
{
    const int NX = g->NX, NY = g->NY, NZ = g->NZ, jy = p->jy;
    const size_t N = (size_t)NX * NY * NZ;
#define IX(i, j, k) ((((size_t)(i)) * NY + (size_t)(j)) * NZ + (size_t)(k))
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
            for (int k = 0; k < NZ; k++) {
                const int ghost = i == 0 || i == NX - 1 || j == 0 || j == NY - 1 || k == 0 || k == NZ - 1;
                if (!ghost) continue;
                int sk = k == 0 ? NZ - 2 : (k == NZ - 1 ? 1 : k), si, sj;
                unsigned sign = 0;
                if (i == 0) {
                    si = 1;
                    sj = j < jy / 2 + 1 ? j + jy / 2 : j - jy / 2;
                    sign = (1u << 1) | (1u << 2) | (1u << 4) | (1u << 5);
                } else {
                    si = i == NX - 1 ? NX - 2 : i;
                    sj = j == 0 ? NY - 2 : (j == NY - 1 ? 1 : j);
                }
                const size_t d = IX(i, j, k), s = IX(si, sj, sk);
                for (int m = 0; m < 9; m++) U[m * N + d] = (sign >> m) & 1u ? -U[m * N + s] : U[m * N + s];
                const double *u = U;
                const double oneox = g->oneox[i];
                const double oneorho = 1.0 / u[d];
                const double b2 = u[4 * N + d] * u[4 * N + d] + u[5 * N + d] * u[5 * N + d] + u[6 * N + d] * u[6 * N + d];
                const double r2v2 = u[1 * N + d] * u[1 * N + d] + u[2 * N + d] * u[2 * N + d] * oneox * oneox
                                    + u[3 * N + d] * u[3 * N + d];
                V[d] = u[d];
                V[1 * N + d] = u[1 * N + d] * oneorho;
                V[2 * N + d] = u[2 * N + d] * oneorho * oneox;
                V[3 * N + d] = u[3 * N + d] * oneorho;
                V[4 * N + d] = u[4 * N + d];
                V[5 * N + d] = u[5 * N + d];
                V[6 * N + d] = u[6 * N + d];
                V[7 * N + d] = p->gm1 * (u[7 * N + d] - 0.5 * (r2v2 * oneorho + b2));
                V[8 * N + d] = u[8 * N + d];
            }
#undef IX
}

int cylmhd_model_disc_state(const cylmhd_params *p, double *U, double *V)
{
    int rc;
    if ((rc = cylmhd_model_disc(p, V))) return rc;
    if ((rc = cylmhd_vtou_host(p, V, U))) return rc;
    cyl_grid g;
    if ((rc = cyl_grid_build(p, &g))) return rc;
    fill_ghosts_rk3(p, &g, U, V);
    cyl_grid_free(&g);
    return CYLMHD_OK;
}

int cylmhd_fill_ghosts_host(const cylmhd_params *p, double *U, double *V)
{
    cyl_grid g;
    int rc;
    if (!U || !V) return CYLMHD_EINVAL;
    if ((rc = cyl_grid_build(p, &g))) return rc;
    fill_ghosts_rk3(p, &g, U, V);
    cyl_grid_free(&g);
    return CYLMHD_OK;
}

int cylmhd_model_galactic_disc3(const cylmhd_params *p, double *V)
{
    /* model.h:164-292 (the z boundary calls on U at :312-313 are left to the
     * caller, which owns U) */
    cyl_grid g;
    int rc = cyl_grid_build(p, &g);
    if (rc) return rc;
    const int NX = g.NX, NY = g.NY, NZ = g.NZ;
    const size_t N = (size_t)NX * NY * NZ;
    const double v0 = 6.558e+6, kpc = 3.08568e+21, msun = 1.989e+33, mp = 1.673e-24;
    const double l0 = (10 * kpc);
    const double pi4 = 4.0 * M_PI;
    const double vhalo0 = 131.5e+5 / v0, dh = 1.2;
    const double a = 0.65, b = 0.026, md = 10.0;
    const double db = 0.07, mb = 3.4;
    const double mdisc = 0.4 * 0.75;
    const double rdisc = 0.7, zdisc = 0.03;
    const double discrho0 = (1.0e+10 * msun / (mp * 1000 * kpc * kpc * kpc) * mdisc / (pi4 * rdisc * rdisc * zdisc));
    const double csdisc2 = 1.33e+8 * 1.0e+4 / (v0 * v0);
    const double discbeta = 1.0e+4;
    const double halorho0 = 1.0e-4 * (1.0e+10 * msun / (mp * 1000 * kpc * kpc * kpc) * mdisc / (pi4 * rdisc * rdisc * zdisc));
    const double cshalo2 = 1.33e+8 * 2.4e+6 / (v0 * v0);
    const double oneocshalo2 = 1.0 / cshalo2;
    const double oneocsdisc2 = 1.0 / csdisc2;
    const double discbphi0 = sqrt(2.0 * csdisc2 * discrho0 / discbeta);
    for (int i = 0; i < NX; i++) {
        for (int j = 0; j < NY; j++) {
            for (int k = 0; k < NZ; k++) {
                const size_t c = ((size_t)i * NY + j) * NZ + k;
                const double x = g.x[i], zz = g.z[k];
                const double x2 = x * x, z2 = zz * zz;
                const double potaxi = ((vhalo0 * vhalo0 * (2.0 * log(l0) + log(x2 + z2 + dh * dh)))
                                       - md / sqrt(x2 + (a + sqrt(z2 + b * b)) * (a + sqrt(z2 + b * b)))
                                       - mb / (sqrt(x2 + z2) + db));
                const double potaxir = ((vhalo0 * vhalo0 * (2.0 * log(l0) + log(x2 + dh * dh)))
                                        - md / sqrt(x2 + (a + b) * (a + b))
                                        - mb / (x + db));
                const double potaxi0 = ((vhalo0 * vhalo0 * (2.0 * log(l0) + log(dh * dh)))
                                        - md / (a + b)
                                        - mb / db);
                const double grx = ((-2.0 * vhalo0 * vhalo0 / (x2 + z2 + dh * dh)
                                     - md / pow(x2 + (a + sqrt(z2 + b * b)) * (a + sqrt(z2 + b * b)), 1.5)
                                     - mb / ((sqrt(x2 + z2) + db) * (sqrt(x2 + z2) + db) * sqrt(x2 + z2)))
                                    * x);
                const double halorho = halorho0 * exp(-oneocshalo2 * (potaxi - potaxi0));
                const double halopr = cshalo2 * halorho;
                const double discrho = (discrho0 * exp(-x / rdisc) * exp(-oneocsdisc2 * (potaxi - potaxir)));
                const double discpr = csdisc2 * discrho;
                V[0 * N + c] = discrho + halorho;
                V[1 * N + c] = 0.0;
                V[2 * N + c] = (sqrt(-x * (grx + csdisc2 / rdisc)) * discrho / V[0 * N + c] - p->omegasp * x);
                V[3 * N + c] = 0.0;
                V[4 * N + c] = 0.0;
                V[5 * N + c] = (discbphi0 * exp(-x / rdisc) / (cosh(zz / zdisc) * cosh(zz / zdisc)));
                V[6 * N + c] = 0.0;
                V[7 * N + c] = discpr + halopr;
                V[8 * N + c] = 0.0;
            }
        }
    }
    cyl_grid_free(&g);
    return CYLMHD_OK;
}

int cylmhd_vtou_host(const cylmhd_params *p, const double *V, double *U)
{
    /* vtou.h:20-62 */
    cyl_grid g;
    int rc = cyl_grid_build(p, &g);
    if (rc) return rc;
    const size_t N = (size_t)g.NX * g.NY * g.NZ;
    for (int i = 0; i < g.NX; i++) {
        for (size_t jk = 0; jk < (size_t)g.NY * g.NZ; jk++) {
            const size_t c = (size_t)i * g.NY * g.NZ + jk;
            const double v2 = V[1 * N + c] * V[1 * N + c] + V[2 * N + c] * V[2 * N + c] + V[3 * N + c] * V[3 * N + c];
            const double b2 = V[4 * N + c] * V[4 * N + c] + V[5 * N + c] * V[5 * N + c] + V[6 * N + c] * V[6 * N + c];
            U[0 * N + c] = V[0 * N + c];
            U[1 * N + c] = V[0 * N + c] * V[1 * N + c];
            U[2 * N + c] = g.x[i] * V[0 * N + c] * V[2 * N + c];
            U[3 * N + c] = V[0 * N + c] * V[3 * N + c];
            U[4 * N + c] = V[4 * N + c];
            U[5 * N + c] = V[5 * N + c];
            U[6 * N + c] = V[6 * N + c];
            U[7 * N + c] = 0.5 * (V[0 * N + c] * v2 + b2) + p->oneogm1 * V[7 * N + c];
            U[8 * N + c] = V[8 * N + c];
        }
    }
    cyl_grid_free(&g);
    return CYLMHD_OK;
}

int cylmhd_check_host(const cylmhd_params *p, const double *rho_init, double *U, double *V, int counts[3])
{
    /* check.h:23-188 over every cell, ghosts included; rhofloor[c] =
     * rho_init[c] * rhocr as SaveInitial forms it (saveinitial.h:44), or 0
     * where no SaveInitial has run yet (the reference's global array is then
     * still zero-initialised, so the density floor cannot fire) */
    if (!p || !U || !V) return CYLMHD_EINVAL;
    cyl_grid g;
    int rc = cyl_grid_build(p, &g);
    if (rc) return rc;
    const size_t N = (size_t)g.NX * g.NY * g.NZ;
    const double oneobetacr = (1.0 / p->betacr), oneogm1 = p->oneogm1;
    int rhoflag = 0, prflag = 0, oneobetaflag = 0;
    for (int i = 0; i < g.NX; i++) {
        const double x = g.x[i];
        for (size_t jk = 0; jk < (size_t)g.NY * g.NZ; jk++) {
            const size_t c = (size_t)i * g.NY * g.NZ + jk;
            double *v0 = V + c, *v1 = V + N + c, *v3 = V + 3 * N + c, *v7 = V + 7 * N + c;
            const double v2 = V[2 * N + c] * V[2 * N + c];
            const double b2 = (V[4 * N + c] * V[4 * N + c] + V[5 * N + c] * V[5 * N + c] + V[6 * N + c] * V[6 * N + c]);
            const double rf = rho_init ? rho_init[c] * p->rhocr : 0.0;
            if (*v0 < rf) {
                rhoflag++;
                *v0 = rf; *v1 = 0.0; *v3 = 0.0;
                U[c] = *v0; U[N + c] = 0.0; U[2 * N + c] = x * *v0 * V[2 * N + c]; U[3 * N + c] = 0.0;
                U[7 * N + c] = 0.5 * (*v0 * v2 + b2) + oneogm1 * *v7;
            }
            const double prfloor = *v0 * p->prcr;
            if (*v7 < prfloor) {
                prflag++;
                *v1 = 0.0; *v3 = 0.0; *v7 = prfloor;
                U[N + c] = 0.0; U[3 * N + c] = 0.0;
                U[7 * N + c] = (0.5 * (*v0 * v2 + b2) + oneogm1 * *v7);
            }
            const double oneobeta = 0.5 * b2 / *v7;
            if (oneobeta > oneobetacr) {
                oneobetaflag++;
                *v0 = p->betacr * oneobeta * *v0; *v1 = 0.0; *v3 = 0.0;
                *v7 = p->betacr * oneobeta * *v7;
                U[c] = *v0; U[N + c] = 0.0; U[2 * N + c] = x * *v0 * V[2 * N + c]; U[3 * N + c] = 0.0;
                U[7 * N + c] = (0.5 * (*v0 * v2 + b2) + oneogm1 * *v7);
            }
        }
    }
    cyl_grid_free(&g);
    if (counts) { counts[0] = rhoflag; counts[1] = prflag; counts[2] = oneobetaflag; }
    return CYLMHD_OK;
}

/* ------------------------------------------------------------------------ */
static int write_raw(const char *dir, const char *name, const void *buf, size_t bytes)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *fo = fopen(path, "wb");
    if (!fo) return CYLMHD_EIO;
    size_t w = fwrite(buf, bytes, 1, fo);
    fclose(fo);
    return w == 1 ? CYLMHD_OK : CYLMHD_EIO;
}

int cylmhd_binary_output(const char *dir, int nx, int ny, int nz, double t,
                         const double *x, const double *y, const double *z, const double *V)
{
    /* output.h:1202-1414: directory, GridTime.dat, then 12 raw FP64 files */
    static const char *names[9] = {"rho", "vr", "vphi", "vz", "br", "bphi", "bz", "pr", "psi"};
    char path[4096];
    if (!dir || !x || !y || !z || !V || nx < 1 || ny < 1 || nz < 1) return CYLMHD_EINVAL;
    if (mkdir(dir, S_IRUSR | S_IWUSR | S_IXUSR | S_IRGRP | S_IWGRP | S_IXGRP | S_IROTH | S_IXOTH) != 0 &&
        errno != EEXIST)
        return CYLMHD_EIO;
    snprintf(path, sizeof path, "%s/GridTime.dat", dir);
    FILE *fo = fopen(path, "w");
    if (!fo) return CYLMHD_EIO;
    fprintf(fo, "%d\t%d\t%d\t%.3e", nx, ny, nz, t);
    fclose(fo);
    const size_t N = (size_t)nx * ny * nz;
    int rc;
    if ((rc = write_raw(dir, "r", x, sizeof(double) * nx))) return rc;
    if ((rc = write_raw(dir, "phi", y, sizeof(double) * ny))) return rc;
    if ((rc = write_raw(dir, "z", z, sizeof(double) * nz))) return rc;
    for (int m = 0; m < 9; m++)
        if ((rc = write_raw(dir, names[m], V + (size_t)m * N, sizeof(double) * N))) return rc;
    return CYLMHD_OK;
}

int cylmhd_restart_read(const char *dir, int nx, int ny, int nz, double *V)
{
    /* restartread.h:16-139: rho, vr, vphi, vz, br, bphi, bz, pr, psi */
    static const char *names[9] = {"rho", "vr", "vphi", "vz", "br", "bphi", "bz", "pr", "psi"};
    char path[4096];
    if (!dir || !V || nx < 1 || ny < 1 || nz < 1) return CYLMHD_EINVAL;
    const size_t N = (size_t)nx * ny * nz;
    for (int m = 0; m < 9; m++) {
        snprintf(path, sizeof path, "%s/%s", dir, names[m]);
        FILE *fi = fopen(path, "rb");
        if (!fi) return CYLMHD_EIO;
        size_t r = fread(V + (size_t)m * N, sizeof(double) * N, 1, fi);
        fclose(fi);
        if (r != 1) return CYLMHD_EIO;
    }
    return CYLMHD_OK;
}
