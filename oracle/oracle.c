/*
 * oracle.c -- CPU FP64 restatement of the reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the GPU library
 * and the timed CPU baseline.  Build: oracle/Makefile (gcc -O2 -fopenmp
 * -ffp-contract=off, the flags that keep x86-64 arithmetic identical to the
 * reference compiled the same way).
 *
 * Every routine cites the reference file:line it restates.  Expressions are
 * written in the reference's evaluation order (C left-to-right association)
 * so results are bit-identical to the reference's on the same inputs; the
 * only documented deviations are:
 *   - cot(pitch) := 1/tan(pitch) (the reference calls an undeclared cot());
 *   - racy OpenMP reductions are replaced by exact ones (Setdt keeps the
 *     first maximum in (i,j,k) order, the serial answer; floor counters are
 *     exact);
 *   - the seeded initial state comes from the caller (no rand()).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* maxmin.h:16-26 */
static inline double rmax(double a, double b) { if (a > b) return a; else return b; }
static inline double rmin(double a, double b) { if (a > b) return b; else return a; }

struct orc {
    cylmhd_params p;
    int NX, NY, NZ, irin;
    size_t N, NF;             /* cells, faces per component */
    double *x, *xm, *dx, *dxm, *oneox, *oneoxm, *oneodx;
    double *y, dy, oneody;
    double *z, *zm, *dz, *dzm, *oneodz;
    double *g1, *g2, *g3, *g4, *g5, *g6, *dvolume, *minlength;   /* 3-D */
    double minminlength;
    double xabs[11], zabs[11], *cabs;
    double *U, *U1, *V, *S, *F, *G, *H;
    double *Vl, *Vr;
    double *Grx, *Gry, *Grz, *Grxaxi, *Gryaxi, *Grzaxi;
    double *Uinit, *Vinit, *rhofloor;
    double dt, ch, psicoef;
    int counts[9];
};

#define C3(o, i, j, k) ((((size_t)(i)) * (o)->NY + (size_t)(j)) * (o)->NZ + (size_t)(k))
#define F3(o, i, j, k) ((((size_t)(i)) * ((o)->NY - 1) + (size_t)(j)) * ((o)->NZ - 1) + (size_t)(k))

int orc_threads(void) { return omp_get_max_threads(); }

static double *alloc0(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

void orc_destroy(orc *o)
{
    if (!o) return;
    double *ptrs[] = {o->x, o->xm, o->dx, o->dxm, o->oneox, o->oneoxm, o->oneodx, o->y, o->z,
                      o->zm, o->dz, o->dzm, o->oneodz, o->g1, o->g2, o->g3, o->g4, o->g5, o->g6,
                      o->dvolume, o->minlength, o->cabs, o->U, o->U1, o->V, o->S, o->F, o->G,
                      o->H, o->Vl, o->Vr, o->Grx, o->Gry, o->Grz, o->Grxaxi, o->Gryaxi,
                      o->Grzaxi, o->Uinit, o->Vinit, o->rhofloor};
    for (size_t n = 0; n < sizeof ptrs / sizeof ptrs[0]; ++n) free(ptrs[n]);
    free(o);
}

/* CylindricalGrid (cylindricalgrid.h:25-227) */
static void build_grid(orc *o)
{
    const cylmhd_params *p = &o->p;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ;
    int i, j, k;
    o->dx[0] = p->dx0;
    o->x[0] = p->x0 - p->dx0;
    for (i = 0; i < NX - 1; i++) {
        o->x[i + 1] = o->x[i] + o->dx[i];
        o->dx[i + 1] = p->stretchratio * o->dx[i];
        if (o->x[i] > p->lx) o->dx[i + 1] = rmin(p->stretchratio * o->dx[i], p->dxmax);
        else o->dx[i + 1] = o->dx[i];
        o->xm[i] = (o->x[i] + o->x[i + 1]) * 0.5;
        o->dxm[i] = (o->dx[i] + o->dx[i + 1]) * 0.5;
        o->oneox[i] = 1.0 / o->x[i];
        o->oneoxm[i] = 1.0 / o->xm[i];
        o->oneodx[i] = 1.0 / o->dx[i];
    }
    o->xm[NX - 1] = o->x[NX - 1] + 0.5 * o->dx[NX - 1];
    o->dxm[NX - 1] = o->dx[NX - 1];
    o->oneox[NX - 1] = 1.0 / o->x[NX - 1];
    o->oneodx[NX - 1] = 1.0 / o->dx[NX - 1];

    o->dy = p->dy0;
    o->y[0] = p->y0 - o->dy;
    o->oneody = 1.0 / o->dy;
    for (j = 0; j < NY - 1; j++) o->y[j + 1] = o->y[j] + o->dy;

    o->dz[0] = p->dz0;
    o->z[0] = p->z0 - p->dz0;
    for (k = 0; k < NZ - 1; k++) {
        o->z[k + 1] = o->z[k] + o->dz[k];
        if (o->z[k] > p->lz) o->dz[k + 1] = rmin(p->stretchratio * o->dz[k], p->dzmax);
        else o->dz[k + 1] = o->dz[k];
        o->zm[k] = (o->z[k] + o->z[k + 1]) * 0.5;
        o->dzm[k] = (o->dz[k] + o->dz[k + 1]) * 0.5;
        o->oneodz[k] = 1.0 / o->dz[k];
    }
    o->zm[NZ - 1] = o->z[NZ - 1] + 0.5 * o->dz[NZ - 1];
    o->oneodz[NZ - 1] = 1.0 / o->dz[NZ - 1];

#pragma omp parallel for private(j, k)
    for (i = 1; i < NX - 1; i++) {
        for (j = 1; j < NY - 1; j++) {
            for (k = 1; k < NZ - 1; k++) {
                const size_t c = C3(o, i, j, k);
                double oneodvol;
                o->dvolume[c] = 0.5 * (o->xm[i] * o->xm[i] - o->xm[i - 1] * o->xm[i - 1]) * o->dy * o->dzm[k - 1];
                oneodvol = 1.0 / o->dvolume[c];
                o->g1[c] = o->xm[i - 1] * o->dy * o->dzm[k - 1] * oneodvol;
                o->g2[c] = o->xm[i] * o->dy * o->dzm[k - 1] * oneodvol;
                o->g3[c] = o->dxm[i - 1] * o->dzm[k - 1] * oneodvol;
                o->g4[c] = o->g3[c];
                o->g5[c] = 0.5 * (o->xm[i] * o->xm[i] - o->xm[i - 1] * o->xm[i - 1]) * o->dy * oneodvol;
                o->g6[c] = o->g5[c];
            }
        }
    }
    double oneominminlength = 0.0;
    for (i = 0; i < NX; i++)
        for (j = 0; j < NY; j++)
            for (k = 0; k < NZ; k++) {
                const size_t c = C3(o, i, j, k);
                double oneominlength;
                o->minlength[c] = rmin(o->dx[i], rmin(o->x[i] * o->dy, o->dz[k]));
                oneominlength = 1.0 / o->minlength[c];
                if (oneominlength > oneominminlength) oneominminlength = oneominlength;
            }
    o->minminlength = 1.0 / oneominminlength;

    for (i = NX - 1 - 10; i < NX; i++)
        o->xabs[i - (NX - 1 - 10)] =
            (0.5 * (1.0 - tanh((o->x[NX - 1] - o->x[i] - 5.0 * o->dx[NX - 1]) / (2.0 * o->dx[NX - 1]))));
    for (k = NZ - 1 - 10; k < NZ; k++)
        if (k >= 0)
            o->zabs[k - (NZ - 1 - 10)] =
                (0.5 * (1.0 - tanh((o->z[NZ - 1] - o->z[k] - 5.0 * o->dz[NZ - 1]) / (2.0 * o->dz[NZ - 1]))));
    for (i = 0; i < o->irin + 1; i++)
        for (k = 0; k < NZ; k++) {
            double r = sqrt(o->x[i] * o->x[i] + o->z[k] * o->z[k]);
            o->cabs[(size_t)i * NZ + k] = (0.5 * (1.0 - tanh((r - p->rin + 5.0 * p->dx0) / (2.0 * p->dx0))));
        }
}

/* InitialGravity (initialgravity.h:22-79) or a point mass; all cells */
static void build_gravity(orc *o)
{
    const cylmhd_params *p = &o->p;
    int i, j, k;
#pragma omp parallel for private(j, k)
    for (i = 0; i < o->NX; i++)
        for (j = 0; j < o->NY; j++)
            for (k = 0; k < o->NZ; k++) {
                const size_t c = C3(o, i, j, k);
                const double x2 = o->x[i] * o->x[i], z2 = o->z[k] * o->z[k];
                if (p->gravity == CYLMHD_GRAV_GALACTIC) {
                    const double vhalo0 = 131.5e+5 / 6.558e+6, dh = 1.2;
                    const double a = 0.65, b = 0.026, md = 10.0;
                    const double db = 0.07, mb = 3.4;
                    o->Grxaxi[c] = ((-2.0 * vhalo0 * vhalo0 / (x2 + z2 + dh * dh)
                                     - md / pow(x2 + (a + sqrt(z2 + b * b)) * (a + sqrt(z2 + b * b)), 1.5)
                                     - mb / ((sqrt(x2 + z2) + db) * (sqrt(x2 + z2) + db) * sqrt(x2 + z2)))
                                    * o->x[i]);
                    o->Gryaxi[c] = 0.0;
                    o->Grzaxi[c] = ((-2.0 * vhalo0 * vhalo0 / (x2 + z2 + dh * dh)
                                     - md / pow(x2 + (a + sqrt(z2 + b * b)) * (a + sqrt(z2 + b * b)), 1.5)
                                           * (1.0 + a / sqrt(z2 + b * b))
                                     - mb / ((sqrt(x2 + z2) + db) * (sqrt(x2 + z2) + db) * sqrt(x2 + z2)))
                                    * o->z[k]);
                } else {
                    const double r2 = x2 + z2;
                    const double inv = 1.0 / (r2 * sqrt(r2));
                    o->Grxaxi[c] = -p->gm_point * o->x[i] * inv;
                    o->Gryaxi[c] = 0.0;
                    o->Grzaxi[c] = -p->gm_point * o->z[k] * inv;
                }
                o->Grx[c] = o->Grxaxi[c];
                o->Gry[c] = o->Gryaxi[c];
                o->Grz[c] = o->Grzaxi[c];
            }
}

orc *orc_create(const cylmhd_params *p)
{
    if (!p || p->ix < 11 || p->jy < 2 || (p->jy & 1) || p->kz < 2) return NULL;
    orc *o = (orc *)calloc(1, sizeof(orc));
    if (!o) return NULL;
    o->p = *p;
    o->NX = p->ix + 2; o->NY = p->jy + 2; o->NZ = p->kz + 2;
    o->irin = (int)(p->rin / p->dx0);
    o->N = (size_t)o->NX * o->NY * o->NZ;
    o->NF = (size_t)(o->NX - 1) * (o->NY - 1) * (o->NZ - 1);
    const size_t N = o->N, NF = o->NF;
    o->x = alloc0(o->NX); o->xm = alloc0(o->NX); o->dx = alloc0(o->NX); o->dxm = alloc0(o->NX);
    o->oneox = alloc0(o->NX); o->oneoxm = alloc0(o->NX); o->oneodx = alloc0(o->NX);
    o->y = alloc0(o->NY);
    o->z = alloc0(o->NZ); o->zm = alloc0(o->NZ); o->dz = alloc0(o->NZ); o->dzm = alloc0(o->NZ);
    o->oneodz = alloc0(o->NZ);
    o->g1 = alloc0(N); o->g2 = alloc0(N); o->g3 = alloc0(N); o->g4 = alloc0(N); o->g5 = alloc0(N);
    o->g6 = alloc0(N); o->dvolume = alloc0(N); o->minlength = alloc0(N);
    o->cabs = alloc0((size_t)(o->irin + 1) * o->NZ);
    o->U = alloc0(9 * N); o->U1 = alloc0(9 * N); o->V = alloc0(9 * N); o->S = alloc0(9 * N);
    o->F = alloc0(9 * NF); o->G = alloc0(9 * NF); o->H = alloc0(9 * NF);
    o->Vl = alloc0(9 * NF); o->Vr = alloc0(9 * NF);
    o->Grx = alloc0(N); o->Gry = alloc0(N); o->Grz = alloc0(N);
    o->Grxaxi = alloc0(N); o->Gryaxi = alloc0(N); o->Grzaxi = alloc0(N);
    o->Uinit = alloc0(9 * N); o->Vinit = alloc0(9 * N); o->rhofloor = alloc0(N);
    if (!o->x || !o->xm || !o->dx || !o->dxm || !o->oneox || !o->oneoxm || !o->oneodx || !o->y ||
        !o->z || !o->zm || !o->dz || !o->dzm || !o->oneodz || !o->g1 || !o->g2 || !o->g3 ||
        !o->g4 || !o->g5 || !o->g6 || !o->dvolume || !o->minlength || !o->cabs || !o->U ||
        !o->U1 || !o->V || !o->S || !o->F || !o->G || !o->H || !o->Vl || !o->Vr || !o->Grx ||
        !o->Gry || !o->Grz || !o->Grxaxi || !o->Gryaxi || !o->Grzaxi || !o->Uinit ||
        !o->Vinit || !o->rhofloor) {
        orc_destroy(o);
        return NULL;
    }
    build_grid(o);
    build_gravity(o);
    return o;
}

void orc_dims(const orc *o, int *nx, int *ny, int *nz) { *nx = o->NX; *ny = o->NY; *nz = o->NZ; }

void orc_grid(const orc *o, double *x, double *xm, double *y, double *z, double *mml)
{
    if (x) memcpy(x, o->x, sizeof(double) * o->NX);
    if (xm) memcpy(xm, o->xm, sizeof(double) * o->NX);
    if (y) memcpy(y, o->y, sizeof(double) * o->NY);
    if (z) memcpy(z, o->z, sizeof(double) * o->NZ);
    if (mml) *mml = o->minminlength;
}

/* dzm[k] = 0.5 (dz[k] + dz[k+1]) (cylindricalgrid.h:94): the height of cell k+1 */
void orc_dzm(const orc *o, double *dzm)
{
    memcpy(dzm, o->dzm, sizeof(double) * o->NZ);
}

void orc_geometry(const orc *o, double *out)
{
    const size_t T = (size_t)o->NX * o->NZ;
    const int j = 1;
    for (int i = 0; i < o->NX; i++)
        for (int k = 0; k < o->NZ; k++) {
            const size_t c = C3(o, i, j, k), t = (size_t)i * o->NZ + k;
            out[0 * T + t] = o->g1[c];
            out[1 * T + t] = o->g2[c];
            out[2 * T + t] = o->g3[c];
            out[3 * T + t] = o->g5[c];
            out[4 * T + t] = o->dvolume[c];
            out[5 * T + t] = o->minlength[c];
        }
}

int orc_absorb_factors(const orc *o, double *xabs, double *zabs, double *cabs)
{
    if (xabs) memcpy(xabs, o->xabs, sizeof o->xabs);
    if (zabs) memcpy(zabs, o->zabs, sizeof o->zabs);
    if (cabs) memcpy(cabs, o->cabs, sizeof(double) * (size_t)(o->irin + 1) * o->NZ);
    return o->irin;
}

void orc_gravity_axi(const orc *o, double *grx, double *grz)
{
    for (int i = 0; i < o->NX; i++)
        for (int k = 0; k < o->NZ; k++) {
            const size_t c = C3(o, i, 1, k);
            grx[(size_t)i * o->NZ + k] = o->Grxaxi[c];
            grz[(size_t)i * o->NZ + k] = o->Grzaxi[c];
        }
}

/* SaveInitial (saveinitial.h:18-50) */
void orc_set_initial(orc *o, const double *U, const double *V)
{
    orc_set_state(o, U, V);
    memcpy(o->Uinit, U, sizeof(double) * 9 * o->N);
    memcpy(o->Vinit, V, sizeof(double) * 9 * o->N);
    for (size_t c = 0; c < o->N; c++) o->rhofloor[c] = o->U[c] * o->p.rhocr;
}

void orc_set_state(orc *o, const double *U, const double *V)
{
    memcpy(o->U, U, sizeof(double) * 9 * o->N);
    memcpy(o->V, V, sizeof(double) * 9 * o->N);
}

void orc_get_state(const orc *o, double *U, double *V)
{
    if (U) memcpy(U, o->U, sizeof(double) * 9 * o->N);
    if (V) memcpy(V, o->V, sizeof(double) * 9 * o->N);
}

void orc_get_gravity(const orc *o, double *grx, double *gry, double *grz)
{
    if (grx) memcpy(grx, o->Grx, sizeof(double) * o->N);
    if (gry) memcpy(gry, o->Gry, sizeof(double) * o->N);
    if (grz) memcpy(grz, o->Grz, sizeof(double) * o->N);
}

void orc_get_fluxes(const orc *o, double *out)
{
    memcpy(out, o->F, sizeof(double) * 9 * o->NF);
    memcpy(out + 9 * o->NF, o->G, sizeof(double) * 9 * o->NF);
    memcpy(out + 18 * o->NF, o->H, sizeof(double) * 9 * o->NF);
}

void orc_floor_counts(const orc *o, int counts[9]) { memcpy(counts, o->counts, sizeof o->counts); }

/* Cooling (cooling.h:22-60): Raymond (1978) fit returned through an int */
double orc_cooling(const cylmhd_params *p, double dummyrho, double dummypr)
{
    if (p->cooling == CYLMHD_COOL_RAYMOND) {
        double temperature = dummypr / dummyrho;
        double tempma, fabstempma, tempmb, fabstempmb;
        double tempmc, fabstempmc, tempmd, fabstempmd;
        double signa, signb, signc, signd;
        double A, B, C;
        double a = 0.031, b = 0.054, c = 0.078, d = 0.235;
        tempma = temperature - a; fabstempma = fabs(tempma);
        tempmb = temperature - b; fabstempmb = fabs(tempmb);
        tempmc = temperature - c; fabstempmc = fabs(tempmc);
        tempmd = temperature - d; fabstempmd = fabs(tempmd);
        signa = tempma / fabstempma;
        signb = tempmb / fabstempmb;
        signc = tempmc / fabstempmc;
        signd = tempmd / fabstempmd;
        A = 0.108e-37 * pow(p->te0 * temperature, 10.26);
        B = 0.0;
        C = 0.0;
        return (double)(int)(0.5 * (A * (1.0 - signa * signb) + B * (1.0 - signb * signc) + C * (1.0 - signc * signd)));
    }
    if (p->cooling == CYLMHD_COOL_POWERLAW) {
        double temperature = dummypr / dummyrho;
        if (temperature > p->cool_tmin)
            return p->cool_index == 0.5 ? p->cool_lambda0 * sqrt(temperature)
                                        : p->cool_lambda0 * pow(temperature, p->cool_index);
        return 0.0;
    }
    return 0.0;
}

/* Setdt (setdt.h:24-94) */
void orc_setdt(orc *o, double *dt, double *ch, int ijk[3])
{
    const cylmhd_params *p = &o->p;
    const size_t N = o->N;
    const double *V = o->V;
    const int nth = omp_get_max_threads();
    double *best = (double *)calloc(nth, sizeof(double));
    long *arg = (long *)calloc(nth, sizeof(long));
    const int ilast = o->NX - 1 - 10;
#pragma omp parallel
    {
        const int t = omp_get_thread_num();
        double oneocfldt = 0.0;
        long a = -1;
#pragma omp for schedule(static)
        for (int i = 1; i < ilast; i++)
            for (int j = 1; j < o->NY - 1; j++)
                for (int k = 1; k < o->NZ - 1; k++) {
                    const size_t c = C3(o, i, j, k);
                    double v2, b2, temponeocfldt, temponeocfldt1, temponeotcool;
                    v2 = V[1 * N + c] * V[1 * N + c] + V[2 * N + c] * V[2 * N + c] + V[3 * N + c] * V[3 * N + c];
                    b2 = V[4 * N + c] * V[4 * N + c] + V[5 * N + c] * V[5 * N + c] + V[6 * N + c] * V[6 * N + c];
                    temponeocfldt1 = sqrt(v2 + (p->gm * V[7 * N + c] + b2) / V[0 * N + c]) / o->minlength[c];
                    temponeotcool = (2.0 * ((V[0 * N + c] * V[0 * N + c] * orc_cooling(p, V[0 * N + c], V[7 * N + c]))
                                            / (1.5 * V[7 * N + c])));
                    temponeocfldt = rmax(temponeocfldt1, temponeotcool);
                    if (temponeocfldt > oneocfldt) { oneocfldt = temponeocfldt; a = (long)c; }
                }
        best[t] = oneocfldt;
        arg[t] = a;
    }
    /* static schedule: thread t owns a contiguous i-range in order, so the
     * first strict maximum over threads in order is the serial answer */
    double m = 0.0;
    long am = -1;
    for (int t = 0; t < nth; t++)
        if (best[t] > m) { m = best[t]; am = arg[t]; }
    free(best); free(arg);
    if (am >= 0) {
        ijk[0] = (int)(am / ((long)o->NY * o->NZ));
        ijk[1] = (int)((am / o->NZ) % o->NY);
        ijk[2] = (int)(am % o->NZ);
    } else {
        ijk[0] = ijk[1] = ijk[2] = -1;
    }
    *dt = p->cfl / m;
    *ch = o->minminlength * m;
}

/* SpiralGravity (spiralgravity.h:20-124) */
void orc_gravity(orc *o, double t)
{
    const cylmhd_params *p = &o->p;
    double m[3] = {0.0, 10.0, 0.0};
    double a[3] = {0.0, 0.65, 0.0};
    double b[3] = {0.0, 0.026, 0.0};
    double oneoscalephase = 1.0 / p->scalephase;
    double oneoscaler = 1.0 / p->scaler;
    double cotpitch = 1.0 / tan(p->pitch);
    double switchfunc = (t - p->tausp) / fabs(t - p->tausp);
    double epssp = 0.5 * p->epssp0 * (t / p->tausp * (1.0 - switchfunc) + (1.0 + switchfunc));
    const int armnum = p->armnum;
    int i;
#pragma omp parallel for
    for (i = 0; i < o->NX; i++) {
        double x2 = o->x[i] * o->x[i];
        double xoscalephase = o->x[i] * oneoscalephase;
        double xoscaler = o->x[i] * oneoscaler;
        double xoscaler2 = xoscaler * oneoscaler;
        double x2oscaler2 = xoscaler * xoscaler;
        double oneoonex2oscaler2 = 1.0 / (1.0 + x2oscaler2);
        double oneoonex2oscaler215 = 1.0 * pow(oneoonex2oscaler2, 1.5);
        double xoscaler2oonex2oscaler215 = xoscaler2 * oneoonex2oscaler215;
        double x2oscaler2oonex2oscaler215 = x2oscaler2 * oneoonex2oscaler215;
        for (int j = 0; j < o->NY; j++) {
            double phase = armnum * (o->y[j] - (p->omegapattern - p->omegasp) * t + cotpitch * log(xoscalephase));
            double sinphase, cosphase;
            sincos(phase, &sinphase, &cosphase);
            for (int k = 0; k < o->NZ; k++) {
                const size_t c = C3(o, i, j, k);
                double z2 = o->z[k] * o->z[k];
                double a1zk2b12 = ((a[1] + sqrt(z2 + b[1] * b[1])) * (a[1] + sqrt(z2 + b[1] * b[1])));
                double oneosqrtx2a1zk2b12 = 1.0 / sqrt(x2 + a1zk2b12);
                double scalezosqrtz2scalez2 = (p->scalez / sqrt(z2 + p->scalez * p->scalez));
                double spiralfr = (-epssp * m[1] * oneosqrtx2a1zk2b12 * scalezosqrtz2scalez2
                                   * xoscaler2oonex2oscaler215
                                   * ((o->x[i] * oneosqrtx2a1zk2b12 * oneosqrtx2a1zk2b12
                                       - ((2.0 - x2oscaler2) * oneoonex2oscaler2)) * cosphase
                                      + armnum * cotpitch * sinphase));
                double spiralfphi = (epssp * m[1] * oneosqrtx2a1zk2b12 * scalezosqrtz2scalez2
                                     * xoscaler2oonex2oscaler215 * armnum * sinphase);
                double spiralfz = (-epssp * m[1] * oneosqrtx2a1zk2b12 * o->z[k] * scalezosqrtz2scalez2
                                   * x2oscaler2oonex2oscaler215
                                   * ((1.0 + a[1] / sqrt(z2 + b[1] * b[1])) * oneosqrtx2a1zk2b12 * oneosqrtx2a1zk2b12
                                      + 1.0 / (z2 + p->scalez * p->scalez)) * cosphase);
                o->Grx[c] = o->Grxaxi[c] + spiralfr;
                o->Gry[c] = o->Gryaxi[c] + spiralfphi;
                o->Grz[c] = o->Grzaxi[c] + spiralfz;
            }
        }
    }
}

/* GLMSourceSolver (glmsourcesolver.h:19-47) */
static void glm_source(orc *o, double *dU, double *dV)
{
    const size_t N = o->N;
    size_t c;
#pragma omp parallel for
    for (c = 0; c < N; c++) {
        dU[8 * N + c] = o->psicoef * dU[8 * N + c];
        dV[8 * N + c] = o->psicoef * dV[8 * N + c];
    }
}

/* MinmodLimiter{X,Y,Z} (limiterx.h:171-217, limitery.h:46-94,
 * limiterz.h:146-190), MCLimiter{X,Y,Z} (limiterx.h:335-385,
 * limitery.h:197-247, limiterz.h:294-358), VanLeerLimiter{X,Y,Z}
 * (limiterx.h:266-326, limitery.h:129-188, limiterz.h:232-285) and
 * SuperbeeLimiter{X,Y,Z} (limiterx.h:394-449, limitery.h:256-311,
 * limiterz.h:367-432), each followed by the interface-state boundaries it
 * calls. */
static void limiter(orc *o, int dir)
{
    const cylmhd_params *p = &o->p;
    const size_t N = o->N, NF = o->NF;
    const double *V = o->V;
    double *Vl = o->Vl, *Vr = o->Vr;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ;
    const int lim = p->limiter;
    const int ilo = dir == 0 ? 1 : 0, jlo = dir == 1 ? 1 : 0, klo = dir == 2 ? 1 : 0;
    const size_t st = dir == 0 ? (size_t)NY * NZ : (dir == 1 ? (size_t)NZ : 1);
    const int nb = 4 + dir, nv = 1 + dir;
    int i;
#pragma omp parallel for
    for (i = ilo; i < NX - 1; i++) {
        for (int j = jlo; j < NY - 1; j++) {
            for (int k = klo; k < NZ - 1; k++) {
                const size_t c = C3(o, i, j, k);
                const size_t fl = F3(o, i, j, k);
                const size_t fr = dir == 0 ? F3(o, i - 1, j, k) : (dir == 1 ? F3(o, i, j - 1, k) : F3(o, i, j, k - 1));
                double lambmax = 0.0, lambmin = 0.0;
                if (lim == CYLMHD_LIMITER_MINMOD) {
                    double b2 = (V[4 * N + c] * V[4 * N + c] + V[5 * N + c] * V[5 * N + c] + V[6 * N + c] * V[6 * N + c]);
                    double gmpr = p->gm * V[7 * N + c];
                    double fast = sqrt(((b2 + gmpr) + sqrt((b2 + gmpr) * (b2 + gmpr) - 4.0 * gmpr * V[nb * N + c] * V[nb * N + c]))
                                       / (2.0 * V[0 * N + c]));
                    lambmax = rmax(0, V[nv * N + c] + fast);
                    lambmin = rmin(0, V[nv * N + c] - fast);
                }
                for (int m = 0; m < 9; m++) {
                    const double *Vm = V + (size_t)m * N;
                    double a = Vm[c] - Vm[c - st];
                    double b = Vm[c + st] - Vm[c];
                    double grad;
                    if (lim == CYLMHD_LIMITER_MC) {
                        double cc = 0.25 * (Vm[c + st] - Vm[c - st]);
                        if (a * b <= 0.0) grad = 0.0;
                        else {
                            if (a > 0.0) grad = rmin(a, rmin(b, cc));
                            else grad = rmax(a, rmax(b, cc));
                        }
                        Vl[m * NF + fl] = Vm[c] + grad;
                        Vr[m * NF + fr] = Vm[c] - grad;
                    } else if (lim == CYLMHD_LIMITER_VANLEER) {
                        if (a * b <= 0.0) grad = 0.0;
                        else grad = a * b / (a + b);
                        Vl[m * NF + fl] = Vm[c] + grad;
                        Vr[m * NF + fr] = Vm[c] - grad;
                    } else if (lim == CYLMHD_LIMITER_SUPERBEE) {
                        if (a * b <= 0.0) grad = 0.0;
                        else {
                            if (fabs(a) >= fabs(b)) {
                                if (a > 0.0) grad = rmin(a, 2.0 * b);
                                else grad = rmax(a, 2.0 * b);
                            } else {
                                if (a > 0.0) grad = rmin(2.0 * a, b);
                                else grad = rmax(2.0 * a, b);
                            }
                        }
                        Vl[m * NF + fl] = Vm[c] + 0.5 * grad;
                        Vr[m * NF + fr] = Vm[c] - 0.5 * grad;
                    } else {
                        if (a * b <= 0.0) grad = 0.0;
                        else {
                            if (a > 0.0) grad = rmin(a, b);
                            else grad = rmax(a, b);
                        }
                        if (dir == 0) {
                            Vl[m * NF + fl] = (Vm[c] + 0.5 * (1.0 - lambmax * o->dt * o->oneodx[i]) * grad);
                            Vr[m * NF + fr] = (Vm[c] - 0.5 * (1.0 + lambmin * o->dt * o->oneodx[i - 1]) * grad);
                        } else if (dir == 1) {
                            Vl[m * NF + fl] = (Vm[c] + 0.5 * (1.0 - lambmax * o->dt * o->oneox[i] * o->oneody) * grad);
                            Vr[m * NF + fr] = (Vm[c] - 0.5 * (1.0 + lambmin * o->dt * o->oneox[i] * o->oneody) * grad);
                        } else {
                            Vl[m * NF + fl] = (Vm[c] + 0.5 * (1.0 - lambmax * o->dt * o->oneodz[k]) * grad);
                            Vr[m * NF + fr] = (Vm[c] - 0.5 * (1.0 + lambmin * o->dt * o->oneodz[k - 1]) * grad);
                        }
                    }
                }
            }
        }
    }
    /* interface-state boundaries */
    if (dir == 0) {
        if (lim == CYLMHD_LIMITER_VANLEER) {
            /* XVlrFreeBoundary (xvlrboundary.h:189-216) */
            for (int m = 0; m < 9; m++)
                for (int j = 0; j < NY - 1; j++)
                    for (int k = 0; k < NZ - 1; k++) {
                        Vl[m * NF + F3(o, 0, j, k)] = Vr[m * NF + F3(o, 0, j, k)];
                        Vr[m * NF + F3(o, NX - 2, j, k)] = Vl[m * NF + F3(o, NX - 2, j, k)];
                    }
        } else {
            /* XVlrPeriodicBoundary (xvlrboundary.h:153-180) */
            for (int m = 0; m < 9; m++)
                for (int j = 0; j < NY - 1; j++)
                    for (int k = 0; k < NZ - 1; k++) {
                        Vl[m * NF + F3(o, 0, j, k)] = Vl[m * NF + F3(o, NX - 2, j, k)];
                        Vr[m * NF + F3(o, NX - 2, j, k)] = Vr[m * NF + F3(o, 0, j, k)];
                    }
        }
    } else if (dir == 1) {
        /* YVlrPeriodicBoundary (yvlrboundary.h:61-87), every limiter */
        for (int m = 0; m < 9; m++)
            for (int ii = 0; ii < NX - 1; ii++)
                for (int k = 0; k < NZ - 1; k++) {
                    Vl[m * NF + F3(o, ii, 0, k)] = Vl[m * NF + F3(o, ii, NY - 2, k)];
                    Vr[m * NF + F3(o, ii, NY - 2, k)] = Vr[m * NF + F3(o, ii, 0, k)];
                }
        if (lim == CYLMHD_LIMITER_VANLEER)
            /* then YVlrInitializeBoundary (yvlrboundary.h:26-52) */
            for (int m = 0; m < 9; m++)
                for (int ii = 0; ii < NX - 1; ii++)
                    for (int k = 0; k < NZ - 1; k++) {
                        Vl[m * NF + F3(o, ii, 0, k)] = o->Vinit[m * N + C3(o, ii, 0, k)];
                        Vr[m * NF + F3(o, ii, NY - 2, k)] = o->Vinit[m * N + C3(o, ii, NY - 1, k)];
                    }
    } else if (lim == CYLMHD_LIMITER_MINMOD) {
        /* ZLeftVlVzAntisymBzsymBoundary (zvlrboundary.h:63-92) then
         * ZRightVrFreeBoundary (zvlrboundary.h:28-54) */
        for (int ii = 0; ii < NX - 1; ii++)
            for (int j = 0; j < NY - 1; j++)
                for (int m = 0; m < 9; m++)
                    Vl[m * NF + F3(o, ii, j, 0)] = (m == 3 ? -Vr[m * NF + F3(o, ii, j, 1)] : Vr[m * NF + F3(o, ii, j, 1)]);
        for (int m = 0; m < 9; m++)
            for (int ii = 0; ii < NX - 1; ii++)
                for (int j = 0; j < NY - 1; j++)
                    Vr[m * NF + F3(o, ii, j, NZ - 2)] = Vl[m * NF + F3(o, ii, j, NZ - 2)];
    } else if (lim == CYLMHD_LIMITER_VANLEER) {
        /* ZVlrFreeBoundary (zvlrboundary.h:137-164) */
        for (int m = 0; m < 9; m++)
            for (int ii = 0; ii < NX - 1; ii++)
                for (int j = 0; j < NY - 1; j++) {
                    Vl[m * NF + F3(o, ii, j, 0)] = Vr[m * NF + F3(o, ii, j, 0)];
                    Vr[m * NF + F3(o, ii, j, NZ - 2)] = Vl[m * NF + F3(o, ii, j, NZ - 2)];
                }
    } else if (lim == CYLMHD_LIMITER_SUPERBEE) {
        /* ZVlrPeriodicBoundary (zvlrboundary.h:101-128) */
        for (int m = 0; m < 9; m++)
            for (int ii = 0; ii < NX - 1; ii++)
                for (int j = 0; j < NY - 1; j++) {
                    Vl[m * NF + F3(o, ii, j, 0)] = Vl[m * NF + F3(o, ii, j, NZ - 2)];
                    Vr[m * NF + F3(o, ii, j, NZ - 2)] = Vr[m * NF + F3(o, ii, j, 0)];
                }
    } else {
        /* ZVlrFixedBoundary (zvlrboundary.h:173-200) */
        for (int m = 0; m < 9; m++)
            for (int ii = 0; ii < NX - 1; ii++)
                for (int j = 0; j < NY - 1; j++) {
                    Vl[m * NF + F3(o, ii, j, 0)] = o->Vinit[m * N + C3(o, ii, j, 0)];
                    Vr[m * NF + F3(o, ii, j, NZ - 2)] = o->Vinit[m * N + C3(o, ii, j, NZ - 1)];
                }
    }
}

/* VlrtoUlr (vlrtoulr.h:25-43) */
static void vlr_to_ulr(double v2, double b2, double r, const double *Vx, double *Ux, double oneogm1)
{
    Ux[0] = Vx[0];
    Ux[1] = Vx[0] * Vx[1];
    Ux[2] = r * Vx[0] * Vx[2];
    Ux[3] = Vx[0] * Vx[3];
    Ux[4] = Vx[4];
    Ux[5] = Vx[5];
    Ux[6] = Vx[6];
    Ux[7] = 0.5 * (Vx[0] * v2 + b2) + oneogm1 * Vx[7];
}

/* VlrUlrto{F,G,H}lr (vlrulrtoflr.h:24-51, vlrulrtoglr.h:25-53,
 * vlrulrtohlr.h:24-52) */
static void phys_flux(const orc *o, int dir, int i, double b2, double vdotb, const double *Vx,
                      const double *Ux, double *Fx)
{
    if (dir == 0) {
        const double xm = o->xm[i];
        Fx[0] = Ux[1];
        Fx[1] = Ux[1] * Vx[1] + (Vx[7] + 0.5 * b2) - Vx[4] * Vx[4];
        Fx[2] = Ux[2] * Vx[1] - xm * Vx[5] * Vx[4];
        Fx[3] = Ux[3] * Vx[1] - Vx[6] * Vx[4];
        Fx[4] = 0.0;
        Fx[5] = Vx[5] * Vx[1] - Vx[2] * Vx[4];
        Fx[6] = Vx[6] * Vx[1] - Vx[3] * Vx[4];
        Fx[7] = (Ux[7] + Vx[7] + 0.5 * b2) * Vx[1] - vdotb * Vx[4];
    } else if (dir == 1) {
        const double x = o->x[i];
        Fx[0] = Ux[2] * o->oneox[i];
        Fx[1] = Ux[1] * Vx[2] - Vx[4] * Vx[5];
        Fx[2] = Ux[2] * Vx[2] + x * ((Vx[7] + 0.5 * b2) - Vx[5] * Vx[5]);
        Fx[3] = Ux[3] * Vx[2] - Vx[6] * Vx[5];
        Fx[4] = Vx[4] * Vx[2] - Vx[1] * Vx[5];
        Fx[5] = 0.0;
        Fx[6] = Vx[6] * Vx[2] - Vx[3] * Vx[5];
        Fx[7] = (Ux[7] + Vx[7] + 0.5 * b2) * Vx[2] - vdotb * Vx[5];
    } else {
        const double x = o->x[i];
        Fx[0] = Ux[3];
        Fx[1] = Ux[1] * Vx[3] - Vx[4] * Vx[6];
        Fx[2] = Ux[2] * Vx[3] - x * Vx[5] * Vx[6];
        Fx[3] = Ux[3] * Vx[3] + (Vx[7] + 0.5 * b2) - Vx[6] * Vx[6];
        Fx[4] = Vx[4] * Vx[3] - Vx[1] * Vx[6];
        Fx[5] = Vx[5] * Vx[3] - Vx[2] * Vx[6];
        Fx[6] = 0.0;
        Fx[7] = (Ux[7] + Vx[7] + 0.5 * b2) * Vx[3] - vdotb * Vx[6];
    }
}

/* Primitiveto{F,G,H} (primitivetof.h:22-41, primitivetog.h, primitivetoh.h);
 * arguments are physical (r, phi, z) components */
static void prim_flux(const orc *o, int dir, int i, double vdotb, double rho, double vx, double vy,
                      double vz, double bx, double by, double bz, double en, double ptot, double *Fx)
{
    if (dir == 0) {
        Fx[0] = rho * vx;
        Fx[1] = rho * vx * vx + ptot - bx * bx;
        Fx[2] = o->xm[i] * (rho * vy * vx - by * bx);
        Fx[3] = rho * vz * vx - bz * bx;
        Fx[4] = 0.0;
        Fx[5] = by * vx - vy * bx;
        Fx[6] = bz * vx - vz * bx;
        Fx[7] = (en + ptot) * vx - vdotb * bx;
    } else if (dir == 1) {
        Fx[0] = rho * vy;
        Fx[1] = rho * vx * vy - bx * by;
        Fx[2] = o->x[i] * (rho * vy * vy + ptot - by * by);
        Fx[3] = rho * vz * vy - bz * by;
        Fx[4] = bx * vy - vx * by;
        Fx[5] = 0.0;
        Fx[6] = bz * vy - vz * by;
        Fx[7] = (en + ptot) * vy - vdotb * by;
    } else {
        Fx[0] = rho * vz;
        Fx[1] = rho * vx * vz - bx * bz;
        Fx[2] = o->x[i] * (rho * vy * vz - by * bz);
        Fx[3] = rho * vz * vz + ptot - bz * bz;
        Fx[4] = bx * vz - vx * bz;
        Fx[5] = by * vz - vy * bz;
        Fx[6] = 0.0;
        Fx[7] = (en + ptot) * vz - vdotb * bz;
    }
}

/* GLM part of the flux (hlldx.h:482-494, hlldy.h:490-503, hlldz.h:486-498) */
static void glm_flux(int dir, double ch, const double *vl, const double *vr, double *Fx)
{
    const int nb = 4 + dir;
    Fx[nb] += 0.5 * (vl[8] + vr[8] + ch * (vl[nb] - vr[nb]));
    Fx[8] = 0.5 * ch * (ch * (vl[nb] + vr[nb]) + (vl[8] - vr[8]));
}

/* HLLD{X,Y,Z} (hlldx.h:53-477, hlldy.h:55-485, hlldz.h:53-481) for one face.
 * n = normal component; (t1, t2) = the tangential pair in the order each
 * direction forms its dot products: x (phi, z), y (z, r), z (r, phi). */
void orc_hlld_face(const orc *o, int dir, int i, double ch, const double *Vl, const double *Vr, double *F)
{
    const double gm = o->p.gm, oneogm1 = o->p.oneogm1;
    const int n = 1 + dir, nb = 4 + dir;
    const int t1 = dir == 0 ? 2 : (dir == 1 ? 3 : 1);
    const int t2 = dir == 0 ? 3 : (dir == 1 ? 1 : 2);
    const double r = dir == 0 ? o->xm[i] : o->x[i];
    double Ul[8], Ur[8];
    double vl2, bl2, vldotbl, vr2, br2, vrdotbr, gmpr, ptl, ptr, vfl, vfr, sl, sr;

    vl2 = Vl[1] * Vl[1] + Vl[2] * Vl[2] + Vl[3] * Vl[3];
    bl2 = Vl[4] * Vl[4] + Vl[5] * Vl[5] + Vl[6] * Vl[6];
    vldotbl = Vl[1] * Vl[4] + Vl[2] * Vl[5] + Vl[3] * Vl[6];
    gmpr = gm * Vl[7];
    ptl = Vl[7] + 0.5 * bl2;
    vfl = sqrt(((bl2 + gmpr) + sqrt((bl2 + gmpr) * (bl2 + gmpr) - 4.0 * gmpr * Vl[nb] * Vl[nb])) / (2.0 * Vl[0]));
    vr2 = Vr[1] * Vr[1] + Vr[2] * Vr[2] + Vr[3] * Vr[3];
    br2 = Vr[4] * Vr[4] + Vr[5] * Vr[5] + Vr[6] * Vr[6];
    vrdotbr = Vr[1] * Vr[4] + Vr[2] * Vr[5] + Vr[3] * Vr[6];
    gmpr = gm * Vr[7];
    ptr = Vr[7] + 0.5 * br2;
    vfr = sqrt(((br2 + gmpr) + sqrt((br2 + gmpr) * (br2 + gmpr) - 4.0 * gmpr * Vr[nb] * Vr[nb])) / (2.0 * Vr[0]));

    sl = rmin(Vl[n], Vr[n]) - rmax(vfl, vfr);
    sr = rmax(Vl[n], Vr[n]) + rmax(vfl, vfr);

    if (sl >= 0.0) {
        vlr_to_ulr(vl2, bl2, r, Vl, Ul, oneogm1);
        phys_flux(o, dir, i, bl2, vldotbl, Vl, Ul, F);
    } else if (sr <= 0.0) {
        vlr_to_ulr(vr2, br2, r, Vr, Ur, oneogm1);
        phys_flux(o, dir, i, br2, vrdotbr, Vr, Ur, F);
    } else {
        double slvl, srvr, sm, oneoslsm, oneosrsm, ptint, bn, rholl, rhorr;
        double sgn, D, den, num;
        double vt1ll, vt2ll, bt1ll, bt2ll, vt1rr, vt2rr, bt1rr, bt2rr;
        double vlldotbll, vrrdotbrr, enll, enrr, sqrtrholl, sqrtrhorr, sll, srr;
        vlr_to_ulr(vl2, bl2, r, Vl, Ul, oneogm1);
        vlr_to_ulr(vr2, br2, r, Vr, Ur, oneogm1);
        slvl = sl - Vl[n];
        srvr = sr - Vr[n];
        if (dir == 1)
            sm = ((srvr * Vr[0] * Vr[2] - slvl * Vl[0] * Vl[2]) - ptr + ptl) / (srvr * Ur[0] - slvl * Ul[0]);
        else
            sm = (srvr * Ur[n] - slvl * Ul[n] - ptr + ptl) / (srvr * Ur[0] - slvl * Ul[0]);
        oneoslsm = 1.0 / (sl - sm);
        oneosrsm = 1.0 / (sr - sm);
        ptint = (srvr * Ur[0] * ptl - slvl * Ul[0] * ptr + Ur[0] * Ul[0] * srvr * slvl * (Vr[n] - Vl[n]))
                / (srvr * Ur[0] - slvl * Ul[0]);
        bn = (sr * Ur[nb] - sl * Ul[nb]) / (sr - sl);
        rholl = Ul[0] * slvl * oneoslsm;
        rhorr = Ur[0] * srvr * oneosrsm;

        /* left-left state with the branchless degeneracy switch */
        D = Ul[0] * slvl * (sl - sm) - bn * bn;
        sgn = (fabs(fabs(D) - 1.0e-8 * bn * bn) / (fabs(D) - 1.0e-8 * bn * bn));
        den = ((1.0 / D + 1.0) + (1.0 / D - 1.0) * sgn) * 0.5;
        num = ((((Ul[0] * slvl * slvl - bn * bn) + 1.0) + ((Ul[0] * slvl * slvl - bn * bn) - 1.0) * sgn) * 0.5);
        vt1ll = Vl[t1] - (bn * Ul[t1 + 3] * (sm - Vl[n]) * ((den + den * sgn) * 0.5));
        vt2ll = Vl[t2] - (bn * Ul[t2 + 3] * (sm - Vl[n]) * ((den + den * sgn) * 0.5));
        bt1ll = Ul[t1 + 3] * num * den;
        bt2ll = Ul[t2 + 3] * num * den;
        vlldotbll = sm * bn + vt1ll * bt1ll + vt2ll * bt2ll;

        /* right-right state */
        D = Ur[0] * srvr * (sr - sm) - bn * bn;
        sgn = (fabs(fabs(D) - 1.0e-8 * bn * bn) / (fabs(D) - 1.0e-8 * bn * bn));
        den = ((1.0 / D + 1.0) + (1.0 / D - 1.0) * sgn) * 0.5;
        num = ((((Ur[0] * srvr * srvr - bn * bn) + 1.0) + ((Ur[0] * srvr * srvr - bn * bn) - 1.0) * sgn) * 0.5);
        vt1rr = Vr[t1] - (bn * Ur[t1 + 3] * (sm - Vr[n]) * ((den + den * sgn) * 0.5));
        vt2rr = Vr[t2] - (bn * Ur[t2 + 3] * (sm - Vr[n]) * ((den + den * sgn) * 0.5));
        bt1rr = Ur[t1 + 3] * num * den;
        bt2rr = Ur[t2 + 3] * num * den;
        vrrdotbrr = sm * bn + vt1rr * bt1rr + vt2rr * bt2rr;

        enll = (slvl * Ul[7] - ptl * Vl[n] + ptint * sm + bn * (vldotbl - vlldotbll)) * oneoslsm;
        enrr = (srvr * Ur[7] - ptr * Vr[n] + ptint * sm + bn * (vrdotbr - vrrdotbrr)) * oneosrsm;
        sqrtrholl = sqrt(rholl);
        sqrtrhorr = sqrt(rhorr);
        sll = sm - fabs(bn) / sqrtrholl;
        srr = sm + fabs(bn) / sqrtrhorr;

        double v[4], b[4], vdotb, rho, en;
        if (sm >= 0.0 && sll >= 0.0) {
            v[n] = sm; v[t1] = vt1ll; v[t2] = vt2ll;
            b[n] = bn; b[t1] = bt1ll; b[t2] = bt2ll;
            vdotb = vlldotbll; rho = rholl; en = enll;
        } else if (sm < 0.0 && srr <= 0.0) {
            v[n] = sm; v[t1] = vt1rr; v[t2] = vt2rr;
            b[n] = bn; b[t1] = bt1rr; b[t2] = bt2rr;
            vdotb = vrrdotbrr; rho = rhorr; en = enrr;
        } else {
            double oneosqrtrho, vt1int, vt2int, bt1int, bt2int, vintdotbint;
            sgn = fabs(bn) / bn;
            oneosqrtrho = 1.0 / (sqrtrholl + sqrtrhorr);
            vt1int = (sqrtrholl * vt1ll + sqrtrhorr * vt1rr + (bt1rr - bt1ll) * sgn) * oneosqrtrho;
            vt2int = (sqrtrholl * vt2ll + sqrtrhorr * vt2rr + (bt2rr - bt2ll) * sgn) * oneosqrtrho;
            bt1int = (sqrtrholl * bt1rr + sqrtrhorr * bt1ll + sqrtrholl * sqrtrhorr * (vt1rr - vt1ll) * sgn) * oneosqrtrho;
            bt2int = (sqrtrholl * bt2rr + sqrtrhorr * bt2ll + sqrtrholl * sqrtrhorr * (vt2rr - vt2ll) * sgn) * oneosqrtrho;
            vintdotbint = sm * bn + vt1int * bt1int + vt2int * bt2int;
            v[n] = sm; v[t1] = vt1int; v[t2] = vt2int;
            b[n] = bn; b[t1] = bt1int; b[t2] = bt2int;
            vdotb = vintdotbint;
            if (sm >= 0.0) { rho = rholl; en = enll - sqrtrholl * (vlldotbll - vintdotbint) * sgn; }
            else           { rho = rhorr; en = enrr + sqrtrhorr * (vrrdotbrr - vintdotbint) * sgn; }
        }
        prim_flux(o, dir, i, vdotb, rho, v[1], v[2], v[3], b[1], b[2], b[3], en, ptint, F);
    }
    glm_flux(dir, ch, Vl, Vr, F);
}

/* HLL{X,Y,Z} (hllx.h:43-179, hlly.h, hllz.h) for one face */
void orc_hll_face(const orc *o, int dir, int i, double ch, const double *Vl, const double *Vr, double *F)
{
    const double gm = o->p.gm, oneogm1 = o->p.oneogm1;
    const int n = 1 + dir, nb = 4 + dir;
    const double r = dir == 0 ? o->xm[i] : o->x[i];
    double Ul[8], Ur[8], Fl[8], Fr[8];
    double v2, b2, vdotb, gmpr, vfl, vfr, sl, sr;
    v2 = Vl[1] * Vl[1] + Vl[2] * Vl[2] + Vl[3] * Vl[3];
    b2 = Vl[4] * Vl[4] + Vl[5] * Vl[5] + Vl[6] * Vl[6];
    vdotb = Vl[1] * Vl[4] + Vl[2] * Vl[5] + Vl[3] * Vl[6];
    vlr_to_ulr(v2, b2, r, Vl, Ul, oneogm1);
    phys_flux(o, dir, i, b2, vdotb, Vl, Ul, Fl);
    gmpr = gm * Vl[7];
    vfl = sqrt(((b2 + gmpr) + sqrt((b2 + gmpr) * (b2 + gmpr) - 4.0 * gmpr * Vl[nb] * Vl[nb])) / (2.0 * Vl[0]));
    v2 = Vr[1] * Vr[1] + Vr[2] * Vr[2] + Vr[3] * Vr[3];
    b2 = Vr[4] * Vr[4] + Vr[5] * Vr[5] + Vr[6] * Vr[6];
    vdotb = Vr[1] * Vr[4] + Vr[2] * Vr[5] + Vr[3] * Vr[6];
    vlr_to_ulr(v2, b2, r, Vr, Ur, oneogm1);
    phys_flux(o, dir, i, b2, vdotb, Vr, Ur, Fr);
    gmpr = gm * Vr[7];
    vfr = sqrt(((b2 + gmpr) + sqrt((b2 + gmpr) * (b2 + gmpr) - 4.0 * gmpr * Vr[nb] * Vr[nb])) / (2.0 * Vr[0]));
    sl = rmin(Vl[n], Vr[n]) - rmax(vfl, vfr);
    sr = rmax(Vl[n], Vr[n]) + rmax(vfl, vfr);
    if (sl >= 0.0) {
        for (int m = 0; m < 8; m++) F[m] = Fl[m];
    } else if (sr >= 0.0) {
        for (int m = 0; m < 8; m++) F[m] = (sr * Fl[m] - sl * Fr[m] + sr * sl * (Ur[m] - Ul[m])) / (sr - sl);
    } else {
        for (int m = 0; m < 8; m++) F[m] = Fr[m];
    }
    glm_flux(dir, ch, Vl, Vr, F);
}

/* Riemann pass over every face of the reference's loop ranges */
static void riemann(orc *o, int dir)
{
    const size_t NF = o->NF;
    double *Fo = dir == 0 ? o->F : (dir == 1 ? o->G : o->H);
    const int hll = o->p.solver == CYLMHD_SOLVER_HLL;
    int i;
#pragma omp parallel for
    for (i = 0; i < o->NX - 1; i++)
        for (int j = 0; j < o->NY - 1; j++)
            for (int k = 0; k < o->NZ - 1; k++) {
                const size_t f = F3(o, i, j, k);
                double vl[9], vr[9], fx[9];
                for (int m = 0; m < 9; m++) { vl[m] = o->Vl[m * NF + f]; vr[m] = o->Vr[m * NF + f]; }
                if (hll) orc_hll_face(o, dir, i, o->ch, vl, vr, fx);
                else orc_hlld_face(o, dir, i, o->ch, vl, vr, fx);
                for (int m = 0; m < 9; m++) Fo[m * NF + f] = fx[m];
            }
}

/* Source terms, by p->source: SpiralGravityCooling (source.h:26-134, the
 * default), AxiGravityCooling (source.h:143-200: Grxaxi/Grzaxi at every i),
 * AxiGravity (source.h:207-261: the same without the cooling term),
 * UniformGz (source.h:269-300) and NoSource (source.h:308-349, which leaves
 * S8 untouched: 0 from the start) */
static void sources(orc *o)
{
    const cylmhd_params *p = &o->p;
    const size_t N = o->N;
    const double *V = o->V;
    double *S = o->S;
    const double omegasp = p->omegasp;
    const int mode = p->source;
    int i;
    if (mode == CYLMHD_SOURCE_UNIFORM_GZ || mode == CYLMHD_SOURCE_NONE) {
#pragma omp parallel for
        for (i = 0; i < o->NX; i++)
            for (int j = 0; j < o->NY; j++)
                for (int k = 0; k < o->NZ; k++) {
                    const size_t c = C3(o, i, j, k);
                    const double V0 = V[c], V1 = V[N + c], V2 = V[2 * N + c], V3 = V[3 * N + c];
                    const double V4 = V[4 * N + c], V5 = V[5 * N + c], V6 = V[6 * N + c], V7 = V[7 * N + c];
                    for (int m = 0; m < 8; m++) S[m * N + c] = 0.0;
                    if (mode == CYLMHD_SOURCE_UNIFORM_GZ) {
                        S[3 * N + c] = -V0 * 1.0;
                        S[7 * N + c] = -V3 * V0 * 1.0;
                        S[8 * N + c] = 0.0;
                    } else {
                        S[1 * N + c] = (V0 * V2 * V2 - V5 * V5 + (V7 + 0.5 * (V4 * V4 + V5 * V5 + V6 * V6))) * o->oneox[i];
                        S[5 * N + c] = (V5 * V1 - V2 * V4) * o->oneox[i];
                    }
                }
        return;
    }
#pragma omp parallel for
    for (i = 0; i < o->NX; i++) {
        const int spiral = i >= o->irin && mode == CYLMHD_SOURCE_SPIRAL_COOLING;
        for (int j = 0; j < o->NY; j++)
            for (int k = 0; k < o->NZ; k++) {
                const size_t c = C3(o, i, j, k);
                const double V0 = V[c], V1 = V[N + c], V2 = V[2 * N + c], V3 = V[3 * N + c];
                const double V4 = V[4 * N + c], V5 = V[5 * N + c], V6 = V[6 * N + c], V7 = V[7 * N + c];
                const double gx = spiral ? o->Grx[c] : o->Grxaxi[c];
                const double gz = spiral ? o->Grz[c] : o->Grzaxi[c];
                const double x = o->x[i];
                S[0 * N + c] = 0.0;
                S[1 * N + c] = ((V0 * V2 * V2 - V5 * V5 + (V7 + 0.5 * (V4 * V4 + V5 * V5 + V6 * V6))) * o->oneox[i]
                                + V0 * gx + V0 * omegasp * omegasp * x + (2.0 * V0 * V2 * omegasp));
                if (spiral) S[2 * N + c] = x * (V0 * o->Gry[c] - (2.0 * V0 * V1 * omegasp));
                else        S[2 * N + c] = x * (-2.0 * V0 * V1 * omegasp);
                S[3 * N + c] = V0 * gz;
                S[4 * N + c] = 0.0;
                S[5 * N + c] = (V5 * V1 - V2 * V4) * o->oneox[i];
                S[6 * N + c] = 0.0;
                if (spiral)
                    S[7 * N + c] = (V0 * (V1 * gx + V2 * o->Gry[c] + V3 * gz)
                                    - (V0 * V0 * orc_cooling(p, V0, V7))
                                    + V0 * x * V1 * omegasp * omegasp);
                else if (mode == CYLMHD_SOURCE_AXI)
                    S[7 * N + c] = (V0 * (V1 * gx + V3 * gz)
                                    + V0 * x * V1 * omegasp * omegasp);
                else
                    S[7 * N + c] = (V0 * (V1 * gx + V3 * gz)
                                    - (V0 * V0 * orc_cooling(p, V0, V7))
                                    + V0 * x * V1 * omegasp * omegasp);
                S[8 * N + c] = 0.0;
            }
    }
}

/* stage update loops: firststep.h:97-118, secondstep.h:100-121,
 * thirdstep.h:100-123, halfstep.h:91-110, fullstep.h:93-112 */
static void update(orc *o, int kind)
{
    const size_t N = o->N, NF = o->NF;
    const double dt = o->dt, oneo3 = 1.0 / 3.0;
    int i;
#pragma omp parallel for
    for (i = 1; i < o->NX - 1; i++)
        for (int j = 1; j < o->NY - 1; j++)
            for (int k = 1; k < o->NZ - 1; k++) {
                const size_t c = C3(o, i, j, k);
                const size_t f = F3(o, i, j, k);
                const size_t fi = F3(o, i - 1, j, k), fj = F3(o, i, j - 1, k), fk = F3(o, i, j, k - 1);
                for (int m = 0; m < 9; m++) {
                    const double *Fm = o->F + m * NF, *Gm = o->G + m * NF, *Hm = o->H + m * NF;
                    const double div = ((o->g2[c] * Fm[f] - o->g1[c] * Fm[fi])
                                        + (o->g4[c] * Gm[f] - o->g3[c] * Gm[fj])
                                        + (o->g6[c] * Hm[f] - o->g5[c] * Hm[fk]));
                    double *U = o->U + m * N, *U1 = o->U1 + m * N;
                    const double S = o->S[m * N + c];
                    switch (kind) {
                    case 1: U1[c] = U[c] - dt * div + dt * S; break;                        /* FirstStep  */
                    case 2: U1[c] = 0.25 * (3.0 * U[c] + U1[c] - dt * div + dt * S); break; /* SecondStep */
                    case 3: U[c] = oneo3 * (U[c] + 2.0 * (U1[c] - dt * div + dt * S)); break; /* ThirdStep */
                    case 4: U1[c] = (U[c] - dt * div + dt * S); break;                      /* Halfstep   */
                    case 5: U[c] = 0.5 * (U[c] + U1[c] - dt * div + dt * S); break;         /* Fullstep   */
                    }
                }
            }
}

/* cell boundaries: yboundary.h:97-154, xboundary.h:33-115, 322-346,
 * zboundary.h:67-96 (VzAntisymBSym), 105-161 (periodic), 203-226 (free) */
static void cell_bcs(orc *o, double *dU, int rk3)
{
    const size_t N = o->N;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ, jy = o->p.jy;
    int i, j, k, m;
    /* YLeftPeriodicBoundary, YRightPeriodicBoundary */
    for (m = 0; m < 9; m++)
        for (i = 0; i < NX; i++)
            for (k = 0; k < NZ; k++) dU[m * N + C3(o, i, 0, k)] = dU[m * N + C3(o, i, NY - 2, k)];
    for (m = 0; m < 9; m++)
        for (i = 0; i < NX; i++)
            for (k = 0; k < NZ; k++) dU[m * N + C3(o, i, NY - 1, k)] = dU[m * N + C3(o, i, 1, k)];
    /* XLeftAxisBoundary: across the axis, j -> j +- jy/2 */
    for (j = 0; j < NY; j++) {
        const int js = j < jy / 2 + 1 ? j + jy / 2 : j - jy / 2;
        for (k = 0; k < NZ; k++) {
            dU[0 * N + C3(o, 0, j, k)] = dU[0 * N + C3(o, 1, js, k)];
            dU[1 * N + C3(o, 0, j, k)] = -dU[1 * N + C3(o, 1, js, k)];
            dU[2 * N + C3(o, 0, j, k)] = -dU[2 * N + C3(o, 1, js, k)];
            dU[3 * N + C3(o, 0, j, k)] = dU[3 * N + C3(o, 1, js, k)];
            dU[4 * N + C3(o, 0, j, k)] = -dU[4 * N + C3(o, 1, js, k)];
            dU[5 * N + C3(o, 0, j, k)] = -dU[5 * N + C3(o, 1, js, k)];
            dU[6 * N + C3(o, 0, j, k)] = dU[6 * N + C3(o, 1, js, k)];
            dU[7 * N + C3(o, 0, j, k)] = dU[7 * N + C3(o, 1, js, k)];
            dU[8 * N + C3(o, 0, j, k)] = dU[8 * N + C3(o, 1, js, k)];
        }
    }
    /* XRightFreeBoundary */
    for (m = 0; m < 9; m++)
        for (j = 0; j < NY; j++)
            for (k = 0; k < NZ; k++) dU[m * N + C3(o, NX - 1, j, k)] = dU[m * N + C3(o, NX - 2, j, k)];
    if (rk3) {
        /* ZLeftPeriodicBoundary, ZRightPeriodicBoundary */
        for (m = 0; m < 9; m++)
            for (i = 0; i < NX; i++)
                for (j = 0; j < NY; j++) dU[m * N + C3(o, i, j, 0)] = dU[m * N + C3(o, i, j, NZ - 2)];
        for (m = 0; m < 9; m++)
            for (i = 0; i < NX; i++)
                for (j = 0; j < NY; j++) dU[m * N + C3(o, i, j, NZ - 1)] = dU[m * N + C3(o, i, j, 1)];
    } else {
        /* ZLeftVzAntisymBSymBoundary, ZRightFreeBoundary */
        for (i = 0; i < NX; i++)
            for (j = 0; j < NY; j++) {
                for (m = 0; m < 9; m++)
                    dU[m * N + C3(o, i, j, 0)] = m == 3 ? -dU[m * N + C3(o, i, j, 2)] : dU[m * N + C3(o, i, j, 2)];
                dU[3 * N + C3(o, i, j, 1)] = 0.0;
            }
        for (m = 0; m < 9; m++)
            for (i = 0; i < NX; i++)
                for (j = 0; j < NY; j++) dU[m * N + C3(o, i, j, NZ - 1)] = dU[m * N + C3(o, i, j, NZ - 2)];
    }
}

/* XRightAbsorbBoundary (xboundary.h:124-149), CentralAbsorbBoundary
 * (centralboundary.h:27-52), ZRightAbsorbBoundary (zboundary.h:33-58) */
static void absorb(orc *o, double *dU, int zright)
{
    const size_t N = o->N;
    const int NX = o->NX, NY = o->NY, NZ = o->NZ;
    for (int m = 0; m < 9; m++)
        for (int i = NX - 1 - 10; i < NX; i++)
            for (int j = 0; j < NY; j++)
                for (int k = 0; k < NZ; k++) {
                    const size_t c = m * N + C3(o, i, j, k);
                    dU[c] -= (o->xabs[i - (NX - 1 - 10)] * (dU[c] - o->Uinit[c]));
                }
    for (int m = 0; m < 9; m++)
        for (int i = 0; i < o->irin + 1; i++)
            for (int j = 0; j < NY; j++)
                for (int k = 0; k < NZ; k++) {
                    const size_t c = m * N + C3(o, i, j, k);
                    dU[c] -= (o->cabs[(size_t)i * NZ + k] * (dU[c] - o->Uinit[c]));
                }
    if (zright)
        for (int m = 0; m < 9; m++)
            for (int i = 0; i < NX; i++)
                for (int j = 0; j < NY; j++)
                    for (int k = NZ - 1 - 10; k < NZ; k++) {
                        const size_t c = m * N + C3(o, i, j, k);
                        if (k < 0) continue;
                        dU[c] -= (o->zabs[k - (NZ - 1 - 10)] * (dU[c] - o->Uinit[c]));
                    }
}

/* UtoV (utov.h:24-68) */
static void utov(orc *o, const double *dU, double *dV)
{
    const size_t N = o->N;
    const double gm1 = o->p.gm1;
    int i;
#pragma omp parallel for
    for (i = 0; i < o->NX; i++)
        for (size_t jk = 0; jk < (size_t)o->NY * o->NZ; jk++) {
            const size_t c = (size_t)i * o->NY * o->NZ + jk;
            double oneorho = 1.0 / dU[c];
            double b2 = dU[4 * N + c] * dU[4 * N + c] + dU[5 * N + c] * dU[5 * N + c] + dU[6 * N + c] * dU[6 * N + c];
            double r2v2 = dU[1 * N + c] * dU[1 * N + c] + dU[2 * N + c] * dU[2 * N + c] * o->oneox[i] * o->oneox[i]
                          + dU[3 * N + c] * dU[3 * N + c];
            dV[c] = dU[c];
            dV[1 * N + c] = dU[1 * N + c] * oneorho;
            dV[2 * N + c] = dU[2 * N + c] * oneorho * o->oneox[i];
            dV[3 * N + c] = dU[3 * N + c] * oneorho;
            dV[4 * N + c] = dU[4 * N + c];
            dV[5 * N + c] = dU[5 * N + c];
            dV[6 * N + c] = dU[6 * N + c];
            dV[7 * N + c] = gm1 * (dU[7 * N + c] - 0.5 * (r2v2 * oneorho + b2));
            dV[8 * N + c] = dU[8 * N + c];
        }
}

/* Check (check.h:23-188) */
static void check(orc *o, double *dU, double *dV, int *counts)
{
    const cylmhd_params *p = &o->p;
    const size_t N = o->N;
    const double oneobetacr = (1.0 / p->betacr), oneogm1 = p->oneogm1;
    int rhoflag = 0, prflag = 0, oneobetaflag = 0;
    int i;
#pragma omp parallel for reduction(+ : rhoflag, prflag, oneobetaflag)
    for (i = 0; i < o->NX; i++)
        for (size_t jk = 0; jk < (size_t)o->NY * o->NZ; jk++) {
            const size_t c = (size_t)i * o->NY * o->NZ + jk;
            const double x = o->x[i];
            double v2 = dV[2 * N + c] * dV[2 * N + c];
            double b2 = (dV[4 * N + c] * dV[4 * N + c] + dV[5 * N + c] * dV[5 * N + c] + dV[6 * N + c] * dV[6 * N + c]);
            double prfloor, oneobeta;
            if (dV[c] < o->rhofloor[c]) {
                rhoflag++;
                dV[c] = o->rhofloor[c];
                dV[1 * N + c] = 0.0;
                dV[3 * N + c] = 0.0;
                dU[c] = dV[c];
                dU[1 * N + c] = 0.0;
                dU[2 * N + c] = x * dV[c] * dV[2 * N + c];
                dU[3 * N + c] = 0.0;
                dU[7 * N + c] = 0.5 * (dV[c] * v2 + b2) + oneogm1 * dV[7 * N + c];
            }
            prfloor = dV[c] * p->prcr;
            if (dV[7 * N + c] < prfloor) {
                prflag++;
                dV[1 * N + c] = 0.0;
                dV[3 * N + c] = 0.0;
                dV[7 * N + c] = prfloor;
                dU[1 * N + c] = 0.0;
                dU[3 * N + c] = 0.0;
                dU[7 * N + c] = (0.5 * (dV[c] * v2 + b2) + oneogm1 * dV[7 * N + c]);
            }
            oneobeta = 0.5 * b2 / dV[7 * N + c];
            if (oneobeta > oneobetacr) {
                oneobetaflag++;
                dV[c] = p->betacr * oneobeta * dV[c];
                dV[1 * N + c] = 0.0;
                dV[3 * N + c] = 0.0;
                dV[7 * N + c] = p->betacr * oneobeta * dV[7 * N + c];
                dU[c] = dV[c];
                dU[1 * N + c] = 0.0;
                dU[2 * N + c] = x * dV[c] * dV[2 * N + c];
                dU[3 * N + c] = 0.0;
                dU[7 * N + c] = (0.5 * (dV[c] * v2 + b2) + oneogm1 * dV[7 * N + c]);
            }
        }
    counts[0] = rhoflag; counts[1] = prflag; counts[2] = oneobetaflag;
}

/* RestartRead's tail (restartread.h:132-139): VtoU (vtou.h:20-62) of the
 * primitives just read, then Check(U, V) with the current rhofloor -- still
 * zero before the first orc_set_initial, as the reference's zero-initialised
 * global is at RestartRead(0) (main.c:250).  The result is the current state. */
void orc_restart_state(orc *o, const double *V, int counts[3])
{
    const size_t N = o->N;
    const double oneogm1 = o->p.oneogm1;
    memcpy(o->V, V, sizeof(double) * 9 * N);
    double *U = o->U, *W = o->V;
    for (int i = 0; i < o->NX; i++)
        for (size_t jk = 0; jk < (size_t)o->NY * o->NZ; jk++) {
            const size_t c = (size_t)i * o->NY * o->NZ + jk;
            const double v2 = W[1 * N + c] * W[1 * N + c] + W[2 * N + c] * W[2 * N + c] + W[3 * N + c] * W[3 * N + c];
            const double b2 = W[4 * N + c] * W[4 * N + c] + W[5 * N + c] * W[5 * N + c] + W[6 * N + c] * W[6 * N + c];
            U[c] = W[c];
            U[1 * N + c] = W[c] * W[1 * N + c];
            U[2 * N + c] = o->x[i] * W[c] * W[2 * N + c];
            U[3 * N + c] = W[c] * W[3 * N + c];
            U[4 * N + c] = W[4 * N + c];
            U[5 * N + c] = W[5 * N + c];
            U[6 * N + c] = W[6 * N + c];
            U[7 * N + c] = 0.5 * (W[c] * v2 + b2) + oneogm1 * W[7 * N + c];
            U[8 * N + c] = W[8 * N + c];
        }
    int cnt[3];
    check(o, U, W, cnt);
    if (counts) { counts[0] = cnt[0]; counts[1] = cnt[1]; counts[2] = cnt[2]; }
}

/* one RK stage: FirstStep (firststep.h:19-176), SecondStep (secondstep.h:
 * 21-190), ThirdStep (thirdstep.h:19-221), Halfstep (halfstep.h:22-168),
 * Fullstep (fullstep.h:21-195) */
void orc_stage(orc *o, int stage, double dt, double ch)
{
    const int rk3 = o->p.rk != 2;
    o->dt = dt;
    o->ch = ch;
    o->psicoef = exp(-dt * ch * (1.0 / o->p.cr) * 0.5);   /* tvdrungekutta.h:22 */
    double *Uin = stage == 1 ? o->U : o->U1;
    glm_source(o, Uin, o->V);
    limiter(o, 0); riemann(o, 0);
    limiter(o, 1); riemann(o, 1);
    limiter(o, 2); riemann(o, 2);
    sources(o);
    int kind = rk3 ? stage : 3 + stage;
    update(o, kind);
    const int last = rk3 ? stage == 3 : stage == 2;
    double *Uout = last ? o->U : o->U1;
    cell_bcs(o, Uout, rk3);
    if (last) absorb(o, Uout, !rk3);
    utov(o, Uout, o->V);
    check(o, Uout, o->V, &o->counts[3 * (stage - 1)]);
}

/* TVDRK (tvdrungekutta.h:16-78) */
void orc_step(orc *o, double dt, double ch)
{
    const int nst = o->p.rk == 2 ? 2 : 3;
    memset(o->counts, 0, sizeof o->counts);
    for (int s = 1; s <= nst; s++) orc_stage(o, s, dt, ch);
}

/* conserved totals and div B over interior cells */
void orc_diagnostics(const orc *o, cylmhd_diag *d)
{
    const size_t N = o->N;
    const double *U = o->U;
    double mass = 0.0, energy = 0.0, divb = 0.0, psi = 0.0;
    int i;
#pragma omp parallel for reduction(+ : mass, energy, divb, psi)
    for (i = 1; i < o->NX - 1; i++)
        for (int j = 1; j < o->NY - 1; j++)
            for (int k = 1; k < o->NZ - 1; k++) {
                const size_t c = C3(o, i, j, k);
                const size_t si = (size_t)o->NY * o->NZ, sj = o->NZ;
                const double dv = o->dvolume[c];
                const double brp = 0.5 * (U[4 * N + c] + U[4 * N + c + si]), brm = 0.5 * (U[4 * N + c] + U[4 * N + c - si]);
                const double bpp = 0.5 * (U[5 * N + c] + U[5 * N + c + sj]), bpm = 0.5 * (U[5 * N + c] + U[5 * N + c - sj]);
                const double bzp = 0.5 * (U[6 * N + c] + U[6 * N + c + 1]), bzm = 0.5 * (U[6 * N + c] + U[6 * N + c - 1]);
                const double db = (o->g2[c] * brp - o->g1[c] * brm) + (o->g4[c] * bpp - o->g3[c] * bpm)
                                  + (o->g6[c] * bzp - o->g5[c] * bzm);
                mass += U[c] * dv;
                energy += U[7 * N + c] * dv;
                divb += fabs(db) * dv;
                psi += fabs(U[8 * N + c]) * dv;
            }
    d->mass = mass; d->energy = energy; d->divb_l1 = divb; d->psi_l1 = psi;
}

/* main.c:316-393 without dumps */
double orc_run(orc *o, int nsteps, double t, double *elapsed)
{
    double t0 = omp_get_wtime();
    for (int n = 0; n < nsteps; n++) {
        double dt, ch;
        int ijk[3];
        orc_setdt(o, &dt, &ch, ijk);
        if (isinf(dt) || dt < o->p.dtmin) break;
        orc_gravity(o, t);
        orc_step(o, dt, ch);
        t += dt;
    }
    if (elapsed) *elapsed = omp_get_wtime() - t0;
    return t;
}
