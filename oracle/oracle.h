/*
 * oracle.h -- CPU FP64 restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker / the timed CPU baseline.  The product path never links it.
 *
 * It restates, in the reference's pass structure (limiter pass -> Vl/Vr
 * arrays -> Riemann pass -> F/G/H arrays -> sources -> update -> boundaries
 * -> UtoV -> Check) and with the reference's expression order, the routines
 * of github-nakasho/171026-1-3dhlld-cyli-gal listed at each function in
 * oracle.c.  Sizes are run-time instead of input.h's compile-time #defines;
 * the parameter set is cylmhd_params from include/cylmhd.h.
 */
#ifndef CYLMHD_ORACLE_H
#define CYLMHD_ORACLE_H

#include "../include/cylmhd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc   *orc_create(const cylmhd_params *p);
void   orc_destroy(orc *o);
void   orc_dims(const orc *o, int *nx, int *ny, int *nz);
/* grid vectors (cylindricalgrid.h:25-227) */
void   orc_grid(const orc *o, double *x, double *xm, double *y, double *z, double *minminlength);
/* g-factor / volume tables: out[6][NX][NZ] = g1,g2,g3,g5,dvolume,minlength */
void   orc_geometry(const orc *o, double *out);
void   orc_dzm(const orc *o, double *dzm);
/* absorbing profiles: xabs[11], zabs[11], cabs[(irin+1)*NZ]; returns irin */
int    orc_absorb_factors(const orc *o, double *xabs, double *zabs, double *cabs);
/* axisymmetric gravity [NX*NZ] tables actually used */
void   orc_gravity_axi(const orc *o, double *grx, double *grz);
/* SaveInitial (saveinitial.h:18-50) on the given state: sets U, V, Uinitial,
 * Vinitial and rhofloor. */
void   orc_set_initial(orc *o, const double *U, const double *V);
void   orc_set_state(orc *o, const double *U, const double *V);
/* RestartRead's tail (restartread.h:132-139): U = VtoU(V), then Check(U, V)
 * with the current rhofloor (zero before orc_set_initial); becomes the
 * current state.  counts (optional) = Check's density/pressure/beta hits. */
void   orc_restart_state(orc *o, const double *V, int counts[3]);
void   orc_get_state(const orc *o, double *U, double *V);
/* Setdt (setdt.h:24-94) */
void   orc_setdt(orc *o, double *dt, double *ch, int ijk[3]);
/* SpiralGravity (spiralgravity.h:20-124); out (optional) = Grx,Gry,Grz */
void   orc_gravity(orc *o, double t);
void   orc_get_gravity(const orc *o, double *grx, double *gry, double *grz);
/* TVDRK (tvdrungekutta.h:16-78): RK3 or RK2 by p->rk */
void   orc_step(orc *o, double dt, double ch);
/* one stage only (1..3 RK3, 1..2 RK2); psicoef from dt, ch */
void   orc_stage(orc *o, int stage, double dt, double ch);
/* face fluxes of the last Riemann passes: out[3][9][NX-1][NY-1][NZ-1] */
void   orc_get_fluxes(const orc *o, double *out);
void   orc_floor_counts(const orc *o, int counts[9]);
void   orc_diagnostics(const orc *o, cylmhd_diag *d);
/* main.c:316-393 loop without output: nsteps of Setdt, SpiralGravity, TVDRK.
 * Returns t; *elapsed = wall seconds of the loop (omp_get_wtime). */
double orc_run(orc *o, int nsteps, double t0, double *elapsed);
int    orc_threads(void);

/* pointwise building blocks, exposed for unit tests */
double orc_cooling(const cylmhd_params *p, double rho, double pr);
void   orc_hlld_face(const orc *o, int dir, int i, double ch, const double *vl,
                     const double *vr, double *flux);
void   orc_hll_face(const orc *o, int dir, int i, double ch, const double *vl,
                    const double *vr, double *flux);

#ifdef __cplusplus
}
#endif
#endif
