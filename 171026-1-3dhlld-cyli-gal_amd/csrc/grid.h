/*
 * grid.h -- host-side cylindrical grid and geometry tables.
 *
 * Restates CylindricalGrid (cylindricalgrid.h:25-227) with run-time sizes.
 * The reference stores g1..g6, dvolume and minlength as 3-D arrays, but every
 * one of them depends on (i,k) only; they are kept here as [NX][NZ] tables
 * holding bit-identical values.
 */
#ifndef CYL_GRID_H
#define CYL_GRID_H

#include "../../include/cylmhd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cyl_grid {
    int NX, NY, NZ;          /* ix+2, jy+2, kz+2                          */
    int irin;                /* (int)(rin/dx0)          input.h:72        */
    double *x, *xm, *dx, *dxm, *oneox, *oneoxm, *oneodx;   /* [NX]        */
    double *y;                                             /* [NY]        */
    double dy, oneody;
    double *z, *zm, *dz, *dzm, *oneodz;                    /* [NZ]        */
    double *g1, *g2, *g3, *g5;   /* [NX*NZ]; g4 == g3, g6 == g5 (:120-125) */
    double *dvol;                /* [NX*NZ] dvolume (:114)                 */
    double *minlength;           /* [NX*NZ] (:146)                         */
    double minminlength;         /* 1/max(1/minlength) (:136-163)          */
    double xabs[11], zabs[11];   /* x/z_absorb_factor (:169-187)           */
    double *cabs;                /* [(irin+1)*NZ] c_absorb_factor (:190-201) */
} cyl_grid;

int  cyl_grid_build(const cylmhd_params *p, cyl_grid *g);
void cyl_grid_free(cyl_grid *g);

/* Axisymmetric gravity tables [NX*NZ] (Gryaxi is identically 0).
 * CYLMHD_GRAV_GALACTIC restates InitialGravity (initialgravity.h:43-68). */
int cyl_gravity_axi(const cylmhd_params *p, const cyl_grid *g,
                    double *grx_ik, double *grz_ik);

/* Time-independent factors of SpiralGravity (spiralgravity.h:37-77):
 *   per i : x2oscaler2, oneoonex2oscaler2, xoscaler2oonex2oscaler215,
 *           x2oscaler2oonex2oscaler215, log(x/scalephase)
 *   per k : a1zk2b12, scalezosqrtz2scalez2, 1+a1/sqrt(z2+b1^2),
 *           1/(z2+scalez^2)
 *   per ik: oneosqrtx2a1zk2b12                                         */
typedef struct cyl_spiral_tables {
    double *x2s2, *o1x2s2, *xs2o15, *x2s2o15, *logxs;    /* [NX]   */
    double *szs, *onepa, *oz2s2;                          /* [NZ]   */
    double *q;                                            /* [NX*NZ] */
    double cotpitch;
} cyl_spiral_tables;

int  cyl_spiral_tables_build(const cylmhd_params *p, const cyl_grid *g,
                             cyl_spiral_tables *s);
void cyl_spiral_tables_free(cyl_spiral_tables *s);
/* epssp ramp (spiralgravity.h:41-42) */
double cyl_spiral_eps(const cylmhd_params *p, double t);

#ifdef __cplusplus
}
#endif
#endif
