/*
 * paths.h -- host-side launchers of the device paths (internal).
 */
#ifndef CYLMHD_PATHS_H
#define CYLMHD_PATHS_H

#include "ctx.h"

namespace cyl {

/* state and scratch of one context, all device pointers */
struct DevState {
    size_t N, NF;
    /* plain FP64 path (reference data flow: U, U1, V, F, G, H) */
    double *U, *U1, *V, *F, *G, *H;
    /* initial state: Uinit for the absorbing layers and rho floor,
     * Vinit planes k = 0 and k = NZ-1 for the MC z boundary and
     * j = 0 and j = NY-1 for the vanLeer y boundary */
    double *Uinit, *Vinit_z, *Vinit_y;  /* [9][N], [9][2][NX][NY], [9][2][NX][NZ] */
    int *counts;                   /* [9] floor counters of the last step */
    double *red;                   /* reduction scratch */
    int nred;
    int *absorb_i;                 /* i-planes of the absorbing layers */
    int n_absorb_i;
    int *nonabsorb_i;              /* the other i-planes (RK2 z layer) */
    int n_nonabsorb_i;
};

struct StageArgs {
    int stage;          /* 1..3 (RK3) or 1..2 (RK2) */
    bool rk3, mc, hll;
    int limiter;        /* CYLMHD_LIMITER_* */
};

/* plain FP64 path (plain.hip) */
void plain_stage(hipStream_t s, const DevGeom &g, const DevState &d, const Consts &c, const StageArgs &a,
                 hipEvent_t ev0, hipEvent_t ev1);
/* shared helpers (plain.hip) */
void launch_sincos(hipStream_t s, const DevGeom &g, double *sincos, const double *logxs, double t,
                   double omegapattern, double omegasp);
void launch_setdt(hipStream_t s, const DevGeom &g, const Consts &c, const double *V, double *red, int nred);
void launch_diag(hipStream_t s, const DevGeom &g, const double *U, double *red, int nred);
/* one slice of quantity q (CYLMHD_Q_* or 0..8) over plane (CYLMHD_PLANE_*)
 * of the FP64 primitive state V into out[n] (slices.hip) */
void launch_slice(hipStream_t s, const DevGeom &g, const Consts &c, const double *V, const double *dzm, int q,
                  int plane, int a, int b, double *out, int n);

}  // namespace cyl
#endif
