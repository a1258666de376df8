/*
 * batch.h -- what the batch layer needs from a disc context (internal).
 */
#ifndef CYLMHD_BATCH_H
#define CYLMHD_BATCH_H

#include "fused.h"

namespace cyl {

struct DiscView {
    const cylmhd_params *p;
    DevGeom dg;                 /* device tables of the disc */
    DevState ds;                /* Uinit, Vinit_z, counts, N */
    FusedState *fs;             /* FP32 state buffers */
    double *sincos;             /* [2][NX][NY] spiral phase of this step */
    const double *logxs;        /* ln(x/scalephase) per i */
    double minminlength;
    hipStream_t stream;
};

/* fills *v for a context created with CYLMHD_PATH_FUSED32 */
int disc_view(cylmhd_ctx *c, DiscView *v);
/* cylmhd_create; batch_layout: the state buffers keep their z-side ghost
 * cells packed after the body (the batch kernels' layout) */
int create_ctx(const cylmhd_params *p, cylmhd_ctx **out, bool batch_layout);
/* the stage constants of main.c's step for this disc (dt, ch, psicoef) */
Consts disc_consts(const cylmhd_ctx *c, double dt, double ch);
/* the context's own Setdt partials no longer describe its state */
void disc_invalidate_dt(cylmhd_ctx *c);
/* last error text of the library (error_string adds it) */
void set_errmsg(const char *msg);

}  // namespace cyl
#endif
