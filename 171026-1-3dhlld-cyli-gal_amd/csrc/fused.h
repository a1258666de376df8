/*
 * fused.h -- the fused device paths (FP64 and FP32 state), internal.
 */
#ifndef CYLMHD_FUSED_H
#define CYLMHD_FUSED_H

#include <functional>

#include "paths.h"

namespace cyl {

struct FusedState {
    bool fp32 = false;
    size_t esize = 8;          /* bytes per state value */
    /* state buffers, [9][N] of float or double:
     *   Un    U^n (in/out of a step)      A, B  stage outputs U(1), U(2)
     *   shadow  pre-Check stage values of the cells the ghost map reads */
    void *Un = nullptr, *A = nullptr, *B = nullptr, *shadow = nullptr;
    void *rhofloor = nullptr;  /* [N] rhocr * Uinitial[0] */
    void *edge = nullptr;      /* [2][9][NY][NZ] x-boundary face fluxes */
    double *u64_scratch = nullptr;   /* [9][N] FP64 export buffer */
    double *dtpart = nullptr;  /* Setdt partials of the last stage-3 launch */
    int dt_blocks = 0;
    bool dt_valid = false;
    int nseg = 1, seglen = 1, tiles_j = 1, tiles_k = 1;
};

using AllocFn = std::function<int(size_t, void **)>;

int fused_alloc(FusedState &fs, const DevGeom &g, const DevState &d, bool fp32, AllocFn alloc);
/* rhofloor = rhocr * Uinitial[0] (saveinitial.h:44), from d.Uinit */
int fused_set_rhofloor(hipStream_t s, FusedState &fs, const DevState &d, double rhocr);
int fused_load_state(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const double *U64);
int fused_export(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,
                 double *V64, const double *unused);
int fused_setdt(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c);
int fused_stage(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,
                const StageArgs &a, hipEvent_t ev0, hipEvent_t ev1);
int fused_diag(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d);

}  // namespace cyl
#endif
