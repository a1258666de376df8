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
     *   shadow  pre-Check stage values of the cells the ghost map reads:
     *           [9][NS] for the r and phi sides, then [3][NX][NY][9] for the
     *           z sides (planes k = 1, 2, NZ-2, packed so that the ghost pass
     *           reads them from consecutive addresses) */
    void *Un = nullptr, *A = nullptr, *B = nullptr, *shadow = nullptr;
    void *rhofloor = nullptr;  /* [N] rhocr * Uinitial[0] */
    void *edge = nullptr;      /* [2][9][NY][NZ] x-boundary face fluxes */
    double *u64_scratch = nullptr;   /* [9][N] FP64 export buffer */
    double *dtpart = nullptr;  /* Setdt partials of the last stage-3 launch */
    double *dtpart2 = nullptr; /* [kDtSlices][2]: the partials reduced per slice */
    int dt_blocks = 0;
    int dt_nb = 0;             /* partials the last stage wrote */
    /* the stage as two kernels, the r-march (writing each cell's
     * divergence into div, [9][NS]) and a per-cell update (the default), or
     * as one kernel (CYLMHD_STAGE_KERNELS=1 in the environment at context
     * creation) */
    bool split = false;
    void *div = nullptr;
    bool dt_valid = false;
    int nseg = 1, seglen = 1, tiles_j = 1, tiles_k = 1;
    /* r-segment counts the first stage-1 launch times (the fastest is kept) */
    int ncand = 0, cand[8] = {0};
    bool tuned = false;
    /* padded layout of the state arrays: component stride NS, k rows of
     * NZP cells with k = 1 at offset 32 (128-byte aligned FP32 rows) */
    size_t NS = 0;
    int NZP = 0, NZ = 0;
    /* packed z-side ghosts (the discs of a batch, and a single disc on the
     * two-kernel stage): the z-side ghost cells (k = 0, NZ-1) of each state
     * buffer are not in its [9][NS] body but packed after it,
     * [2][NX][NY][9], so the ghost pass writes 36 contiguous bytes per
     * cell; nstate = elements per state buffer.  batch: a disc of a batch
     * (advanced only by the batch) */
    bool zg = false, batch = false;
    size_t nstate = 0;
    /* geometry packed for the stage kernel (few base pointers):
     *   itab [NX][IT] FP64  x, xm, oneox, oneodx, xabs, spiral r factors
     *   ikg  [NX*NZ][4] T   g1, g2, g3, g5; FP32 state: [NX*NZ][8], then the
     *                       offsets of components 0, 2, 5, 7 (the stored
     *                       values of those are deviations from them: the
     *                       phi-mean of the initial state)
     *   ikd  [NX*NZ][4] FP64 minlength (1/minlength in the FP32 build), Grxaxi, Grzaxi, spiral q
     *   ktab [NZ][KT] FP64  oneodz, z, spiral z factors, zabs */
    double *itab = nullptr, *ikd = nullptr, *ktab = nullptr;
    void *ikg = nullptr;
};

using AllocFn = std::function<int(size_t, void **)>;

/* one set of entry points per state precision (fused.hip built twice) */
#define CYL_FUSED_DECLS                                                                                   \
    int alloc(FusedState &fs, const DevGeom &g, const DevState &d, AllocFn alloc);                       \
    int set_initial_tables(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, double rhocr); \
    int load_state(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const double *U64); \
    int export_state(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,  \
                     double *V64);                                                                        \
    int setdt(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c);       \
    int stage(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,        \
              const StageArgs &a, hipEvent_t ev0, hipEvent_t evm, hipEvent_t ev1);                        \
    int diag(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d);
namespace f32 { CYL_FUSED_DECLS }
namespace f64 { CYL_FUSED_DECLS }
#undef CYL_FUSED_DECLS

inline int fused_alloc(FusedState &fs, const DevGeom &g, const DevState &d, bool fp32, AllocFn a)
{
    return fp32 ? f32::alloc(fs, g, d, a) : f64::alloc(fs, g, d, a);
}
/* from d.Uinit: rhofloor = rhocr * Uinitial[0] (saveinitial.h:44) and,
 * for FP32 state, the offsets ubar */
inline int fused_set_initial_tables(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, double rhocr)
{
    return fs.fp32 ? f32::set_initial_tables(s, fs, g, d, rhocr) : f64::set_initial_tables(s, fs, g, d, rhocr);
}
inline int fused_load_state(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const double *U64)
{
    return fs.fp32 ? f32::load_state(s, fs, g, d, U64) : f64::load_state(s, fs, g, d, U64);
}
inline int fused_export(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,
                        double *V64)
{
    return fs.fp32 ? f32::export_state(s, fs, g, d, c, V64) : f64::export_state(s, fs, g, d, c, V64);
}
inline int fused_setdt(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c)
{
    return fs.fp32 ? f32::setdt(s, fs, g, d, c) : f64::setdt(s, fs, g, d, c);
}
/* ev0 / ev1 bracket the stage's hot kernel(s); evm, between the r-march
 * and the per-cell update of the two-kernel stage, is recorded only there */
inline int fused_stage(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d, const Consts &c,
                       const StageArgs &a, hipEvent_t ev0, hipEvent_t evm, hipEvent_t ev1)
{
    return fs.fp32 ? f32::stage(s, fs, g, d, c, a, ev0, evm, ev1) : f64::stage(s, fs, g, d, c, a, ev0, evm, ev1);
}
inline int fused_diag(hipStream_t s, FusedState &fs, const DevGeom &g, const DevState &d)
{
    return fs.fp32 ? f32::diag(s, fs, g, d) : f64::diag(s, fs, g, d);
}

}  // namespace cyl
#endif
