/*
 * cylmhd.h -- C ABI of the MI355X-native cylindrical HLLD MHD time step.
 *
 * The reference (github-nakasho/171026-1-3dhlld-cyli-gal) has no API: every
 * routine is `int f(void)` over file-scope globals (commonfield.h:114-141,
 * commontemp.h:160-184, commongrid.h:42-62) sized by compile-time #defines
 * (input.h:45-72).  The entry points below replace exactly the calls that
 * main.c makes on the hot path and at start-up; each declaration names the
 * reference routine it stands in for.  All arrays crossing this boundary are
 * host FP64 arrays in the reference layout  A[m][i][j][k]  (component-major,
 * k fastest, ghosts included: NX = ix+2, NY = jy+2, NZ = kz+2), so a driver
 * written against the reference globals can hand its arrays over unchanged
 * and the binary dumps stay byte-identical.
 *
 * Every function returns 0 on success or a negative CYLMHD_E* code; nothing
 * in the library calls exit().  One context = one GPU stream; a context is
 * not re-entrant (like the reference's globals) but contexts are independent.
 */
#ifndef CYLMHD_H
#define CYLMHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define CYLMHD_OK          0
#define CYLMHD_EINVAL     -1   /* bad argument / inconsistent parameters   */
#define CYLMHD_ENOMEM     -2   /* device or host allocation failed         */
#define CYLMHD_EHIP       -3   /* a HIP runtime call failed                */
#define CYLMHD_EIO        -4   /* dump / restart file could not be opened  */
#define CYLMHD_EUNSUP     -5   /* option not available (reserved: every current
                                  option runs on every path)                  */

/* ---- scheme selectors (the reference picks these by commenting calls) -- */
/* limiter: firststep.h:37-38 (MinmodLimiterX is the active one)          */
#define CYLMHD_LIMITER_MINMOD 0   /* limiterx.h:171-217 (MUSCL-Hancock form) */
#define CYLMHD_LIMITER_MC     1   /* limiterx.h:335-385                      */
#define CYLMHD_LIMITER_VANLEER  2 /* limiterx.h:266-326                      */
#define CYLMHD_LIMITER_SUPERBEE 3 /* limiterx.h:394-449                      */
/* Riemann solver: firststep.h:45-46                                       */
#define CYLMHD_SOLVER_HLLD    0   /* hlldx.h:28-499                          */
#define CYLMHD_SOLVER_HLL     1   /* hllx.h:29-184                           */
/* cooling law used by the energy source and the CFL limit                 */
#define CYLMHD_COOL_NONE      0
#define CYLMHD_COOL_RAYMOND   1   /* cooling.h:22-78 (int-truncated)         */
#define CYLMHD_COOL_POWERLAW  2   /* Lambda = lambda0 * T^index for T > tmin */
/* source terms: the stage functions call SpiralGravityCooling
 * (firststep.h:85); the reference's alternatives are its other functions in
 * source.h, chosen by swapping that call                                  */
#define CYLMHD_SOURCE_SPIRAL_COOLING 0  /* source.h:26-134 (default)         */
#define CYLMHD_SOURCE_AXI_COOLING    1  /* source.h:143-200: Grxaxi/Grzaxi everywhere, cooling */
#define CYLMHD_SOURCE_AXI            2  /* source.h:207-261: the same, no cooling term */
#define CYLMHD_SOURCE_UNIFORM_GZ     3  /* source.h:269-300: g_z = -1, no hoop terms */
#define CYLMHD_SOURCE_NONE           4  /* source.h:308-349: hoop and induction terms only */
/* axisymmetric gravity model                                              */
#define CYLMHD_GRAV_GALACTIC  0   /* initialgravity.h:22-79                  */
#define CYLMHD_GRAV_POINTMASS 1   /* -GM r_vec/|r|^3 (Keplerian discs)       */
/* device path                                                             */
#define CYLMHD_PATH_PLAIN64   0   /* pass-by-pass FP64, reference data flow  */
#define CYLMHD_PATH_FUSED64   1   /* fused RK stage (r-march + update), FP64 */
#define CYLMHD_PATH_FUSED32   2   /* fused, FP32 state + FP64 recovery/sums  */

/* ---- run-time parameter set: input.h:45-102 + constant.h:62-89 --------- */
typedef struct cylmhd_params {
    /* grid (input.h:45-76); NX = ix+2, NY = jy+2, NZ = kz+2 (margin = 2) */
    int    ix, jy, kz;
    double lx, lz;               /* start of the stretched zone in r and z */
    double x0, y0, z0;           /* first interior cell centre             */
    double dx0, dy0, dz0;        /* base spacings                          */
    double stretchratio, dxmax, dzmax;
    double rin;                  /* central absorbing radius               */
    /* run control (input.h:58-59, 78-81) */
    double cfl, dtmin;
    double rhocr, prcr, betacr;
    /* constants (constant.h:62-79, 189-196) */
    double gm, gm1, oneogm1;     /* 5/3, 2/3, 1.5 as separate literals     */
    double cr;                   /* GLM: psi damping exp(-dt ch/cr/2)      */
    double te0;                  /* temperature unit of Cooling()          */
    /* spiral arm / rotating frame (input.h:88-96) */
    double epssp0, tausp, scalez, scaler, scalephase;
    int    armnum;
    double omegasp, omegapattern, pitch;
    /* scheme selection */
    int    limiter, solver, rk;  /* rk = 3 (FirstStep..ThirdStep) or 2     */
    int    cooling, gravity;
    int    source;               /* CYLMHD_SOURCE_*                        */
    double gm_point;             /* GM of CYLMHD_GRAV_POINTMASS            */
    double cool_lambda0, cool_tmin, cool_index;
    int    path;                 /* CYLMHD_PATH_*                          */
    /* synthetic disc initial state (benchmark discs; host-side only) */
    double disc_h, disc_floor, disc_cs2, disc_beta, beta_lo, beta_hi;
    int    beta_mode;            /* 0: constant disc_beta, 1: log-uniform per annulus */
    double pert_amp;
    uint64_t seed;
} cylmhd_params;

/* Fill *p with the reference's own input.h / constant.h values. */
int cylmhd_params_default(cylmhd_params *p);

/* Fill *p for one of the benchmark discs (r in [0.1,1], z in [-0.5,0.5],
 * point-mass gravity GM = 1, inertial frame, no cooling unless stated,
 * pressure floor prcr = 1e-3, i.e. 10 % of the disc's p/rho):
 *   1: 64x64x32, minmod, CPU-sized;     2: 256x256x128, MC, spiral 5%;
 *   3: 512x512x256, MC, beta log-uniform [0.1,100] per annulus;
 *   4: one disc of the batch (64x64x32 default), minmod, beta [1,100];
 *   5: 768x768x384, MC, power-law cooling, spiral 10%, GLM alpha 0.1.
 * nr/nphi/nz > 0 override the grid size (same disc, other resolution). */
int cylmhd_params_config(cylmhd_params *p, int config, int nr, int nphi, int nz);

/* ---- host-side set-up (no GPU) ----------------------------------------- */
/* Grid: cylindricalgrid.h:25-227.  x[NX], y[NY], z[NZ] cell centres. */
int cylmhd_grid_host(const cylmhd_params *p, double *x, double *y, double *z);
/* Synthetic disc primitives V[9][NX][NY][NZ] (all cells incl. ghosts). */
int cylmhd_model_disc(const cylmhd_params *p, double *V);
/* Synthetic disc as a consistent (U, V) pair: interior from the disc,
 * U = VtoU(V), ghosts by the RK3 cell-boundary set, ghost V = UtoV(U). */
int cylmhd_model_disc_state(const cylmhd_params *p, double *U, double *V);
/* Ghost cells of (U, V) from the interior by the RK3 cell-boundary set
 * (Y periodic, X axis, X right free, Z periodic: firststep.h:125-157),
 * ghost V = UtoV(ghost U).  Host arrays, in place. */
int cylmhd_fill_ghosts_host(const cylmhd_params *p, double *U, double *V);
/* GalacticDisc3 (model.h:149-327) primitives, without the U-side z BCs. */
int cylmhd_model_galactic_disc3(const cylmhd_params *p, double *V);
/* VtoU (vtou.h:20-62) on host arrays. */
int cylmhd_vtou_host(const cylmhd_params *p, const double *V, double *U);
/* Check (check.h:23-188) on host arrays, in place, as RestartRead runs it
 * after its VtoU (restartread.h:132-139).  rho_init = the density U[0] of
 * the state SaveInitial saved (rhofloor = rho_init * rhocr, saveinitial.h:44),
 * or NULL before any SaveInitial (RestartRead(0), main.c:250: the
 * reference's rhofloor is then still all zero).  counts (optional) = the
 * density, pressure and beta floor hits (check.h:180-183). */
int cylmhd_check_host(const cylmhd_params *p, const double *rho_init, double *U, double *V, int counts[3]);

/* ---- device context: the hot path --------------------------------------- */
typedef struct cylmhd_ctx cylmhd_ctx;

/* Allocates device state; builds grid, geometry and gravity tables
 * (CylindricalGrid cylindricalgrid.h:25, InitialGravity initialgravity.h:22). */
int cylmhd_create(const cylmhd_params *p, cylmhd_ctx **out);
int cylmhd_destroy(cylmhd_ctx *c);
int cylmhd_dims(const cylmhd_ctx *c, int *nx, int *ny, int *nz);

/* SaveInitial (saveinitial.h:18-50): U,V become the current state AND the
 * Uinitial/Vinitial/rhofloor references of the absorbing layers and floors. */
int cylmhd_set_initial(cylmhd_ctx *c, const double *U, const double *V);
/* Replace the current state (restart: RestartRead restartread.h:9-143 after
 * its VtoU + Check -- cylmhd_vtou_host + cylmhd_check_host -- or any state
 * the caller holds). */
int cylmhd_upload_state(cylmhd_ctx *c, const double *U, const double *V);
/* Copy the current state back (V feeds BinaryOutput, output.h:1202-1414). */
int cylmhd_download_state(cylmhd_ctx *c, double *U, double *V);

/* Setdt (setdt.h:24-94): dt = cfl/max, ch = minminlength*max, and the
 * (i,j,k) of the limiting cell ("dt factor mesh").  main.c:324. */
int cylmhd_setdt(cylmhd_ctx *c, double *dt, double *ch, int ijk[3]);
/* SpiralGravity (spiralgravity.h:20-124) at time t.  main.c:341. */
int cylmhd_gravity(cylmhd_ctx *c, double t);
/* TVDRK (tvdrungekutta.h:16-78): one full time step of dt with GLM speed
 * ch; RK3 = FirstStep/SecondStep/ThirdStep, RK2 = Halfstep/Fullstep.
 * Enqueued on the context stream; returns without waiting.  main.c:348. */
int cylmhd_step(cylmhd_ctx *c, double dt, double ch);
/* Setdt that allocates nothing, on a caller's device arrays.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture set-up, tight serving loops, harnesses that
 * forbid hipMalloc); the result equals cylmhd_setdt's on the same state.
 * V: device [9][NX][NY][NZ] FP64 primitives; minlength: device [NX][NZ]
 * FP64 (cylindricalgrid.h:146); minminlength its minimum (setdt.h:89);
 * stream: a hipStream_t to run on, or NULL.  The block partials live in a
 * static device array of 1024 (max, index) pairs (16 KB per GPU, held until
 * the process ends); grids of 2^53 cells or more are rejected.  Calls that
 * overlap on one GPU race on that array.  Waits for the stream. */
int cylmhd_setdt_noalloc(const cylmhd_params *p, const double *V, const double *minlength,
                         double minminlength, void *stream, double *dt, double *ch, int ijk[3]);
/* Check() floor counters (check.h:180-183) of the stages of the last step:
 * counts[3*s + 0..2] = rho, pressure, beta hits of stage s. */
int cylmhd_floor_counts(cylmhd_ctx *c, int counts[9]);

typedef struct cylmhd_diag {
    double mass;      /* sum U0 dV over interior cells                    */
    double energy;    /* sum U7 dV                                        */
    double divb_l1;   /* sum |div B| dV (face-averaged finite volume)     */
    double psi_l1;    /* sum |psi| dV                                     */
} cylmhd_diag;
/* Conserved totals and div B (not in the reference; output.h:577-638 is a
 * commented-out Cartesian divB).  FP64 on every path. */
int cylmhd_diagnostics(cylmhd_ctx *c, cylmhd_diag *d);

/* Wait for the context stream; returns the first asynchronous error. */
int cylmhd_sync(cylmhd_ctx *c);
/* Average duration (ms) of the named kernel class over the launches since
 * the last reset, timed with HIP events on the context stream:
 * which = 0: the per-stage kernel(s) of the active path (the flux and update
 * passes of the plain path; the stage kernel, or the r-march and the
 * per-cell update together, of the fused paths); which = 1 / 2: the r-march
 * / the per-cell update of a two-kernel stage alone (launches = 0 where the
 * stage is one kernel). */
int cylmhd_kernel_timing(cylmhd_ctx *c, int which, int reset, double *avg_ms, int *launches);
const char *cylmhd_error_string(int code);

/* ---- batching form: many independent discs in one launch per stage ----
 * Not in the reference (it runs one disc per process).  Every disc keeps a
 * full context (create with CYLMHD_PATH_FUSED32), reachable through
 * cylmhd_batch_disc for set-up, dumps and diagnostics; the batch advances
 * them together: one work list of (disc, tile, r-segment) items, one
 * persistent launch per RK stage over it, per-disc Setdt on the device (each
 * disc its own dt; a disc whose dt falls below dtmin stops, the others go
 * on), and the whole step captured once into a hipGraph and replayed. */
typedef struct cylmhd_batch cylmhd_batch;

/* Parameters of the benchmark batch: n discs, each dimension drawn from
 * {32, 48, 64, 96, 128} with nz <= nr, beta log-uniform in [1, 100] per
 * annulus, spiral amplitude uniform in [0, 0.1], its own seed; minmod +
 * HLLD, RK3, FP32 path.  p must hold n entries. */
int cylmhd_params_batch(cylmhd_params *p, int n, uint64_t seed);
/* All discs must share limiter, solver and RK order (else CYLMHD_EINVAL);
 * each disc may run its own source function (p.source). */
int cylmhd_batch_create(const cylmhd_params *p, int n, cylmhd_batch **out);
int cylmhd_batch_destroy(cylmhd_batch *b);
/* The context of disc d (owned by the batch; do not destroy it).  Set-up,
 * dumps, slices and diagnostics work on it; cylmhd_step does not (the batch
 * keeps its own state layout): cylmhd_batch_run advances it. */
int cylmhd_batch_disc(cylmhd_batch *b, int d, cylmhd_ctx **ctx);
/* Autotuner: times `steps` steps for each stage form x tile shape x
 * r-segment length x launch grid with one tile for every disc, then, per
 * class of disc size (the discs of one phi x z size), the other tiles for
 * that class; works on the current states and each disc's time, steps and
 * alive flag, puts them back after every candidate and keeps the fastest.
 * Stage forms: one kernel per stage (each work item does the whole stage)
 * or two (the r-march per work item, then one per-cell update over every
 * disc).  Tiles (phi x z cells, threads, i.e. waves per tile): 0: 14x16,
 * 256; 1: 6x32, 256; 2: 6x16, 128; 3: 10x16, 192 (2 and 3 for minmod and
 * MC with HLLD only).  choice (optional, 4 ints) = the common tile of the
 * first pass, segment length (0 = whole r), grid (0 = as many blocks as are
 * resident at once), kernels per stage (1, 2). */
int cylmhd_batch_tune(cylmhd_batch *b, int steps, int choice[4]);
/* The tile disc d runs with: tile[0..3] = shape, phi cells, z cells, threads. */
int cylmhd_batch_disc_tile(cylmhd_batch *b, int d, int tile[4]);
/* Advance every live disc nsteps steps (graph replays); *ms (optional) =
 * device time of the replays from HIP events on the batch stream. */
int cylmhd_batch_run(cylmhd_batch *b, int nsteps, double *ms);
/* Time, last dt, steps taken and whether disc d is still running. */
int cylmhd_batch_status(cylmhd_batch *b, int d, double *t, double *dt, int *steps, int *alive);
/* Check() floor hits of disc d summed over the steps of the last runs. */
int cylmhd_batch_floor_counts(cylmhd_batch *b, int d, int counts[9]);

/* ---- dump / restart byte format (host only) ---------------------------- */
/* BinaryOutput (output.h:1202-1414): ./<dir>/GridTime.dat + r, phi, z,
 * rho, vr, vphi, vz, br, bphi, bz, pr, psi as raw native-endian FP64. */
int cylmhd_binary_output(const char *dir, int nx, int ny, int nz, double t,
                         const double *x, const double *y, const double *z,
                         const double *V);
/* RestartRead (restartread.h:9-143): reads the nine primitive files. */
int cylmhd_restart_read(const char *dir, int nx, int ny, int nz, double *V);

/* ---- slices of the resident state and the text writers ----------------
 * The reference's text writer is GDLOutputCylinder (output.h:280-384); its
 * 31 IDL scripts read text slices ("%e" one value a line) from ./idlN/,
 * written by a converter that is not in the reference.  cylmhd_slice cuts
 * a plane out of the state on the device (FP64 primitives; the fused paths
 * convert first), so a slice never needs the whole state on the host. */
/* quantity: 0..8 are the primitives V[m] (rho vr vphi vz br bphi bz pr psi) */
#define CYLMHD_Q_BENERGY    9   /* 0.5 |B|^2 (output.h:350-360)                       */
#define CYLMHD_Q_BETA      10   /* pr / (0.5 |B|^2): gas over magnetic pressure        */
#define CYLMHD_Q_TEMP      11   /* pr / rho, the temperature Cooling uses (cooling.h:25) */
#define CYLMHD_Q_ONEOTCOOL 12   /* 2 rho^2 Lambda / (1.5 pr) (setdt.h:56-58)            */
#define CYLMHD_Q_VDOTB     13   /* v . B                                               */
#define CYLMHD_Q_DIVB      14   /* cell div B as cylmhd_diagnostics (0 on ghost cells) */
#define CYLMHD_Q_RHOBZ     15   /* rho Bz, the Faraday-depth integrand                 */
#define CYLMHD_Q_SPIRALPOT 16   /* arm potential eps m1 q s(z) x^2/sr^2 (1+x^2/sr^2)^-1.5
                                   cos(phase) of the last cylmhd_gravity call: the
                                   potential whose -(1/r) d/dphi is the arm's phi force
                                   (spiralgravity.h:93-98)                              */
#define CYLMHD_Q_VPARA     17   /* velocity along the arm: vphi cos(p) - vr sin(p), p the
                                   pitch angle (the arms are phase = const, input.h:96) */
#define CYLMHD_Q_VPERP     18   /* velocity across the arm: vr cos(p) + vphi sin(p)    */
#define CYLMHD_NQ          19
/* plane (a, b are indices with the ghosts counted, as the reference's) */
#define CYLMHD_PLANE_RPHI        0 /* k = a: i outer, j inner, interior i, j (output.h:318-330) */
#define CYLMHD_PLANE_RPHI_ZMEAN  1 /* mean over the interior k, same order             */
#define CYLMHD_PLANE_RPHI_ZINT   2 /* sum over the interior k of q dzm[k-1]            */
#define CYLMHD_PLANE_RZ          3 /* j = a: k outer, i inner, interior (output.h:368-380) */
#define CYLMHD_PLANE_ZR          4 /* j = a: i outer, k inner, interior                */
#define CYLMHD_PLANE_XZ          5 /* cut through the axis: k outer (interior), 2*NX columns:
                                      the opposite half-plane j' = a + (NY-2)/2 (mod the
                                      interior) from i = NX-1 down to 0, then j = a from
                                      i = 0 up; 1 <= a <= NY-2, NY-2 even              */
#define CYLMHD_PLANE_XZ_PHIMEAN  6 /* as XZ with the mean over the interior j on both halves */
#define CYLMHD_PLANE_PHI         7 /* i = a, k = b: the interior j                     */
/* Number of values of a plane. */
int cylmhd_slice_size(const cylmhd_ctx *c, int plane, size_t *n);
/* One slice of the current state into host out[n] (n = cylmhd_slice_size). */
int cylmhd_slice(cylmhd_ctx *c, int quantity, int plane, int a, int b, double *out, size_t n);

/* GDLOutputCylinder (output.h:280-384) on host arrays: <dir>/r.dat, phi.dat,
 * z.dat (interior coordinates) and 2d-density-rphi-<nout>.dat,
 * 2d-pressure-rphi-, 2d-benergy-rphi- (k = 1), 2d-density-rz- (j = 0). */
int cylmhd_gdl_output(const char *dir, int nout, int nx, int ny, int nz,
                      const double *x, const double *y, const double *z, const double *V);
/* The same files from a context, the planes cut on the device. */
int cylmhd_gdl_output_ctx(cylmhd_ctx *c, const char *dir, int nout);

typedef struct cylmhd_idl_planes {
    int kplane;   /* k of the z-plane slices (zplane*, divb0, vdotb, temp-rphi); 1 */
    int k100;     /* k of z100bphi; 100, clipped to NZ-2                              */
    int jplane;   /* j of the rz cuts and of the xz cut through the axis; 1          */
    int azim_i[4];/* i of the azimuthal profiles azimuth-0..3; 0: 1 + (K+1)(NX-2)/5  */
    int azim_k;   /* k of the azimuthal profiles; 0: 1                                */
} cylmhd_idl_planes;
/* The text slices the reference's IDL scripts read, for dump number nout:
 * <root>/idl<nout>/ (and <root>/idl0/) r.txt, phi.txt, z.txt, x.txt, and
 * <root>/idl<nout>/<name>-<nout>.dat for the names the scripts open, each
 * in the order its fltarr + readf expects; <root>/azimuth-K/ (K = 0..3)
 * phi.dat and {potential,rho,te,vperp,vpara}-<nout>.dat, azimuthal profiles
 * at (azim_i[K], azim_k).  planes may be NULL (defaults).
 * jy (NY-2) must be even: the xz cuts pair each phi with phi + pi. */
int cylmhd_idl_output(cylmhd_ctx *c, const char *root, int nout, const cylmhd_idl_planes *planes);

#ifdef __cplusplus
}
#endif
#endif /* CYLMHD_H */
