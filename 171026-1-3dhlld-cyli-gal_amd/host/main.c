/*
 * cylmhd_run -- the reference driver loop (main.c:224-398) on the GPU library.
 *
 * Start-up follows main.c:232-312: grid, then either the model (GalacticDisc3
 * or one of the benchmark discs) and BinaryOutput(0), or a restart from dumps
 * (RestartRead(0) -> SaveInitial, RestartRead(nout), t = nout*tint).  The
 * loop is main.c:316-393: Setdt, the dt guard, SpiralGravity, TVDRK, t += dt,
 * BinaryOutput every tint, stop past tstop.  The stdout lines are the
 * reference's ("nstep = ", "dt factor mesh", "dt = ", the three Check
 * counters of every stage, the dump and stop notices).
 *
 *   cylmhd_run [--config N] [--grid NR NPHI NZ] [--path plain64|fused64|fused32]
 *              [--rk 2|3] [--nstop N] [--tint T] [--tstop T] [--restart NOUT]
 *              [--outdir DIR] [--quiet] [--gdl] [--idl ROOT] [--convert FIRST LAST]
 *              [--param NAME VALUE]...
 * --config 0 (default) is the reference's own set-up: input.h grid and
 * GalacticDisc3.  --param sets one of the run-time parameters that input.h
 * and constant.h fix at compile time in the reference (cfl, dtmin, rhocr,
 * prcr, betacr, cr, rin, epssp0, tausp, omegasp, omegapattern, pitch).
 * --restart NOUT is main.c:244-288: RestartRead(0) -- the nine primitive
 * files of DIR/0, VtoU, Check with the floors not yet set -- then
 * SaveInitial, then RestartRead(NOUT) with the floors of that initial state,
 * t = NOUT*tint, tout = (NOUT+1)*tint.  With every dump: --gdl also writes GDLOutputCylinder's
 * text files (output.h:280-384; main.c:371-372 has the call commented) into
 * DIR, --idl the text slices the reference's IDL scripts read into
 * ROOT/idl<nout>/ and ROOT/azimuth-K/ (the arm potential at the dump time;
 * with --convert at nout * tint, the restart's time, main.c:269).  --convert FIRST LAST runs no step: it reads the dumps
 * DIR/FIRST .. DIR/LAST (their grid must be the one --config/--grid gives)
 * and writes their text files.  Exit status: 0, or 1 on the dt guard / any
 * library error.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/cylmhd.h"

static int fail(const char *what, int rc)
{
    fprintf(stderr, "%s: %s\n", what, cylmhd_error_string(rc));
    return 1;
}

/* the raw coordinate file <dir>/<name> of a dump equals v[n] */
static int same_coords(const char *dir, const char *name, const double *v, int n)
{
    char path[4200];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) return 0;
    int ok = 1;
    for (int q = 0; q < n && ok; q++) {
        double w;
        ok = fread(&w, sizeof w, 1, f) == 1 && w == v[q];
    }
    fclose(f);
    return ok;
}

static void dirname_of(char *buf, size_t n, const char *outdir, int nout)
{
    snprintf(buf, n, "%s/%d", outdir, nout);
}

/* --param NAME VALUE: the double-valued fields a user of the reference
 * edits in input.h / constant.h */
static int set_param(cylmhd_params *p, const char *name, const char *value)
{
    static const struct { const char *name; size_t off; } tab[] = {
        {"cfl", offsetof(cylmhd_params, cfl)}, {"dtmin", offsetof(cylmhd_params, dtmin)},
        {"rhocr", offsetof(cylmhd_params, rhocr)}, {"prcr", offsetof(cylmhd_params, prcr)},
        {"betacr", offsetof(cylmhd_params, betacr)}, {"cr", offsetof(cylmhd_params, cr)},
        {"rin", offsetof(cylmhd_params, rin)}, {"epssp0", offsetof(cylmhd_params, epssp0)},
        {"tausp", offsetof(cylmhd_params, tausp)}, {"omegasp", offsetof(cylmhd_params, omegasp)},
        {"omegapattern", offsetof(cylmhd_params, omegapattern)}, {"pitch", offsetof(cylmhd_params, pitch)},
    };
    char *end;
    const double v = strtod(value, &end);
    if (end == value || *end) return 0;
    for (size_t q = 0; q < sizeof tab / sizeof tab[0]; q++)
        if (!strcmp(tab[q].name, name)) {
            *(double *)((char *)p + tab[q].off) = v;
            return 1;
        }
    return 0;
}

/* RestartRead(num) (restartread.h:9-143): the nine primitive files, then
 * VtoU and Check; rho_init = NULL while no SaveInitial has set the floors */
static int restart_read(const cylmhd_params *p, const char *dir, int num, int NX, int NY, int NZ,
                        const double *rho_init, double *U, double *V)
{
    int rc, counts[3];
    if ((rc = cylmhd_restart_read(dir, NX, NY, NZ, V))) {
        printf("I can't open %s.\n", dir);
        return fail(dir, rc);
    }
    printf("%d binary read complete.\n", num);
    if ((rc = cylmhd_vtou_host(p, V, U))) return fail("vtou", rc);
    if ((rc = cylmhd_check_host(p, rho_init, U, V, counts))) return fail("check", rc);
    printf("negative density point : %d\nnegative pressure point : %d\ncritical beta point : %d\n", counts[0],
           counts[1], counts[2]);
    return 0;
}

/* ZLeftVzAntisymBSymBoundary then ZLeftFreeBoundary on U (model.h:312-313,
 * zboundary.h:67-96 and 170-190) for the GalacticDisc3 start */
static void galactic_u_bcs(int NX, int NY, int NZ, double *U)
{
    const size_t N = (size_t)NX * NY * NZ;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            const size_t c = ((size_t)i * NY + j) * NZ;
            for (int m = 0; m < 9; m++) U[m * N + c] = m == 3 ? -U[m * N + c + 2] : U[m * N + c + 2];
            U[3 * N + c + 1] = 0.0;
        }
    for (int m = 0; m < 9; m++)
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++) {
                const size_t c = ((size_t)i * NY + j) * NZ;
                U[m * N + c] = U[m * N + c + 1];
            }
}

/* the text files of dump nout from the context's current state at time t
 * (the arm potential of the azimuthal profiles is SpiralGravity's at t; the
 * next step sets it again at the same t) */
static int text_output(cylmhd_ctx *ctx, int gdl, const char *idl, const char *outdir, int nout, double t)
{
    int rc;
    if (gdl && (rc = cylmhd_gdl_output_ctx(ctx, outdir, nout))) return fail("gdl_output", rc);
    if (idl && ((rc = cylmhd_gravity(ctx, t)) || (rc = cylmhd_idl_output(ctx, idl, nout, NULL))))
        return fail("idl_output", rc);
    return 0;
}

int main(int argc, char **argv)
{
    int config = 0, nr = 0, nphi = 0, nz = 0, path = -1, rk = 0, restart = -1, quiet = 0, gdl = 0;
    int conv_first = -1, conv_last = -1;
    const char *idl = NULL;
    const char *pname[32], *pval[32];
    int nparam = 0;
    long nstop = 10000000;
    double tint = 0.05, tstop = 40.0;                      /* input.h:82-84 */
    const char *outdir = ".";
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "--config") && a + 1 < argc) config = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--grid") && a + 3 < argc) { nr = atoi(argv[++a]); nphi = atoi(argv[++a]); nz = atoi(argv[++a]); }
        else if (!strcmp(argv[a], "--path") && a + 1 < argc) {
            const char *s = argv[++a];
            path = !strcmp(s, "plain64") ? CYLMHD_PATH_PLAIN64 : !strcmp(s, "fused64") ? CYLMHD_PATH_FUSED64
                 : !strcmp(s, "fused32") ? CYLMHD_PATH_FUSED32 : -2;
        }
        else if (!strcmp(argv[a], "--rk") && a + 1 < argc) rk = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--nstop") && a + 1 < argc) nstop = atol(argv[++a]);
        else if (!strcmp(argv[a], "--tint") && a + 1 < argc) tint = atof(argv[++a]);
        else if (!strcmp(argv[a], "--tstop") && a + 1 < argc) tstop = atof(argv[++a]);
        else if (!strcmp(argv[a], "--restart") && a + 1 < argc) restart = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--outdir") && a + 1 < argc) outdir = argv[++a];
        else if (!strcmp(argv[a], "--quiet")) quiet = 1;
        else if (!strcmp(argv[a], "--gdl")) gdl = 1;
        else if (!strcmp(argv[a], "--idl") && a + 1 < argc) idl = argv[++a];
        else if (!strcmp(argv[a], "--convert") && a + 2 < argc) { conv_first = atoi(argv[++a]); conv_last = atoi(argv[++a]); }
        else if (!strcmp(argv[a], "--param") && a + 2 < argc && nparam < 32) { pname[nparam] = argv[++a]; pval[nparam++] = argv[++a]; }
        else { fprintf(stderr, "usage: see the header of host/main.c\n"); return 2; }
    }
    if (path == -2) { fprintf(stderr, "unknown --path\n"); return 2; }
    cylmhd_params p;
    int rc = config == 0 ? cylmhd_params_default(&p) : cylmhd_params_config(&p, config, nr, nphi, nz);
    if (rc) return fail("params", rc);
    if (config == 0 && nr > 0) { p.ix = nr; p.jy = nphi; p.kz = nz; p.dy0 = (2.0 * M_PI) / (double)nphi; }
    if (path >= 0) p.path = path;
    if (rk) p.rk = rk;
    for (int q = 0; q < nparam; q++)
        if (!set_param(&p, pname[q], pval[q])) { fprintf(stderr, "--param %s %s: unknown name or bad value\n", pname[q], pval[q]); return 2; }
    const int NX = p.ix + 2, NY = p.jy + 2, NZ = p.kz + 2;
    const size_t N = (size_t)NX * NY * NZ;
    double *x = malloc(sizeof(double) * NX), *y = malloc(sizeof(double) * NY), *z = malloc(sizeof(double) * NZ);
    double *U = malloc(sizeof(double) * 9 * N), *V = malloc(sizeof(double) * 9 * N);
    if (!x || !y || !z || !U || !V) { fprintf(stderr, "out of host memory\n"); return 1; }
    if ((rc = cylmhd_grid_host(&p, x, y, z))) return fail("grid", rc);
    printf("complete grid setting.\n");

    long nstep = 0;
    double t = 0.0, tout = tint;
    int nout = 0;
    char dir[4096];
    cylmhd_ctx *ctx = NULL;
    if ((rc = cylmhd_create(&p, &ctx))) return fail("create", rc);
    if (conv_first >= 0) {
        /* the dump -> text converter: no step */
        if (!gdl && !idl) { fprintf(stderr, "--convert needs --gdl and/or --idl ROOT\n"); return 2; }
        for (int n = conv_first; n <= conv_last; n++) {
            dirname_of(dir, sizeof dir, outdir, n);
            if (!same_coords(dir, "r", x, NX) || !same_coords(dir, "phi", y, NY) || !same_coords(dir, "z", z, NZ)) {
                fprintf(stderr, "%s: missing, or its grid is not the one of --config/--grid\n", dir);
                return 1;
            }
            if ((rc = cylmhd_restart_read(dir, NX, NY, NZ, V))) return fail(dir, rc);
            if ((rc = cylmhd_vtou_host(&p, V, U))) return fail("vtou", rc);
            if ((rc = cylmhd_set_initial(ctx, U, V))) return fail("set_initial", rc);
            if (text_output(ctx, gdl, idl, outdir, n, n * tint)) return 1;   /* main.c:269 */
            printf("converted %s\n", dir);
        }
        cylmhd_destroy(ctx);
        free(x); free(y); free(z); free(U); free(V);
        return 0;
    }
    if (restart >= 0) {
        /* main.c:250-270: RestartRead(0), SaveInitial, RestartRead(nout) */
        double *rho0 = malloc(sizeof(double) * N);
        if (!rho0) { fprintf(stderr, "out of host memory\n"); return 1; }
        dirname_of(dir, sizeof dir, outdir, 0);
        if (restart_read(&p, dir, 0, NX, NY, NZ, NULL, U, V)) return 1;
        if ((rc = cylmhd_set_initial(ctx, U, V))) return fail("set_initial", rc);
        memcpy(rho0, U, sizeof(double) * N);            /* rhofloor = U[0]*rhocr (saveinitial.h:44) */
        nout = restart;
        dirname_of(dir, sizeof dir, outdir, nout);
        if (restart_read(&p, dir, nout, NX, NY, NZ, rho0, U, V)) return 1;
        free(rho0);
        if ((rc = cylmhd_upload_state(ctx, U, V))) return fail("upload", rc);
        t = nout * tint;
        tout = (nout + 1) * tint;
    } else {
        if (config == 0) {
            if ((rc = cylmhd_model_galactic_disc3(&p, V))) return fail("model", rc);
            if ((rc = cylmhd_vtou_host(&p, V, U))) return fail("vtou", rc);
            galactic_u_bcs(NX, NY, NZ, U);
        } else {
            if ((rc = cylmhd_model_disc_state(&p, U, V))) return fail("model", rc);
        }
        if ((rc = cylmhd_set_initial(ctx, U, V))) return fail("set_initial", rc);
        dirname_of(dir, sizeof dir, outdir, nout);
        if ((rc = cylmhd_binary_output(dir, NX, NY, NZ, t, x, y, z, V))) return fail(dir, rc);
        if (text_output(ctx, gdl, idl, outdir, nout, t)) return 1;
    }
    printf("complete model setting.\n");

    const double interior = (double)p.ix * p.jy * p.kz;
    struct timespec t0, t1;
    double loop_s = 0.0;
    long steps_done = 0;
    while (nstep < nstop + 1) {
        nstep++;
        if (!quiet) printf("nstep = %ld\n", nstep);
        double dt, ch;
        int ijk[3];
        clock_gettime(CLOCK_MONOTONIC, &t0);
        if ((rc = cylmhd_setdt(ctx, &dt, &ch, ijk))) return fail("setdt", rc);
        if (!quiet) printf("dt factor mesh %d %d %d \n", ijk[0], ijk[1], ijk[2]);
        if (!quiet) printf("dt = %.3e\n", dt);
        if (isinf(dt) || dt < p.dtmin) {                   /* main.c:331-334 */
            printf("stop due to small dt or inf. program is done.\n");
            cylmhd_destroy(ctx);
            return 1;
        }
        if ((rc = cylmhd_gravity(ctx, t))) return fail("gravity", rc);
        if ((rc = cylmhd_step(ctx, dt, ch))) return fail("step", rc);
        int counts[9];
        if ((rc = cylmhd_floor_counts(ctx, counts))) return fail("floor_counts", rc);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        loop_s += (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
        steps_done++;
        if (!quiet)
            for (int s = 0; s < (p.rk == 2 ? 2 : 3); s++)
                printf("negative density point : %d\nnegative pressure point : %d\ncritical beta point : %d\n",
                       counts[3 * s], counts[3 * s + 1], counts[3 * s + 2]);
        t += dt;
        if (t > tout) {                                       /* main.c:358-378 */
            printf("time=%e.\n output complete.\n", t);
            tout += tint;
            nout++;
            if ((rc = cylmhd_download_state(ctx, NULL, V))) return fail("download", rc);
            dirname_of(dir, sizeof dir, outdir, nout);
            if ((rc = cylmhd_binary_output(dir, NX, NY, NZ, t, x, y, z, V))) return fail(dir, rc);
            if (text_output(ctx, gdl, idl, outdir, nout, t)) return 1;
        }
        if (t > tstop) {                                      /* main.c:383-390 */
            printf("time is over %.3e.\n", tstop);
            break;
        }
    }
    fprintf(stderr, "cylmhd_run: %ld steps, %.3f s in the step loop, %.4g cell-updates/s\n", steps_done, loop_s,
            steps_done ? interior * steps_done / loop_s : 0.0);
    cylmhd_destroy(ctx);
    free(x); free(y); free(z); free(U); free(V);
    return 0;
}
