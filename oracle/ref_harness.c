/*
 * ref_harness.c -- drives the reference's OWN routines (compiled from the
 * headers under /root/reference, never copied) on a small grid, so the
 * oracle restatement can be compared with the real thing bit for bit.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/build_ref.py into oracle/_ref/
 * (git-ignored).  Sizes come from a generated ref_input.h that takes the
 * place of the reference's input.h (its grid sizes are compile-time
 * #defines, input.h:45-72); everything else is the reference's code.
 *
 * Deliberately NOT compiled: spiralgravity.h (it calls cot(), which no libm
 * defines, spiralgravity.h:39).  Gravity is Grxaxi.. (from the file, from
 * the reference's InitialGravity or from GalacticDisc3) with the spiral part
 * off.  Initial states come from the input file, or (flags bit 2) from the
 * reference's GalacticDisc3 (model.h, which needs perturbation.h).
 *
 *   usage: ref_harness IN OUT NSTEPS RK [DUMPDIR]
 *   IN : int64 NX,NY,NZ,flags; V_init[9N]; V_cur[9N]; grx[NX*NZ]; grz[NX*NZ]
 *        flags bit0: use InitialGravity() instead of the tables
 *              bit1: U_init[9N], U_cur[9N] follow (conserved states given)
 *              bit2: initial state and gravity from GalacticDisc3()
 *   OUT: int64 NX,NY,NZ,nsteps; dt[n]; ch[n]; U[9N]; V[9N]; F,G,H[9*NF] of the
 *        last stage; x,xm[NX]; y[NY]; z[NZ]; minminlength; g1,g2,g3,g5,
 *        dvolume,minlength [NX*NZ] (j = 1); grxaxi,grzaxi[NX*NZ];
 *        x_absorb_factor[11]; c_absorb_factor[(irin+1)*NZ]
 *   DUMPDIR: if given, BinaryOutput() of the final state is written there,
 *            and GDLOutputCylinder()'s text files (nout = 7) into DUMPDIR.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include <unistd.h>
#include <sys/stat.h>
#include <omp.h>

#include "maxmin.h"
#include "constant.h"
#include "ref_input.h"
#include "commonfield.h"
#include "commontemp.h"
#include "commongrid.h"
#include "cartesiangrid.h"
#include "cylindricalgrid.h"
#include "setdt.h"
#include "xboundary.h"
#include "yboundary.h"
#include "zboundary.h"
#include "centralboundary.h"
#include "xvlrboundary.h"
#include "yvlrboundary.h"
#include "zvlrboundary.h"
#include "vtou.h"
#include "utov.h"
#include "saveinitial.h"
#include "initialgravity.h"
#include "cooling.h"
#include "source.h"
#include "vlrtoulr.h"
#include "vlrulrtoflr.h"
#include "vlrulrtoglr.h"
#include "vlrulrtohlr.h"
#include "primitivetof.h"
#include "primitivetog.h"
#include "primitivetoh.h"
#include "glmsourcesolver.h"
#include "limiterx.h"
#include "limitery.h"
#include "limiterz.h"
#include "hlldx.h"
#include "hlldy.h"
#include "hlldz.h"
#include "hllx.h"
#include "hlly.h"
#include "hllz.h"
/* scheme selection = the reference's own alternatives, chosen the way its
 * author does (firststep.h:37-46 comment one call in, one out) */
#ifdef REF_LIMITER_MC
#define MinmodLimiterX MCLimiterX
#define MinmodLimiterY MCLimiterY
#define MinmodLimiterZ MCLimiterZ
#endif
#ifdef REF_LIMITER_VL
#define MinmodLimiterX VanLeerLimiterX
#define MinmodLimiterY VanLeerLimiterY
#define MinmodLimiterZ VanLeerLimiterZ
#endif
#ifdef REF_LIMITER_SB
#define MinmodLimiterX SuperbeeLimiterX
#define MinmodLimiterY SuperbeeLimiterY
#define MinmodLimiterZ SuperbeeLimiterZ
#endif
#ifdef REF_SOURCE_AXIC
#define SpiralGravityCooling AxiGravityCooling
#endif
#ifdef REF_SOURCE_AXI
#define SpiralGravityCooling AxiGravity
#endif
#ifdef REF_SOURCE_UGZ
#define SpiralGravityCooling UniformGz
#endif
#ifdef REF_SOURCE_NONE
#define SpiralGravityCooling NoSource
#endif
#ifdef REF_SOLVER_HLL
#define HLLDX HLLX
#define HLLDY HLLY
#define HLLDZ HLLZ
#endif
#include "halfstep.h"
#include "fullstep.h"
#include "firststep.h"
#include "secondstep.h"
#include "thirdstep.h"
#include "check.h"
#include "perturbation.h"
#include "model.h"
#include "output.h"

#define NCELL ((size_t)ixmax * jymax * kzmax)
#define NFACE ((size_t)ixmax1 * jymax1 * kzmax1)

static void rd(FILE *f, void *p, size_t n)
{
    if (fread(p, 1, n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}
static void wr(FILE *f, const void *p, size_t n)
{
    if (fwrite(p, 1, n, f) != n) { fprintf(stderr, "short write\n"); exit(2); }
}

int main(int argc, char **argv)
{
    long long hdr[4];
    int i, j, k, m;
    if (argc < 5) { fprintf(stderr, "usage: %s IN OUT NSTEPS RK [DUMPDIR]\n", argv[0]); return 2; }
    const int nsteps = atoi(argv[3]);
    const int rk = atoi(argv[4]);
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) { perror(argv[1]); return 2; }
    rd(fi, hdr, sizeof hdr);
    if (hdr[0] != ixmax || hdr[1] != jymax || hdr[2] != kzmax) {
        fprintf(stderr, "size mismatch: file %lld %lld %lld, build %d %d %d\n", hdr[0], hdr[1], hdr[2],
                ixmax, jymax, kzmax);
        return 2;
    }
    static double Vcur[9][ixmax][jymax][kzmax];
    static double Uini[9][ixmax][jymax][kzmax], Ucur[9][ixmax][jymax][kzmax];
    static double grx[ixmax][kzmax], grz[ixmax][kzmax];
    rd(fi, V, sizeof V);
    rd(fi, Vcur, sizeof Vcur);
    rd(fi, grx, sizeof grx);
    rd(fi, grz, sizeof grz);
    if (hdr[3] & 2) {            /* conserved states given as well */
        rd(fi, Uini, sizeof Uini);
        rd(fi, Ucur, sizeof Ucur);
    }
    fclose(fi);

    CylindricalGrid();
    if (hdr[3] & 4) {
        /* the reference's own initial state: GalacticDisc3 (model.h:149-327)
         * sets V, U (with its z boundaries), the axisymmetric gravity and
         * SaveInitial; the input states are ignored */
        GalacticDisc3();
    } else {
        /* initial state -> Uinitial/Vinitial/rhofloor (saveinitial.h) */
        if (hdr[3] & 2) memcpy(U, Uini, sizeof U);
        else VtoU();
        SaveInitial();
        /* current state */
        memcpy(V, Vcur, sizeof V);
        if (hdr[3] & 2) memcpy(U, Ucur, sizeof U);
        else VtoU();
    }
    if (hdr[3] & 4) {
        /* gravity as GalacticDisc3 left it */
    } else if (hdr[3] & 1) {
        InitialGravity();
    } else {
        for (i = 0; i < ixmax; i++)
            for (j = 0; j < jymax; j++)
                for (k = 0; k < kzmax; k++) {
                    Grxaxi[i][j][k] = grx[i][k];
                    Gryaxi[i][j][k] = 0.0;
                    Grzaxi[i][j][k] = grz[i][k];
                }
    }
    double *dts = calloc(nsteps + 1, sizeof(double)), *chs = calloc(nsteps + 1, sizeof(double));
    int done = 0;
    t = 0.0;
    for (int n = 0; n < nsteps; n++) {
        Setdt();
        if (isinf(dt) || dt < dtmin) break;
        /* SpiralGravity with the spiral part switched off */
        for (i = 0; i < ixmax; i++)
            for (j = 0; j < jymax; j++)
                for (k = 0; k < kzmax; k++) {
                    Grx[i][j][k] = Grxaxi[i][j][k];
                    Gry[i][j][k] = Gryaxi[i][j][k];
                    Grz[i][j][k] = Grzaxi[i][j][k];
                }
        /* TVDRK (tvdrungekutta.h:16-78) with the stage set chosen by RK */
        psicoef = exp(-dt * ch * oneocr * 0.5);
        if (rk == 2) {
            Halfstep();
            Fullstep();
        } else {
            FirstStep();
            SecondStep();
            ThirdStep();
        }
        dts[n] = dt;
        chs[n] = ch;
        t += dt;
        done++;
    }
    FILE *fo = fopen(argv[2], "wb");
    if (!fo) { perror(argv[2]); return 2; }
    long long oh[4] = {ixmax, jymax, kzmax, done};
    wr(fo, oh, sizeof oh);
    wr(fo, dts, sizeof(double) * done);
    wr(fo, chs, sizeof(double) * done);
    wr(fo, U, sizeof U);
    wr(fo, V, sizeof V);
    wr(fo, F, sizeof F);
    wr(fo, G, sizeof G);
    wr(fo, H, sizeof H);
    wr(fo, x, sizeof x);
    wr(fo, xm, sizeof xm);
    wr(fo, y, sizeof y);
    wr(fo, z, sizeof z);
    wr(fo, &minminlength, sizeof minminlength);
    {
        static double tab[6][ixmax][kzmax];
        for (i = 0; i < ixmax; i++)
            for (k = 0; k < kzmax; k++) {
                const int jj = 1;
                const int inside = i >= 1 && i < ixmax1 && k >= 1 && k < kzmax1;
                tab[0][i][k] = inside ? g1[i][jj][k] : 0.0;
                tab[1][i][k] = inside ? g2[i][jj][k] : 0.0;
                tab[2][i][k] = inside ? g3[i][jj][k] : 0.0;
                tab[3][i][k] = inside ? g5[i][jj][k] : 0.0;
                tab[4][i][k] = dvolume[i][jj][k];
                tab[5][i][k] = minlength[i][jj][k];
                grx[i][k] = Grxaxi[i][jj][k];
                grz[i][k] = Grzaxi[i][jj][k];
            }
        wr(fo, tab, sizeof tab);
        wr(fo, grx, sizeof grx);
        wr(fo, grz, sizeof grz);
    }
    wr(fo, x_absorb_factor, sizeof x_absorb_factor);
    wr(fo, c_absorb_factor, sizeof c_absorb_factor);
    fclose(fo);
    if (argc > 5) {
        if (chdir(argv[5]) != 0) { perror(argv[5]); return 2; }
        nout = 0;
        BinaryOutput();
        /* the reference's text writer, into DUMPDIR itself, dump number 7 */
        nout = 7;
        GDLOutputCylinder();
    }
    (void)m;
    return 0;
}
