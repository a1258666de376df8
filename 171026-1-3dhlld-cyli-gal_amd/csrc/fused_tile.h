/*
 * fused_tile.h -- device side of the fused stage (internal): the work-item
 * function of one RK stage over a (phi, z) tile column, the r-boundary faces
 * and the ghost-cell map.  Included by fused.hip (single disc; built once per
 * state precision) and batch.hip (many discs in one launch).  The includer
 * defines FUSED_NS and FT before including it.
 *
 * The stage on one tile marches through r.  At plane i a 3-slot LDS ring
 * holds the primitives of planes i, i+1, i+2 (each plane is loaded and
 * converted UtoV once) plus the phi/z halo of plane i, and
 *   - x: computes the states of cell i+1 (one limiter call gives both of its
 *        faces) and the r-face flux at i+1/2 (the one at i-1/2 is carried
 *        in registers from the previous plane);
 *   - phi, z: the limiter states of every cell of the tile and its one-cell
 *        halo, once per direction, into LDS; then the faces (HLLD/HLL);
 *   - accumulates the divergence in the reference's order
 *        ((g2 F+ - g1 F-) + (g4 G+ - g3 G-)) + (g6 H+ - g5 H-),
 *     adds SpiralGravityCooling, forms the RK stage combination, applies
 *     the absorbing layers (last stage), Check, and stores U;
 *   - writes the pre-Check values of the cells the boundary map reads into
 *     a shadow array (for the ghost pass), and on the last stage the Setdt
 *     partial.
 * Reads per cell and stage: U_in once (plus the halo, mostly L2 hits) and
 * once more at the update, U^n on stages 2-3; writes U_out once.  V, the
 * face fluxes and S never touch HBM.
 *
 * Against the plain path the only change in the arithmetic is that V is
 * recomputed from the stored U each stage instead of being carried from
 * the previous Check (in floored cells the two differ in the last bits).
 * FP32 state keeps the pressure recovery of UtoV, the rebuild of the energy
 * in Check, the RK combination, the CFL reduction and the diagnostics in
 * FP64.
 *
 * Boundary faces: phi is periodic in the cell boundaries, so halo rows wrap
 * (-1 -> NY-3, NY -> 2).  The interface boundaries follow the limiter, as
 * limiter{x,y,z}.h call them:
 *   minmod   r XVlrPeriodic, phi YVlrPeriodic, z ZLeftVlVzAntisymBzsym /
 *            ZRightVrFree
 *   MC       r XVlrPeriodic, phi YVlrPeriodic, z ZVlrFixed (Vinitial)
 *   vanLeer  r XVlrFree, phi YVlrInitialize (Vinitial at j = 0, NY-1),
 *            z ZVlrFree
 *   superbee r XVlrPeriodic, phi YVlrPeriodic, z ZVlrPeriodic
 * The r faces 0 and NX-2 are computed by edge_cell before the stage kernel.
 */
#ifndef CYLMHD_FUSED_TILE_H
#define CYLMHD_FUSED_TILE_H

#include <climits>
#include <type_traits>

#include "fused.h"

namespace cyl {
namespace FUSED_NS {

using FT = FUSED_FT;

constexpr int KOFF = 31;     /* k = 1 at offset 32 of a padded row */

/* C3: cell index in the reference layout (Uinitial, the dt argmax);
 * SI: cell index in the state arrays of this path, whose k rows are padded
 * so that k = 1 (the first interior cell) starts a 128-byte line: a tile row
 * of 16 FP32 cells is then one aligned 64-byte access, and a wavefront's 64
 * consecutive interior k are two whole 128-byte lines (with k = 1 at offset
 * 16, the interior of a 256-cell row spanned nine lines, two of them half
 * padding). */
#define C3(i, j, k) ((((size_t)(i)) * g.NY + (size_t)(j)) * g.NZ + (size_t)(k))
#define SI(i, j, k) ((((size_t)(i)) * g.NY + (size_t)(j)) * (size_t)a.NZP + (size_t)(k) + KOFF)
/* ZG: with the batch layout the z-side ghost cells (k = 0: side 0, k = NZ-1:
 * side 1) of a state buffer live after its [9][NS] body, [2][NX][NY][9], the
 * nine values of a cell together (component m at ZG(...) + m), so that the
 * ghost pass writes them to consecutive addresses */
#define ZG(side, i, j) (9 * a.NS + ((((size_t)(side)) * g.NX + (size_t)(i)) * g.NY + (size_t)(j)) * 9)


/* Loads from tables that no launch writes (geometry rows, packed source
 * terms, boundary data) go through the constant address space: a
 * wave-uniform index then becomes a scalar load instead of a vector load
 * the wave has to wait for. */
using cdp = const __attribute__((address_space(4))) double *;
template <typename X>
__device__ __forceinline__ X ro(const X *p, size_t idx)
{
    return ((const __attribute__((address_space(4))) X *)p)[idx];
}


/* ---- precision-specific UtoV and Check ----------------------------------
 * The FP32 state is stored as its deviation from a fixed per-(i,k) offset,
 * u = U - ubar(i,k), ubar the phi-mean of the initial state for the four
 * components that carry a disc's axisymmetric background -- rho, r rho
 * vphi, B_phi and E (NOFF below); the others are stored as they are, being
 * small themselves: the RK combinations are affine with weights summing
 * to one, so they act on the deviations unchanged, and the rounding of each
 * stored value is relative to the deviation, not to U.  In a high-Mach
 * low-beta cell E is hundreds of times p/(gamma-1): rounding E itself to
 * FP32 every stage would put ~1e-5 relative noise into p, which the flow
 * near the inner edge amplifies.  The full values are formed in FP64 where
 * they feed the pressure recovery.  FP64 state keeps no offset (ub unused). */
constexpr int NOFF = 4;                  /* offset slots: components 0, 2, 5, 7 */
/* the offsets of (i, k) sit in the (i, k) row of the ikg table, after its
 * four g-factors: ikg [NX*NZ][IKW] of T = g1, g2, g3, g5 (, ub0, ub2, ub5,
 * ub7 for FP32 state), so no further table pointer lives in the kernel */
template <typename T> constexpr int ikg_w() { return sizeof(T) == 4 ? 8 : 4; }
__host__ __device__ constexpr int oslot(int m) { return m == 0 ? 0 : (m == 2 ? 1 : (m == 5 ? 2 : (m == 7 ? 3 : -1))); }
template <typename T> struct Prec;

template <> struct Prec<double> {
    static __device__ __forceinline__ void utov(const Consts &c, const double *U, const double *, double oneox, double *V)
    {
        cyl::utov<double>(c, U, oneox, V);
    }
    static __device__ __forceinline__ int check(const Consts &c, double *U, const double *, double *V, double x, double rf)
    {
        return cyl::check<double>(c, U, V, x, rf);
    }
};

/* FP32 state: p = gm1 (E - (rho v)^2/2rho - B^2/2) and the energy rebuild
 * of Check are formed in FP64 (the cancellation points of low-beta and
 * high-Mach cells); the rest stays FP32.  U holds deviations, ub the
 * offsets of the cell. */
template <> struct Prec<float> {
    static __device__ __forceinline__ void utov(const Consts &c, const float *U, const float *ub, double oneox, float *V)
    {
        /* full values of the offset components, formed in FP64 */
        const double d0 = (double)U[0] + (double)ub[0];
        const double d2 = (double)U[2] + (double)ub[1];
        const double d5 = (double)U[5] + (double)ub[2];
        const float rho = (float)d0, mphi = (float)d2;
        const float oneorho = 1.0f / rho;
        const double d1 = U[1], d3 = U[3], d4 = U[4], d6 = U[6];
        const double b2 = d4 * d4 + d5 * d5 + d6 * d6;
        const double r2v2 = d1 * d1 + d2 * d2 * oneox * oneox + d3 * d3;
        V[0] = rho;
        V[1] = U[1] * oneorho;
        V[2] = mphi * oneorho * (float)oneox;
        V[3] = U[3] * oneorho;
        V[4] = U[4];
        V[5] = (float)d5;
        V[6] = U[6];
        V[7] = (float)(c.gm1 * (((double)U[7] + (double)ub[3]) - 0.5 * (r2v2 / d0 + b2)));
        V[8] = U[8];
    }
    static __device__ __forceinline__ int check(const Consts &c, float *U, const float *ub, float *V, double x, float rf)
    {
        int flags = 0;
        const double v2 = (double)V[2] * V[2];
        const double b2 = ((double)V[4] * V[4] + (double)V[5] * V[5] + (double)V[6] * V[6]);
        if (V[0] < rf) {
            flags |= 1;
            V[0] = rf; V[1] = 0.0f; V[3] = 0.0f;
            U[0] = (float)((double)V[0] - ub[0]); U[1] = 0.0f; U[2] = (float)(x * V[0] * V[2] - ub[1]); U[3] = 0.0f;
            U[7] = (float)(0.5 * ((double)V[0] * v2 + b2) + c.oneogm1 * V[7] - ub[3]);
        }
        const float prfloor = V[0] * (float)c.prcr;
        if (V[7] < prfloor) {
            flags |= 2;
            V[1] = 0.0f; V[3] = 0.0f; V[7] = prfloor;
            U[1] = 0.0f; U[3] = 0.0f;
            U[7] = (float)(0.5 * ((double)V[0] * v2 + b2) + c.oneogm1 * V[7] - ub[3]);
        }
        const double oneobeta = 0.5 * b2 / (double)V[7];
        if (oneobeta > c.oneobetacr) {
            flags |= 4;
            V[0] = (float)(c.betacr * oneobeta * V[0]); V[1] = 0.0f; V[3] = 0.0f;
            V[7] = (float)(c.betacr * oneobeta * V[7]);
            U[0] = (float)((double)V[0] - ub[0]); U[1] = 0.0f; U[2] = (float)(x * V[0] * V[2] - ub[1]); U[3] = 0.0f;
            U[7] = (float)(0.5 * ((double)V[0] * v2 + b2) + c.oneogm1 * V[7] - ub[3]);
        }
        return flags;
    }
};

/* packed geometry rows (FusedState::itab, ktab) */
enum { IT_X, IT_XM, IT_OX, IT_ODX, IT_XABS, IT_X2S2, IT_O1X2S2, IT_XS2O15, IT_X2S2O15, IT = 12 };
enum { KT_ODZ, KT_Z, KT_SZS, KT_ONEPA, KT_OZ2S2, KT_ZABS, KT = 8 };

template <typename T> struct FArgs {
    DevGeom g;
    Consts c;
    const double *itab, *ikd, *ktab;
    const T *ikg;
    const T *Uin, *Un;
    T *Uout, *shadow;
    const T *rhofloor;
    const double *Uinit;
    /* initial-state planes an interface boundary reads: Vinit_z [9][2][NX][NY]
     * (MC, ZVlrFixed) or Vinit_y [9][2][NX][NZ] (vanLeer, YVlrInitialize) */
    const double *Vinit_b;
    size_t NS;                 /* component stride of the state arrays */
    int NZP;                   /* padded k-row length of the state arrays */
    const T *edge;
    T *div;                    /* split stage: the divergence of each cell [9][NS] */
    int *counts;
    double *dtpart;
    size_t N;
    int kind, seglen, tiles_j, tiles_k, ntiles, nblocks;
    bool last, rk3, want_dt;
};

__device__ __forceinline__ int jmap(const DevGeom &g, int jg)
{
    if (jg == -1) return g.NY - 3;                  /* periodic wrap of the phi halo */
    if (jg == g.NY) return 2;
    return jg < 0 ? 0 : (jg > g.NY - 1 ? g.NY - 1 : jg);
}
__device__ __forceinline__ int kmap(const DevGeom &g, int kg) { return kg < 0 ? 0 : (kg > g.NZ - 1 ? g.NZ - 1 : kg); }

/* offsets of cell (i, k) (FP32 state; the FP64 build keeps none) */
template <typename T>
__device__ __forceinline__ void ub_load(const FArgs<T> &a, int i, int k, T *ub)
{
    if constexpr (sizeof(T) == 4) {
        const float4 q = *(const float4 *)(a.ikg + ((size_t)i * a.g.NZ + k) * 8 + 4);
        ub[0] = q.x; ub[1] = q.y; ub[2] = q.z; ub[3] = q.w;
    }
}

/* Element idx of component m of a state-layout array (idx < NS): the
 * component base arr + m NS is uniform and idx a 32-bit byte offset (a
 * component stays below 4 GiB), so the access takes the scalar-base form
 * with one 32-bit offset for all nine components instead of a 64-bit
 * address per component and lane. */
template <typename X>
__device__ __forceinline__ X *comp_at(X *arr, size_t mNS, unsigned byte_off)
{
    return (X *)((char *)(arr + mNS) + byte_off);
}

/* stage input of one cell (i, k) as primitives: psi damped
 * (GLMSourceSolver) then UtoV */
template <typename T>
__device__ __forceinline__ void cell_V_u(const FArgs<T> &a, T *u, int i, int k, T *v)
{
    T ub[NOFF];
    ub_load(a, i, k, ub);
    u[8] = T(a.c.psicoef) * u[8];
    Prec<T>::utov(a.c, u, ub, ro(a.itab, (size_t)i * IT + IT_OX), v);
}
template <typename T>
__device__ __forceinline__ void cell_V(const FArgs<T> &a, size_t cc, int i, int k, T *v, size_t stride)
{
    T u[9];
#pragma unroll
    for (int m = 0; m < 9; m++) u[m] = a.Uin[m * stride + cc];
    cell_V_u(a, u, i, k, v);
}
template <typename T>
__device__ __forceinline__ void cell_V(const FArgs<T> &a, size_t cc, int i, int k, T *v)
{
    T u[9];
    const unsigned off = (unsigned)(cc * sizeof(T));
#pragma unroll
    for (int m = 0; m < 9; m++) u[m] = *comp_at(a.Uin, m * a.NS, off);
    cell_V_u(a, u, i, k, v);
}
/* cell (i, j, k) of any k; ZGL: the batch layout's z-side ghosts */
template <bool ZGL, typename T>
__device__ __forceinline__ void cell_V_at(const FArgs<T> &a, int i, int j, int k, T *v)
{
    const DevGeom &g = a.g;
    if (ZGL && (k == 0 || k == g.NZ - 1)) cell_V(a, ZG(k != 0, i, j), i, k, v, 1);
    else cell_V(a, SI(i, j, k), i, k, v);
}

/* Work-group barrier for LDS traffic only.  __syncthreads() fences global
 * memory as well, so every wave would wait at each barrier for all of its
 * outstanding global loads, stores and direct-to-LDS loads (vmcnt 0).  The
 * stage kernel's threads share data only through LDS, so here a wave waits
 * only for its own LDS accesses (lgkmcnt 0: LDS is coherent within the
 * work-group once an access has completed) and the compiler is kept from
 * moving memory accesses across the barrier. */
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);          /* lgkmcnt(0); vmcnt, expcnt unconstrained */
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

/* XCD-aware block order: consecutive tiles (which share halo planes) go to
 * the same XCD's L2 (blocks are dealt round-robin over the 8 XCDs).  The
 * grid is a multiple of 8, so this is a bijection of [0, gridDim.x). */
__device__ __forceinline__ int remap_block(int b, int nb)
{
    const int per = nb / 8;
    return (b % 8) * per + b / 8;
}

/* ---- packed geometry ---------------------------------------------------- */
template <typename T>
__global__ void k_pack(DevGeom g, double *itab, T *ikg, double *ikd, double *ktab)
{
    const size_t nik = (size_t)g.NX * g.NZ;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nik; q += (size_t)gridDim.x * blockDim.x) {
        constexpr int W = ikg_w<T>();
        ikg[W * q + 0] = T(g.g1[q]); ikg[W * q + 1] = T(g.g2[q]); ikg[W * q + 2] = T(g.g3[q]); ikg[W * q + 3] = T(g.g5[q]);
        for (int s = 4; s < W; s++) ikg[W * q + s] = T(0);      /* offsets: set with the initial state */
        /* FP32 build: 1/minlength (its Setdt multiplies); FP64: minlength */
        ikd[4 * q + 0] = sizeof(T) == 4 ? 1.0 / g.minlength[q] : g.minlength[q]; ikd[4 * q + 1] = g.grxaxi[q]; ikd[4 * q + 2] = g.grzaxi[q]; ikd[4 * q + 3] = g.sp_q[q];
        if (q < (size_t)g.NX) {
            const int i = (int)q;
            double *r = itab + (size_t)i * IT;
            r[IT_X] = g.x[i]; r[IT_XM] = g.xm[i]; r[IT_OX] = g.oneox[i]; r[IT_ODX] = g.oneodx[i];
            r[IT_XABS] = i >= g.NX - 11 ? g.xabs[i - (g.NX - 11)] : 0.0;
            r[IT_X2S2] = g.sp_x2s2[i]; r[IT_O1X2S2] = g.sp_o1x2s2[i];
            r[IT_XS2O15] = g.sp_xs2o15[i]; r[IT_X2S2O15] = g.sp_x2s2o15[i];
        }
        if (q < (size_t)g.NZ) {
            const int k = (int)q;
            double *r = ktab + (size_t)k * KT;
            r[KT_ODZ] = g.oneodz[k]; r[KT_Z] = g.z[k]; r[KT_SZS] = g.sp_szs[k];
            r[KT_ONEPA] = g.sp_onepa[k]; r[KT_OZ2S2] = g.sp_oz2s2[k];
            r[KT_ZABS] = k >= g.NZ - 11 ? g.zabs[k - (g.NZ - 11)] : 0.0;
        }
    }
}

/* gravity_cell (ctx.h) from the packed rows, same expression order:
 * ikd = (Grx axisymmetric, Grz axisymmetric, q) of cell (i, k), kt = the
 * k row (szs, z, 1 + a, z^2 s^2 term), sc = (sin, cos) of the arm phase
 * at (i, j) */
template <typename TT>
__device__ __forceinline__ bool gravity_staged(const DevGeom &g, int i, cdp it, const TT *ikd, const TT *kt,
                                               const double *sc, double &gx, double &gy, double &gz, bool axi_only)
{
    if (i < g.irin || axi_only) {
        gx = ikd[1];
        gy = 0.0;
        gz = ikd[2];
        return false;
    }
    const double m1 = 10.0;
    const double epssp = g.epssp;
    const double q = ikd[3];
    const double szs = kt[0];
    const double sinphase = sc[0];
    const double cosphase = sc[1];
    const double x = it[IT_X], z = kt[1];
    const double spiralfr = (-epssp * m1 * q * szs * it[IT_XS2O15]
                             * ((x * q * q - ((2.0 - it[IT_X2S2]) * it[IT_O1X2S2])) * cosphase
                                + g.armnum * g.cotpitch * sinphase));
    const double spiralfphi = (epssp * m1 * q * szs * it[IT_XS2O15] * g.armnum * sinphase);
    const double spiralfz = (-epssp * m1 * q * z * szs * it[IT_X2S2O15]
                             * (kt[2] * q * q + kt[3]) * cosphase);
    gx = ikd[1] + spiralfr;
    gy = 0.0 + spiralfphi;
    gz = ikd[2] + spiralfz;
    return true;
}

/* The update of one interior cell (i, gj, gk) once its divergence acc is
 * formed: SpiralGravityCooling (or the source ALT selects), the RK stage
 * combination, the shadow values the ghost pass reads, the absorbing
 * layers (last stage), UtoV and Check, the store and, on the last stage,
 * the cell's Setdt candidate.  v holds the stage-input primitives of the
 * cell (overwritten), uin / unn its stage input and U^n (unn zero where the
 * stage kind reads no U^n), rfl its density floor; the table rows of the
 * cell come in as ikd (1/minlength or minlength, Grx, Grz, q of (i, k)),
 * kt (szs, z, 1 + a, z^2 s^2 term, z absorption of k), sc (sin, cos of
 * the arm phase at (i, gj)), cabs (central absorption of (i, k)) and ub
 * (its offsets, FP32 state).  cc / ccr: the cell in the state and in the
 * reference layout; it: the i row of itab. */
template <typename T, bool ALT, typename TT>
__device__ __forceinline__ void cell_update(const FArgs<T> &a, int i, int gj, int gk, size_t cc, size_t ccr, cdp it,
                                            const T *acc, T *v, const T *uin, const T *unn, T rfl, const TT *ikd,
                                            const TT *kt, const double *sc, const TT *cabs, const T *ub, int &f0,
                                            int &f1, int &f2, double &best, long long &bix)
{
    const DevGeom &g = a.g;
    const Consts &c = a.c;
    const size_t N = a.N, NS = a.NS;
    T S[9], u[9], un[9];
    double gx, gy, gz;
    const bool spiral = gravity_staged(g, i, it, ikd, kt, sc, gx, gy, gz, ALT && (c.source == 1 || c.source == 2));
    sources<T, ALT>(c, v, T(it[IT_X]), T(it[IT_OX]), T(gx), T(gy), T(gz), spiral, S);
    T ubc[NOFF];                 /* offsets of the cell */
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < NOFF; q++) ubc[q] = ub[q];
    }
#pragma unroll
    for (int m = 0; m < 9; m++) {
        /* the empty asm keeps the (zero or loaded) Un value in its own
         * precision up to here: otherwise the compiler widens it to FP64
         * at the load, in front of the update barrier, and waits for the
         * loads there */
        T w = unn[m];
        asm("" : "+v"(w));
        u[m] = uin[m];
        un[m] = w;
    }
    u[8] = T(c.psicoef) * u[8];
    un[8] = T(c.psicoef) * un[8];             /* Un is zero in stages that do not read it */
    /* The RK combination is formed in FP64 and rounded once: with FP32
     * state the stage-2/3 forms would otherwise round four or five times
     * per update (for FP64 state this is the same arithmetic).  One
     * branch-free form covers the stage kinds (the coefficients are
     * uniform), s = ((u + ka Un) - dt div) + dt S, r = kf (kb Un + ke s):
     * kinds 1, 4: s;  2: 0.25 ((3 Un + U) - dt div + dt S);  3: (1/3) (Un +
     * 2 s);  5: 0.5 ((Un + U) - dt div + dt S). */
    const int kind = a.kind;
    const double dt = c.dt;
    const double ka = kind == 2 ? 3.0 : (kind == 5 ? 1.0 : 0.0);
    const double kb = kind == 3 ? 1.0 : 0.0, ke = kind == 3 ? 2.0 : 1.0;
    const double kf = kind == 2 ? 0.25 : (kind == 3 ? 1.0 / 3.0 : (kind == 5 ? 0.5 : 1.0));
#pragma unroll
    for (int m = 0; m < 9; m++) {
        const double div = acc[m], um = u[m], unm = un[m], sm = S[m];
        const double sv = ((um + ka * unm) - dt * div) + dt * sm;
        u[m] = T(kf * (kb * unm + ke * sv));
    }
    if (!a.rk3 && gk == 1) u[3] = T(0.0);       /* ZLeftVzAntisymBSym, U3[k=1] */
    if (i == 1 || i == g.NX - 2 || gj == 1 || gj == g.NY - 2) {
#pragma unroll
        for (int m = 0; m < 9; m++) a.shadow[m * NS + cc] = u[m];
    }
    if (gk <= 2 || gk == g.NZ - 2) {
        /* z sides: packed planes k = 1, 2, NZ-2 after the shadow, the nine
         * values of a cell together */
        const size_t P = (size_t)g.NX * g.NY;
        T *zs = a.shadow + 9 * NS + 9 * ((size_t)i * g.NY + gj);
        if (gk == 1) {
#pragma unroll
            for (int m = 0; m < 9; m++) zs[m] = u[m];
        }
        if (gk == 2) {
#pragma unroll
            for (int m = 0; m < 9; m++) zs[9 * P + m] = u[m];
        }
        if (gk == g.NZ - 2) {
#pragma unroll
            for (int m = 0; m < 9; m++) zs[18 * P + m] = u[m];
        }
    }
    if (a.last) {
        const bool ax = i >= g.NX - 11, ac = i <= g.irin, az = !a.rk3 && gk >= g.NZ - 11;
        if (ax || ac || az) {
            double ui[9];                 /* every load issued before the first use */
#pragma unroll
            for (int m = 0; m < 9; m++) ui[m] = a.Uinit[m * N + ccr];
            if constexpr (sizeof(T) == 4) {
#pragma unroll
                for (int m = 0; m < 9; m++)
                    if (oslot(m) >= 0) ui[m] -= (double)ubc[oslot(m)];   /* the target as a deviation */
            }
#pragma unroll
            for (int m = 0; m < 9; m++) {
                if (ax) u[m] -= (T(it[IT_XABS]) * (u[m] - T(ui[m])));
                if (ac) u[m] -= (T(*cabs) * (u[m] - T(ui[m])));
                if (az) u[m] -= (T(kt[4]) * (u[m] - T(ui[m])));
            }
        }
    }
    Prec<T>::utov(c, u, ubc, it[IT_OX], v);
    const int fl = Prec<T>::check(c, u, ubc, v, it[IT_X], rfl);
    f0 += fl & 1; f1 += (fl >> 1) & 1; f2 += (fl >> 2) & 1;
#pragma unroll
    for (int m = 0; m < 9; m++) a.Uout[m * NS + cc] = u[m];
    if (a.want_dt && i < g.NX - 11) {
        /* Setdt (setdt.h:33-69) on the checked primitives, FP64 */
        const double V0 = v[0], V1 = v[1], V2 = v[2], V3 = v[3], V4 = v[4], V5 = v[5], V6 = v[6], V7 = v[7];
        const double v2 = V1 * V1 + V2 * V2 + V3 * V3;
        const double b2 = V4 * V4 + V5 * V5 + V6 * V6;
        double t1, tc;
        if constexpr (sizeof(T) == 4) {
            /* FP64 still, with the row's stored 1/minlength and the cooling
             * term only when a cooling function is on (otherwise it is
             * 0 / (1.5 p) = 0) */
            t1 = sqrt(v2 + (c.gm * V7 + b2) / V0) * ikd[0];
            tc = c.cooling ? (2.0 * ((V0 * V0 * cooling(c, V0, V7)) / (1.5 * V7))) : 0.0;
        } else {
            t1 = sqrt(v2 + (c.gm * V7 + b2) / V0) / ikd[0];
            tc = (2.0 * ((V0 * V0 * cooling(c, V0, V7)) / (1.5 * V7)));
        }
        const double tt = rmax(t1, tc);
        const long long ix = (long long)ccr;
        if ((tt > best || (tt == best && ix < bix)) && tt > 0.0) { best = tt; bix = ix; }
    }
}

/* Tile shapes: a block owns TJ x TK (phi x z) cells and has (TJ+2) x TK
 * threads, one per phi-state task (the TJ rows plus one halo row each
 * side): FP32 14 x 16 / 256 threads, FP64 6 x 16 / 128 threads.  The LDS
 * of one block (3-plane ring + phi and z state buffers) is 75 KB (FP32) or
 * 77 KB (FP64), so two blocks share a CU. */

template <typename T, int SHAPE> struct Tile;
template <> struct Tile<float, 0>  { static constexpr int TJ = 14, TK = 16, WAVES = 2; };
template <> struct Tile<double, 0> { static constexpr int TJ = 6, TK = 16, WAVES = 1; };
/* the batch autotuner's other shapes: wider in z (1); a two-wave tile (2:
 * 40 KB of LDS, four blocks per CU) and a three-wave tile (3: 60 KB, two
 * blocks per CU) */
template <> struct Tile<float, 1>  { static constexpr int TJ = 6, TK = 32, WAVES = 2; };
template <> struct Tile<double, 1> { static constexpr int TJ = 6, TK = 16, WAVES = 1; };
template <> struct Tile<float, 2>  { static constexpr int TJ = 6, TK = 16, WAVES = 2; };
template <> struct Tile<float, 3>  { static constexpr int TJ = 10, TK = 16, WAVES = 2; };
template <typename T, int SHAPE> constexpr int tile_threads() { return (Tile<T, SHAPE>::TJ + 2) * Tile<T, SHAPE>::TK; }
template <typename T, int SHAPE> constexpr int tile_waves() { return Tile<T, SHAPE>::WAVES; }

/* One RK stage on the work item (tile, segment): the r-march over planes
 * ibeg..iend of a TJ x TK (phi x z) column.  slot indexes the Setdt
 * partial the item writes on the last stage.  ALT: any source function
 * (Consts::source); false compiles SpiralGravityCooling only. */
template <typename T, int LIM, bool HLL, int SHAPE, bool ALT, bool BATCH = false>
__device__ __forceinline__ void fused_tile(const FArgs<T> &a, int tile, int seg, int slot)
{
    constexpr int TJ = Tile<T, SHAPE>::TJ, TK = Tile<T, SHAPE>::TK;
    constexpr int NTH = (TJ + 2) * TK;           /* threads = phi-state tasks */
    constexpr int NZT = TJ * (TK + 2);           /* z-state tasks */
    constexpr int NCELL = TJ * TK, NHALO = 4 * TK + 4 * TJ;
    static_assert(NZT <= NTH && TJ * (TK + 1) <= NTH && NHALO <= NTH, "tile does not fit the block");
    constexpr bool MC = LIM == 1;
    /* padded plane: (TJ+4) x (TK+4) cells (corners unused), 9 components,
     * 3 ring slots; every stencil read is a per-thread base + an immediate */
    constexpr int PK = TK + 4, PS = (TJ + 4) * PK, SLOT = 9 * PS;
    const DevGeom &g = a.g;
    const Consts &c = a.c;
    __shared__ T ring[3 * SLOT];
    /* limiter states of plane i: phi direction, task q = (row + 1) * TK + col
     * for rows -1..TJ; z direction, task q = row * (TK + 2) + col + 1 for
     * cols -1..TK.  [0] upper (+face) state, [1] lower (-face) state; the
     * face pass overwrites a cell's lower state with the flux through its
     * -face (only that face reads it). */
    __shared__ __align__(16) T sy[2][9 * NTH];
    __shared__ T sz[2][9 * NZT];
    /* table rows of the update, staged into LDS with the plane loads so
     * that the update phase waits on no global load: per plane (slot i % 2)
     * the (i, k) rows ikg, ikd, cabs and the (i, j) arm phase sin/cos; per
     * work item the k rows (szs, z, 1 + a, z^2 s^2 term, z absorption) and
     * the z-state 1/dz pair */
    __shared__ T s_ikg[2][TK][ikg_w<T>()];
    /* FP64 rows; in FP32 for the two-wave FP32 tile, whose LDS then fits
     * four blocks to a CU (gravity, 1/minlength and the absorbing factors of
     * the FP32 path, rounded once) */
    using TT = typename std::conditional<sizeof(T) == 4 && NTH <= 128, float, double>::type;
    __shared__ TT s_ikd[2][TK][4];
    __shared__ TT s_cabs[2][TK];
    __shared__ double s_sc[2][TJ][2];
    __shared__ TT s_kt[TK][5];
    __shared__ T s_odz[TK + 2][2];
    /* MC z boundary (ZVlrFixedBoundary): Vinitial at k = 0 (side 0) and
     * k = NZ-1 (side 1) for the tile's phi rows of plane i; staged at the
     * top of the plane in tiles that touch a z boundary, read by the z-face
     * pass before the update barrier, so one buffer serves */
    __shared__ T s_vz[MC ? 2 : 1][MC ? TJ : 1][9];
    static_assert(TK + TJ <= NTH - NCELL, "no spare threads for the table rows");
    static_assert(2 * TJ * 9 <= NTH, "no threads for the z-boundary rows");

    const int t = threadIdx.x;
    const int tj = tile / a.tiles_k, tk = tile % a.tiles_k;
    const int j0 = 1 + tj * TJ, k0 = 1 + tk * TK;
    const int kend = min(k0 + TK - 1, g.NZ - 2);      /* last interior k of the tile */
    const bool zedge = k0 == 1 || kend == g.NZ - 2;   /* tile has a z-boundary face */
    const int ibeg = 1 + seg * a.seglen;
    const int iend = min(g.NX - 2, ibeg + a.seglen - 1);
    if (ibeg > iend) {                            /* empty segment: neutral Setdt partial */
        if (a.want_dt && t == 0) { a.dtpart[2 * slot] = 0.0; a.dtpart[2 * slot + 1] = 9.3e18; }
        return;
    }
    lds_barrier();                                /* LDS of a previous work item is free */
    if (t < TK) {
        const cdp kr = (cdp)a.ktab + (size_t)min(k0 + t, g.NZ - 1) * KT;
        s_kt[t][0] = kr[KT_SZS]; s_kt[t][1] = kr[KT_Z]; s_kt[t][2] = kr[KT_ONEPA];
        s_kt[t][3] = kr[KT_OZ2S2]; s_kt[t][4] = kr[KT_ZABS];
    } else if (t < 2 * TK + 2) {
        const int zg = kmap(g, k0 + (t - TK) - 1);
        s_odz[t - TK][0] = T(ro(a.ktab, (size_t)zg * KT + KT_ODZ));
        s_odz[t - TK][1] = T(ro(a.ktab, (size_t)(zg > 0 ? zg - 1 : 0) * KT + KT_ODZ));
    }
    const bool is_cell = t < NCELL;
    const int cj = t / TK, ck = t % TK;                 /* cell coordinates in the tile */
    const int gj = j0 + cj, gk = k0 + ck;
    const bool valid = is_cell && gj <= g.NY - 2 && gk <= g.NZ - 2;
    const size_t N = a.N, NS = a.NS;
    const int cbase = (cj + 2) * PK + ck + 2;           /* own cell in a padded plane */
    /* phi-state task: row yr (-1..TJ), col yc */
    const int yr = t / TK - 1, yc = t % TK;
    const int ybase = (yr + 1) * PK + yc + 2;           /* rows yr-1..yr+1 of col yc */
    /* z-state task: row zr, col zc (-1..TK) */
    const bool zstate = t < NZT;
    const int zr = t / (TK + 2), zc = t % (TK + 2) - 1;
    const int zbase = (zr + 2) * PK + zc + 1;           /* cols zc-1..zc+1 of row zr */
    const int zcg = kmap(g, k0 + zc);
    /* faces: phi face (fyj, fyk) between rows fyj-1 and fyj; z face
     * (fzj, fzk) between cols fzk-1 and fzk, global face index kf */
    const bool yface = t < (TJ + 1) * TK;
    const int fyj = t / TK, fyk = t % TK;
    const bool zface = t < TJ * (TK + 1);
    const int fzj = t / (TK + 1), fzk = t % (TK + 1);
    const int kf = k0 + fzk - 1;

    auto ld9 = [&](int off, T *v) {
#pragma unroll
        for (int m = 0; m < 9; m++) v[m] = ring[off + m * PS];
    };
    auto load_center = [&](int p) {
        if (is_cell && p >= 0 && p <= g.NX - 1) {
            T v[9];
            if constexpr (BATCH) cell_V_at<true>(a, p, jmap(g, gj), kmap(g, gk), v);
            else cell_V(a, SI(p, jmap(g, gj), kmap(g, gk)), p, kmap(g, gk), v);
            const int off = (p % 3) * SLOT + cbase;
#pragma unroll
            for (int m = 0; m < 9; m++) ring[off + m * PS] = v[m];
        }
    };
    int hj = 0, hk = 0;
    if (t < 2 * TK) { hj = -2 + t / TK; hk = t % TK; }
    else if (t < 4 * TK) { hj = TJ + (t - 2 * TK) / TK; hk = (t - 2 * TK) % TK; }
    else if (t < 4 * TK + 2 * TJ) { const int u = t - 4 * TK; hj = u / 2; hk = -2 + u % 2; }
    else { const int u = t - 4 * TK - 2 * TJ; hj = u / 2; hk = TK + u % 2; }
    const size_t hcol = (size_t)jmap(g, j0 + hj) * a.NZP + kmap(g, k0 + hk) + KOFF;
    const int hoff = (hj + 2) * PK + hk + 2;
    auto load_halo = [&](int p) {
        if (t < NHALO) {
            T v[9];
            if constexpr (BATCH) cell_V_at<true>(a, p, jmap(g, j0 + hj), kmap(g, k0 + hk), v);
            else cell_V(a, (size_t)p * g.NY * a.NZP + hcol, p, kmap(g, k0 + hk), v);
            const int off = (p % 3) * SLOT + hoff;
#pragma unroll
            for (int m = 0; m < 9; m++) ring[off + m * PS] = v[m];
        }
    };

    /* table rows of plane p, by the threads without a cell */
    auto stage_rows = [&](int p) {
        const int sl = p & 1, u = t - NCELL;
        if (u >= 0 && u < TK) {
            const size_t ik = (size_t)p * g.NZ + min(k0 + u, g.NZ - 1);
#pragma unroll
            for (int q = 0; q < 4; q++) s_ikd[sl][u][q] = ro(a.ikd, 4 * ik + q);
#pragma unroll
            for (int q = 0; q < ikg_w<T>(); q++) s_ikg[sl][u][q] = ro(a.ikg, ikg_w<T>() * ik + q);
            if (p <= g.irin) s_cabs[sl][u] = ro(g.cabs, ik);
        } else if (u >= TK && u < TK + TJ) {
            const int j = min(j0 + u - TK, g.NY - 1);
            s_sc[sl][u - TK][0] = ro(g.sincos, (size_t)p * g.NY + j);
            s_sc[sl][u - TK][1] = ro(g.sincos, (size_t)g.NX * g.NY + (size_t)p * g.NY + j);
        }
    };

    T up_prev[9], fx_prev[9], fxn[9];
    int f0 = 0, f1 = 0, f2 = 0;
    double best = 0.0;
    long long bix = LLONG_MAX;
    const size_t NE = (size_t)g.NY * g.NZ;

    /* ---- prologue: upper state of cell ibeg and the flux at ibeg - 1/2 --- */
    if (ibeg == 1) {
        load_center(0); load_center(1); load_center(2);
        __syncthreads();
        if (is_cell) {
            T vm[9], v0[9], vp[9];
            ld9(0 * SLOT + cbase, vm); ld9(1 * SLOT + cbase, v0); ld9(2 * SLOT + cbase, vp);
            cell_states<0, LIM>(c, vm, v0, vp, T(ro(a.itab, 1 * IT + IT_ODX)), T(ro(a.itab, 0 * IT + IT_ODX)), T(1.0), up_prev, (T *)nullptr);
            const size_t e = (size_t)jmap(g, gj) * g.NZ + kmap(g, gk);
#pragma unroll
            for (int m = 0; m < 9; m++) fx_prev[m] = a.edge[m * NE + e];   /* face 0 */
        }
    } else {
        T upm[9];
        load_center(ibeg - 2); load_center(ibeg - 1); load_center(ibeg);
        __syncthreads();
        if (is_cell) {
            T vm[9], v0[9], vp[9];
            ld9(((ibeg - 2) % 3) * SLOT + cbase, vm); ld9(((ibeg - 1) % 3) * SLOT + cbase, v0);
            ld9((ibeg % 3) * SLOT + cbase, vp);
            cell_states<0, LIM>(c, vm, v0, vp, T(ro(a.itab, (ibeg - 1) * IT + IT_ODX)), T(ro(a.itab, (ibeg - 2) * IT + IT_ODX)), T(1.0), upm, (T *)nullptr);
        }
        __syncthreads();
        load_center(ibeg + 1);
        __syncthreads();
        if (is_cell) {
            T vm[9], v0[9], vp[9], lo[9];
            ld9(((ibeg - 1) % 3) * SLOT + cbase, vm); ld9((ibeg % 3) * SLOT + cbase, v0);
            ld9(((ibeg + 1) % 3) * SLOT + cbase, vp);
            cell_states<0, LIM>(c, vm, v0, vp, T(ro(a.itab, ibeg * IT + IT_ODX)), T(ro(a.itab, (ibeg - 1) * IT + IT_ODX)), T(1.0), up_prev, lo);
            riemann<0, HLL>(c, upm, lo, T(ro(a.itab, (ibeg - 1) * IT + IT_XM)), T(ro(a.itab, (ibeg - 1) * IT + IT_OX)), fx_prev);
        }
    }

    /* No barrier is needed before the loads of plane i: the center of
     * plane i+2 goes into the slot of plane i-1, whose cells other threads
     * last read in the state pass of iteration i-1 (before its face
     * barrier), and the halo of plane i goes into halo positions that only
     * the state pass of iteration i reads. */
    for (int i = ibeg; i <= iend; i++) {
        /* the centre cells of plane i+2 go into the ring slot of plane i-1.
         * BATCH: their loads are issued here and converted and written
         * after the state pass (a thread's own ring column, read next by its
         * own r-states after the face barrier and by the others two planes
         * on), so that the latency overlaps the halo, the first barrier and
         * the state pass; this pays where the extra live registers do not
         * spill (the batch kernel), and not in the single-disc kernel */
        const bool pre = i + 2 <= g.NX - 1 && i < g.NX - 2;
        T raw[9], rub[NOFF];
        if constexpr (BATCH) {
            if (pre && is_cell) {
                const int kk = kmap(g, gk);
                ub_load(a, i + 2, kk, rub);
                const bool zg = kk == g.NZ - 1;        /* past the end of a partial tile */
                const size_t cp = zg ? ZG(1, i + 2, jmap(g, gj)) : SI(i + 2, jmap(g, gj), kk);
                const size_t st = zg ? (size_t)1 : NS;
#pragma unroll
                for (int m = 0; m < 9; m++) raw[m] = a.Uin[m * st + cp];
            }
        } else {
            if (pre) load_center(i + 2);
        }
        load_halo(i);
        stage_rows(i);
        if (MC && zedge) {
            const int side = t / (TJ * 9), q = t % (TJ * 9), rj = q / 9, m = q % 9;
            if (side < 2 && ((side == 0 && k0 == 1) || (side == 1 && kend == g.NZ - 2))) {
                const size_t NP = (size_t)g.NX * g.NY;
                s_vz[side][rj][m] = T(ro(a.Vinit_b, (size_t)m * 2 * NP + side * NP + (size_t)i * g.NY + jmap(g, j0 + rj)));
            }
        }
        lds_barrier();
        const int s0 = (i % 3) * SLOT, s1 = ((i + 1) % 3) * SLOT, s2 = ((i + 2) % 3) * SLOT;
        const cdp it = (cdp)a.itab + (size_t)i * IT;
        /* ---- limiter states, once per cell and direction ---------------- */
        {
            T vm[9], v0[9], vp[9], up[9], lo[9];
            const T f1 = T(it[IT_OX]), f2 = T(g.oneody);
            ld9(s0 + ybase, vm); ld9(s0 + ybase + PK, v0); ld9(s0 + ybase + 2 * PK, vp);
            cell_states<1, LIM>(c, vm, v0, vp, f1, f1, f2, up, lo);
#pragma unroll
            for (int m = 0; m < 9; m++) { sy[0][m * NTH + t] = up[m]; sy[1][m * NTH + t] = lo[m]; }
        }
        if (zstate) {
            T vm[9], v0[9], vp[9], up[9], lo[9];
            ld9(s0 + zbase, vm); ld9(s0 + zbase + 1, v0); ld9(s0 + zbase + 2, vp);
            cell_states<2, LIM>(c, vm, v0, vp, s_odz[zc + 1][0], s_odz[zc + 1][1], T(1.0), up, lo);
#pragma unroll
            for (int m = 0; m < 9; m++) { sz[0][m * NZT + t] = up[m]; sz[1][m * NZT + t] = lo[m]; }
        }
        if constexpr (BATCH) {
            if (pre && is_cell) {
                T v[9];
                raw[8] = T(c.psicoef) * raw[8];
                Prec<T>::utov(c, raw, rub, ro(a.itab, (size_t)(i + 2) * IT + IT_OX), v);
                const int off = ((i + 2) % 3) * SLOT + cbase;
#pragma unroll
                for (int m = 0; m < 9; m++) ring[off + m * PS] = v[m];
            }
        }
        lds_barrier();
        /* ---- phi faces: (TJ+1) x TK ------------------------------------- */
        if (yface) {
            T vl[9], vr[9], fy[9];
            const int qu = fyj * TK + fyk, ql = (fyj + 1) * TK + fyk;
#pragma unroll
            for (int m = 0; m < 9; m++) { vl[m] = sy[0][m * NTH + qu]; vr[m] = sy[1][m * NTH + ql]; }
            if (LIM == 2) {
                /* YVlrInitializeBoundary: Vinitial at j = 0 is the left state
                 * of face 0, at j = NY-1 the right state of face NY-2 */
                const int jf = j0 + fyj - 1;
                const size_t NQ = (size_t)g.NX * g.NZ, ik = (size_t)i * g.NZ + min(k0 + fyk, g.NZ - 1);
                if (jf == 0) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vl[m] = T(ro(a.Vinit_b, (size_t)m * 2 * NQ + ik));
                } else if (jf == g.NY - 2) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vr[m] = T(ro(a.Vinit_b, (size_t)m * 2 * NQ + NQ + ik));
                }
            }
            riemann<1, HLL>(c, vl, vr, T(it[IT_X]), T(it[IT_OX]), fy);
#pragma unroll
            for (int m = 0; m < 9; m++) sy[1][m * NTH + ql] = fy[m];
        }
        /* ---- z faces: TJ x (TK+1) --------------------------------------- */
        if (zface && kf <= g.NZ - 2) {
            T vl[9], vr[9], fz[9];
            const int qu = fzj * (TK + 2) + fzk, ql = qu + 1;
            if (kf == 0) {
                if (MC) {
                    /* ZVlrFixedBoundary: Vinitial at k = 0 */
#pragma unroll
                    for (int m = 0; m < 9; m++) vl[m] = s_vz[0][fzj][m];
                } else if (LIM == 2) {
                    /* ZVlrFreeBoundary: Vl = Vr, the lower state of cell 1 */
#pragma unroll
                    for (int m = 0; m < 9; m++) vl[m] = sz[1][m * NZT + ql];
                } else if (LIM == 3) {
                    /* ZVlrPeriodicBoundary: the upper state of cell NZ-2 */
                    T vm[9], v0[9], vp[9];
                    const int jj = jmap(g, j0 + fzj), L = g.NZ - 2;
                    cell_V(a, SI(i, jj, L - 1), i, L - 1, vm); cell_V(a, SI(i, jj, L), i, L, v0);
                    if constexpr (BATCH) cell_V_at<true>(a, i, jj, L + 1, vp);
                    else cell_V(a, SI(i, jj, L + 1), i, L + 1, vp);
                    cell_states<2, LIM>(c, vm, v0, vp, T(1.0), T(1.0), T(1.0), vl, (T *)nullptr);
                } else {
                    /* ZLeftVlVzAntisymBzsym: lower state of cell 2, vz negated */
                    T vm[9], v0[9], vp[9];
                    const int zb = s0 + (fzj + 2) * PK + 2;       /* cols 0..2 of the tile = k 1..3 */
                    ld9(zb, vm); ld9(zb + 1, v0); ld9(zb + 2, vp);
                    cell_states<2, LIM>(c, vm, v0, vp, T(ro(a.ktab, 2 * KT + KT_ODZ)), T(ro(a.ktab, 1 * KT + KT_ODZ)), T(1.0), (T *)nullptr, vl);
                    vl[3] = -vl[3];
                }
            } else {
#pragma unroll
                for (int m = 0; m < 9; m++) vl[m] = sz[0][m * NZT + qu];
            }
            if (kf == g.NZ - 2) {
                if (MC) {
                    /* ZVlrFixedBoundary: Vinitial at k = NZ-1 */
#pragma unroll
                    for (int m = 0; m < 9; m++) vr[m] = s_vz[1][fzj][m];
                } else if (LIM == 3) {
                    /* ZVlrPeriodicBoundary: the lower state of cell 1 */
                    T vm[9], v0[9], vp[9];
                    const int jj = jmap(g, j0 + fzj);
                    if constexpr (BATCH) cell_V_at<true>(a, i, jj, 0, vm);
                    else cell_V(a, SI(i, jj, 0), i, 0, vm);
                    cell_V(a, SI(i, jj, 1), i, 1, v0); cell_V(a, SI(i, jj, 2), i, 2, vp);
                    cell_states<2, LIM>(c, vm, v0, vp, T(1.0), T(1.0), T(1.0), (T *)nullptr, vr);
                } else {
                    /* ZRightVrFreeBoundary (minmod) / ZVlrFreeBoundary (vanLeer): Vr = Vl */
#pragma unroll
                    for (int m = 0; m < 9; m++) vr[m] = vl[m];
                }
            } else {
#pragma unroll
                for (int m = 0; m < 9; m++) vr[m] = sz[1][m * NZT + ql];
            }
            riemann<2, HLL>(c, vl, vr, T(it[IT_X]), T(it[IT_OX]), fz);
#pragma unroll
            for (int m = 0; m < 9; m++) sz[1][m * NZT + ql] = fz[m];
        }
        /* ---- x: flux at i+1/2 (computed after the phi and z faces, so that
         * it is not held in registers through those passes) */
        if (is_cell) {
            if (i == g.NX - 2) {
                const size_t e = (size_t)jmap(g, gj) * g.NZ + kmap(g, gk);
#pragma unroll
                for (int m = 0; m < 9; m++) fxn[m] = a.edge[9 * NE + m * NE + e];   /* face NX-2 */
            } else {
                T vm[9], v0[9], vp[9], lo[9], up[9];
                ld9(s0 + cbase, vm); ld9(s1 + cbase, v0); ld9(s2 + cbase, vp);
                cell_states<0, LIM>(c, vm, v0, vp, T(ro(a.itab, (i + 1) * IT + IT_ODX)), T(it[IT_ODX]), T(1.0), up, lo);
                riemann<0, HLL>(c, up_prev, lo, T(it[IT_XM]), T(it[IT_OX]), fxn);
#pragma unroll
                for (int m = 0; m < 9; m++) up_prev[m] = up[m];
            }
        }
        /* inputs of the update, fetched before the barrier so that their
         * latency overlaps it (set on every lane: dead again after use) */
        const size_t cc = SI(i, gj, gk), ccr = C3(i, gj, gk);
        T uin[9], unn[9], rfl = T(0);
#pragma unroll
        for (int m = 0; m < 9; m++) { uin[m] = T(0); unn[m] = T(0); }
        if (valid) {
#pragma unroll
            for (int m = 0; m < 9; m++) uin[m] = a.Uin[m * NS + cc];
            rfl = a.rhofloor[cc];
            if (a.kind == 2 || a.kind == 3 || a.kind == 5) {
#pragma unroll
                for (int m = 0; m < 9; m++) unn[m] = a.Un[m * NS + cc];
            }
        }
        lds_barrier();
        /* ---- divergence, sources, stage update, absorb, Check, store ----- */
        if (valid) {
            const int sl = i & 1;
            T acc[9];
            {
                const T g1 = s_ikg[sl][ck][0], g2 = s_ikg[sl][ck][1];
#pragma unroll
                for (int m = 0; m < 9; m++) { acc[m] = g2 * fxn[m] - g1 * fx_prev[m]; fx_prev[m] = fxn[m]; }
            }
            {
                const T g3 = s_ikg[sl][ck][2], g4 = g3;
                const int qm = (cj + 1) * TK + ck, qp = qm + TK;
#pragma unroll
                for (int m = 0; m < 9; m++) acc[m] = acc[m] + (g4 * sy[1][m * NTH + qp] - g3 * sy[1][m * NTH + qm]);
            }
            {
                const T g5 = s_ikg[sl][ck][3], g6 = g5;
                const int qm = cj * (TK + 2) + ck + 1, qp = qm + 1;
#pragma unroll
                for (int m = 0; m < 9; m++) acc[m] = acc[m] + (g6 * sz[1][m * NZT + qp] - g5 * sz[1][m * NZT + qm]);
            }
            T v[9];
            ld9(s0 + cbase, v);
            cell_update<T, ALT>(a, i, gj, gk, cc, ccr, it, acc, v, uin, unn, rfl, s_ikd[sl][ck], s_kt[ck], s_sc[sl][cj],
                                &s_cabs[sl][ck], &s_ikg[sl][ck][4], f0, f1, f2, best, bix);
        } else if (is_cell) {
#pragma unroll
            for (int m = 0; m < 9; m++) fx_prev[m] = fxn[m];
        }
    }
    /* ---- floor counters: one atomic per wave ---------------------------- */
    for (int off = 32; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off); f1 += __shfl_xor(f1, off); f2 += __shfl_xor(f2, off);
    }
    if ((t & 63) == 0) {
        if (f0) atomicAdd(&a.counts[0], f0);
        if (f1) atomicAdd(&a.counts[1], f1);
        if (f2) atomicAdd(&a.counts[2], f2);
    }
    if (a.want_dt) {
        /* Setdt partial of the work item: lane shuffles, then one LDS slot
         * per wavefront (sy is free: the last update has read it) */
        static_assert(NTH % 64 == 0, "whole wavefronts");
        __syncthreads();
        double *sv = (double *)&sy[0][0];
        long long *sx = (long long *)(sv + NTH / 64);
        block_dt_max<NTH / 64>(best, bix, sv, sx);
        if (t == 0) { a.dtpart[2 * slot] = best; a.dtpart[2 * slot + 1] = (double)bix; }
    }
}

/* ---- the stage in two kernels ----------------------------------------------
 * march_tile: the r-march of fused_tile without the update: at plane i the
 * phi- and z-states and faces, the r-states of cell i+1 and the r-face
 * flux at i+1/2, and the divergence of each cell of plane i, stored to
 * a.div; update_cell then does the rest of the stage per cell, streaming
 * (cell_update, the same arithmetic as fused_tile).  Without the update's
 * registers and tables, and with one padded plane in LDS instead of three
 * (each thread keeps the primitives of its own cell at planes i+1 and i+2
 * in registers: only the r-states read them, and the cell at i+1 is
 * written into the plane buffer once plane i is done), the march needs
 * 51 KB of LDS and few enough registers for three 256-thread blocks per CU.
 * FP64 state (6 x 16 tiles of 128 threads, two waves per SIMD at 256
 * VGPRs, so four blocks per CU) needs its LDS below 40 KB: there the phi
 * states, then the phi fluxes and the z states share one buffer -- the phi
 * faces load their two states into registers, a barrier, then the phi
 * faces write their fluxes and the z states are written over the phi
 * states (one barrier more per plane).
 * Results are those of fused_tile bit for bit.  ZGL: the batch layout,
 * whose z-side ghost cells sit in packed planes after the state body. */
template <typename T, int LIM, bool HLL, int SHAPE, bool ZGL = false>
__device__ __forceinline__ void march_tile(const FArgs<T> &a, int tile, int seg)
{
    constexpr int TJ = Tile<T, SHAPE>::TJ, TK = Tile<T, SHAPE>::TK;
    constexpr int NTH = (TJ + 2) * TK;
    constexpr int NZT = TJ * (TK + 2);
    constexpr int NCELL = TJ * TK, NHALO = 4 * TK + 4 * TJ;
    static_assert(NZT <= NTH && TJ * (TK + 1) <= NTH && NHALO <= NTH, "tile does not fit the block");
    static_assert(TK <= NTH - NCELL && TK + 2 <= NTH, "no spare threads for the table rows");
    constexpr bool MC = LIM == 1;
    /* FP32: component strides of the plane and of the z states rounded up
     * to 64 words, so that one ds_read2st64 reads two components of a cell
     * (FP64 keeps them tight: its LDS is at the four-blocks-per-CU limit) */
    constexpr int PK = TK + 4, PSR = (TJ + 4) * PK;
    constexpr int PS = sizeof(T) == 4 ? (PSR + 63) / 64 * 64 : PSR, SLOT = 9 * PS;
    constexpr int NZS = sizeof(T) == 4 ? (NZT + 63) / 64 * 64 : NZT;
    const DevGeom &g = a.g;
    const Consts &c = a.c;
    /* plane i, padded as a ring slot of fused_tile */
    __shared__ T plane[SLOT];
    /* state buffer: phi states [2][9][NTH]; the phi fluxes [9][FYN] (FP64
     * state: below the z states, in the space of the phi states; FP32:
     * over the lower phi states) and the z states [2][9][NZT] */
    constexpr bool SHARE = sizeof(T) == 8;
    constexpr int FYN = (TJ + 1) * TK;
    constexpr int SBUF = SHARE ? (2 * 9 * NTH > 9 * FYN + 2 * 9 * NZS ? 2 * 9 * NTH : 9 * FYN + 2 * 9 * NZS)
                               : 2 * 9 * NTH + 2 * 9 * NZS;
    __shared__ __align__(16) T sbuf[SBUF];
    T *const sy0 = sbuf, *const sy1 = sbuf + 9 * NTH;
    T *const sz0 = SHARE ? sbuf + 9 * FYN : sbuf + 2 * 9 * NTH, *const sz1 = sz0 + 9 * NZS;
    /* phi flux of face fyj (between rows fyj - 1 and fyj) at column fyk:
     * pf[m * PFS + (fyj + FOFF) * TK + fyk] */
    T *const pf = SHARE ? sbuf : sy1;
    constexpr int PFS = SHARE ? FYN : NTH, FOFF = SHARE ? 0 : 1;
    __shared__ T s_g[2][TK][4];                  /* g1, g2, g3, g5 of plane i (slot i % 2) */
    __shared__ T s_odz[TK + 2][2];
    __shared__ T s_vz[MC ? 2 : 1][MC ? TJ : 1][9];
    static_assert(2 * TJ * 9 <= NTH, "no threads for the z-boundary rows");

    const int t = threadIdx.x;
    const int tj = tile / a.tiles_k, tk = tile % a.tiles_k;
    const int j0 = 1 + tj * TJ, k0 = 1 + tk * TK;
    const int kend = min(k0 + TK - 1, g.NZ - 2);
    const bool zedge = k0 == 1 || kend == g.NZ - 2;
    const int ibeg = 1 + seg * a.seglen;
    const int iend = min(g.NX - 2, ibeg + a.seglen - 1);
    if (ibeg > iend) return;
    if (t < TK + 2) {
        const int zg = kmap(g, k0 + t - 1);
        s_odz[t][0] = T(ro(a.ktab, (size_t)zg * KT + KT_ODZ));
        s_odz[t][1] = T(ro(a.ktab, (size_t)(zg > 0 ? zg - 1 : 0) * KT + KT_ODZ));
    }
    const bool is_cell = t < NCELL;
    const int cj = t / TK, ck = t % TK;
    const int gj = j0 + cj, gk = k0 + ck;
    const bool valid = is_cell && gj <= g.NY - 2 && gk <= g.NZ - 2;
    const size_t NS = a.NS;
    const int cbase = (cj + 2) * PK + ck + 2;
    const int yr = t / TK - 1, yc = t % TK;
    const int ybase = (yr + 1) * PK + yc + 2;
    const bool zstate = t < NZT;
    const int zr = t / (TK + 2), zc = t % (TK + 2) - 1;
    const int zbase = (zr + 2) * PK + zc + 1;
    const bool yface = t < (TJ + 1) * TK;
    const int fyj = t / TK, fyk = t % TK;
    const bool zface = t < TJ * (TK + 1);
    const int fzj = t / (TK + 1), fzk = t % (TK + 1);
    const int kf = k0 + fzk - 1;
    const int cjm = jmap(g, gj), ckm = kmap(g, gk);       /* own column, clamped in partial tiles */

    auto ld9 = [&](int off, T *v) {
#pragma unroll
        for (int m = 0; m < 9; m++) v[m] = plane[off + m * PS];
    };
    auto st9 = [&](int off, const T *v) {
#pragma unroll
        for (int m = 0; m < 9; m++) plane[off + m * PS] = v[m];
    };
    auto col_V = [&](int p, T *v) { cell_V_at<ZGL>(a, p, cjm, ckm, v); };
    int hj = 0, hk = 0;
    if (t < 2 * TK) { hj = -2 + t / TK; hk = t % TK; }
    else if (t < 4 * TK) { hj = TJ + (t - 2 * TK) / TK; hk = (t - 2 * TK) % TK; }
    else if (t < 4 * TK + 2 * TJ) { const int u = t - 4 * TK; hj = u / 2; hk = -2 + u % 2; }
    else { const int u = t - 4 * TK - 2 * TJ; hj = u / 2; hk = TK + u % 2; }
    const size_t hcol = (size_t)jmap(g, j0 + hj) * a.NZP + kmap(g, k0 + hk) + KOFF;
    const int hoff = (hj + 2) * PK + hk + 2;

    T up_prev[9], fx_prev[9], vn[9];                 /* vn: own cell at plane i+1 */
    const size_t NE = (size_t)g.NY * g.NZ;
    /* ---- prologue: upper r-state of cell ibeg, the flux at ibeg - 1/2,
     * plane ibeg's own cell into the plane buffer ---------------------------- */
    if (is_cell) {
        if (ibeg == 1) {
            T v0[9], v1[9];
            col_V(0, v0); col_V(1, v1); col_V(2, vn);
            cell_states<0, LIM>(c, v0, v1, vn, T(ro(a.itab, 1 * IT + IT_ODX)), T(ro(a.itab, 0 * IT + IT_ODX)), T(1.0), up_prev, (T *)nullptr);
            const size_t e = (size_t)cjm * g.NZ + ckm;
#pragma unroll
            for (int m = 0; m < 9; m++) fx_prev[m] = a.edge[m * NE + e];   /* face 0 */
            st9(cbase, v1);
        } else {
            T vm2[9], vm1[9], v0[9], upm[9], lo[9];
            col_V(ibeg - 2, vm2); col_V(ibeg - 1, vm1); col_V(ibeg, v0); col_V(ibeg + 1, vn);
            cell_states<0, LIM>(c, vm2, vm1, v0, T(ro(a.itab, (ibeg - 1) * IT + IT_ODX)), T(ro(a.itab, (ibeg - 2) * IT + IT_ODX)), T(1.0), upm, (T *)nullptr);
            cell_states<0, LIM>(c, vm1, v0, vn, T(ro(a.itab, ibeg * IT + IT_ODX)), T(ro(a.itab, (ibeg - 1) * IT + IT_ODX)), T(1.0), up_prev, lo);
            riemann<0, HLL>(c, upm, lo, T(ro(a.itab, (ibeg - 1) * IT + IT_XM)), T(ro(a.itab, (ibeg - 1) * IT + IT_OX)), fx_prev);
            st9(cbase, v0);
        }
    }

    /* Order of the plane buffer: the own cell of plane i+1 is written after
     * the third barrier of iteration i (every read of plane i is before
     * it), the halo of plane i+1 at the top of iteration i+1 (halo cells
     * are read only by the state pass, before the second barrier). */
    for (int i = ibeg; i <= iend; i++) {
        /* own cell of plane i+2: loads issued here, converted after the
         * face passes */
        const bool pre = i + 2 <= g.NX - 1;
        T raw[9], rub[NOFF];
        if (pre && is_cell) {
            /* ZGL: past the end of a partial tile the own column is the
             * packed z-side ghost */
            ub_load(a, i + 2, ckm, rub);
            if (ZGL && ckm == g.NZ - 1) {
                const size_t cp = ZG(1, i + 2, cjm);
#pragma unroll
                for (int m = 0; m < 9; m++) raw[m] = a.Uin[cp + m];
            } else {
                const unsigned off = (unsigned)(SI(i + 2, cjm, ckm) * sizeof(T));
#pragma unroll
                for (int m = 0; m < 9; m++) raw[m] = *comp_at(a.Uin, m * NS, off);
            }
        }
        if (t < NHALO) {
            T v[9];
            if constexpr (ZGL) cell_V_at<true>(a, i, jmap(g, j0 + hj), kmap(g, k0 + hk), v);
            else cell_V(a, (size_t)i * g.NY * a.NZP + hcol, i, kmap(g, k0 + hk), v);
            st9(hoff, v);
        }
        {
            const int u = t - NCELL;
            if (u >= 0 && u < TK) {
                const size_t ik = (size_t)i * g.NZ + min(k0 + u, g.NZ - 1);
#pragma unroll
                for (int q = 0; q < 4; q++) s_g[i & 1][u][q] = ro(a.ikg, ikg_w<T>() * ik + q);
            }
        }
        if (MC && zedge) {
            const int side = t / (TJ * 9), q = t % (TJ * 9), rj = q / 9, m = q % 9;
            if (side < 2 && ((side == 0 && k0 == 1) || (side == 1 && kend == g.NZ - 2))) {
                const size_t NP = (size_t)g.NX * g.NY;
                s_vz[side][rj][m] = T(ro(a.Vinit_b, (size_t)m * 2 * NP + side * NP + (size_t)i * g.NY + jmap(g, j0 + rj)));
            }
        }
        lds_barrier();
        const cdp it = (cdp)a.itab + (size_t)i * IT;
        /* ---- limiter states, once per cell and direction ---------------- */
        {
            T vm[9], v0[9], vp[9], up[9], lo[9];
            const T f1 = T(it[IT_OX]), f2 = T(g.oneody);
            ld9(ybase, vm); ld9(ybase + PK, v0); ld9(ybase + 2 * PK, vp);
            cell_states<1, LIM>(c, vm, v0, vp, f1, f1, f2, up, lo);
#pragma unroll
            for (int m = 0; m < 9; m++) { sy0[m * NTH + t] = up[m]; sy1[m * NTH + t] = lo[m]; }
        }
        auto z_states = [&]() {
            if (zstate) {
                T vm[9], v0[9], vp[9], up[9], lo[9];
                ld9(zbase, vm); ld9(zbase + 1, v0); ld9(zbase + 2, vp);
                cell_states<2, LIM>(c, vm, v0, vp, s_odz[zc + 1][0], s_odz[zc + 1][1], T(1.0), up, lo);
#pragma unroll
                for (int m = 0; m < 9; m++) { sz0[m * NZS + t] = up[m]; sz1[m * NZS + t] = lo[m]; }
            }
        };
        if (!SHARE) z_states();
        lds_barrier();
        /* ---- phi faces: (TJ+1) x TK ------------------------------------- */
        {
            T vl[9], vr[9], fy[9];
            if (yface) {
                const int qu = fyj * TK + fyk, ql = (fyj + 1) * TK + fyk;
#pragma unroll
                for (int m = 0; m < 9; m++) { vl[m] = sy0[m * NTH + qu]; vr[m] = sy1[m * NTH + ql]; }
            }
            if (yface) {
            if (LIM == 2) {
                const int jf = j0 + fyj - 1;
                const size_t NQ = (size_t)g.NX * g.NZ, ik = (size_t)i * g.NZ + min(k0 + fyk, g.NZ - 1);
                if (jf == 0) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vl[m] = T(ro(a.Vinit_b, (size_t)m * 2 * NQ + ik));
                } else if (jf == g.NY - 2) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vr[m] = T(ro(a.Vinit_b, (size_t)m * 2 * NQ + NQ + ik));
                }
            }
            riemann<1, HLL>(c, vl, vr, T(it[IT_X]), T(it[IT_OX]), fy);
            }
            /* every phi state has been read: the phi fluxes and then the z
             * states may overwrite them */
            if (SHARE) lds_barrier();
            if (yface) {
#pragma unroll
            for (int m = 0; m < 9; m++) pf[m * PFS + (fyj + FOFF) * TK + fyk] = fy[m];
            }
        }
        if (SHARE) {
            z_states();
            lds_barrier();
        }
        /* ---- z faces: TJ x (TK+1) --------------------------------------- */
        if (zface && kf <= g.NZ - 2) {
            T vl[9], vr[9], fz[9];
            const int qu = fzj * (TK + 2) + fzk, ql = qu + 1;
            if (kf == 0) {
                if (MC) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vl[m] = s_vz[0][fzj][m];
                } else if (LIM == 2) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vl[m] = sz1[m * NZS + ql];
                } else if (LIM == 3) {
                    T vm[9], v0[9], vp[9];
                    const int jj = jmap(g, j0 + fzj), L = g.NZ - 2;
                    cell_V(a, SI(i, jj, L - 1), i, L - 1, vm); cell_V(a, SI(i, jj, L), i, L, v0);
                    cell_V_at<ZGL>(a, i, jj, L + 1, vp);
                    cell_states<2, LIM>(c, vm, v0, vp, T(1.0), T(1.0), T(1.0), vl, (T *)nullptr);
                } else {
                    T vm[9], v0[9], vp[9];
                    const int zb = (fzj + 2) * PK + 2;
                    ld9(zb, vm); ld9(zb + 1, v0); ld9(zb + 2, vp);
                    cell_states<2, LIM>(c, vm, v0, vp, T(ro(a.ktab, 2 * KT + KT_ODZ)), T(ro(a.ktab, 1 * KT + KT_ODZ)), T(1.0), (T *)nullptr, vl);
                    vl[3] = -vl[3];
                }
            } else {
#pragma unroll
                for (int m = 0; m < 9; m++) vl[m] = sz0[m * NZS + qu];
            }
            if (kf == g.NZ - 2) {
                if (MC) {
#pragma unroll
                    for (int m = 0; m < 9; m++) vr[m] = s_vz[1][fzj][m];
                } else if (LIM == 3) {
                    T vm[9], v0[9], vp[9];
                    const int jj = jmap(g, j0 + fzj);
                    cell_V_at<ZGL>(a, i, jj, 0, vm);
                    cell_V(a, SI(i, jj, 1), i, 1, v0); cell_V(a, SI(i, jj, 2), i, 2, vp);
                    cell_states<2, LIM>(c, vm, v0, vp, T(1.0), T(1.0), T(1.0), (T *)nullptr, vr);
                } else {
#pragma unroll
                    for (int m = 0; m < 9; m++) vr[m] = vl[m];
                }
            } else {
#pragma unroll
                for (int m = 0; m < 9; m++) vr[m] = sz1[m * NZS + ql];
            }
            riemann<2, HLL>(c, vl, vr, T(it[IT_X]), T(it[IT_OX]), fz);
#pragma unroll
            for (int m = 0; m < 9; m++) sz1[m * NZS + ql] = fz[m];
        }
        /* ---- x: r-states of cell i+1 (planes i, i+1, i+2 of the own
         * column) and the flux at i+1/2 ----------------------------------- */
        T fxn[9], vq[9];
        if (is_cell) {
            if (i == g.NX - 2) {
                const size_t e = (size_t)cjm * g.NZ + ckm;
#pragma unroll
                for (int m = 0; m < 9; m++) fxn[m] = a.edge[9 * NE + m * NE + e];   /* face NX-2 */
            } else {
                T vm[9], lo[9], up[9];
                raw[8] = T(c.psicoef) * raw[8];
                Prec<T>::utov(c, raw, rub, ro(a.itab, (size_t)(i + 2) * IT + IT_OX), vq);
                ld9(cbase, vm);
                cell_states<0, LIM>(c, vm, vn, vq, T(ro(a.itab, (i + 1) * IT + IT_ODX)), T(it[IT_ODX]), T(1.0), up, lo);
                riemann<0, HLL>(c, up_prev, lo, T(it[IT_XM]), T(it[IT_OX]), fxn);
#pragma unroll
                for (int m = 0; m < 9; m++) up_prev[m] = up[m];
            }
        }
        lds_barrier();
        /* ---- divergence, in the reference's order ----------------------- */
        if (valid) {
            const int sl = i & 1;
            T acc[9];
            {
                const T g1 = s_g[sl][ck][0], g2 = s_g[sl][ck][1];
#pragma unroll
                for (int m = 0; m < 9; m++) { acc[m] = g2 * fxn[m] - g1 * fx_prev[m]; fx_prev[m] = fxn[m]; }
            }
            {
                const T g3 = s_g[sl][ck][2], g4 = g3;
                const int qm = (cj + FOFF) * TK + ck, qp = qm + TK;
#pragma unroll
                for (int m = 0; m < 9; m++) acc[m] = acc[m] + (g4 * pf[m * PFS + qp] - g3 * pf[m * PFS + qm]);
            }
            {
                const T g5 = s_g[sl][ck][3], g6 = g5;
                const int qm = cj * (TK + 2) + ck + 1, qp = qm + 1;
#pragma unroll
                for (int m = 0; m < 9; m++) acc[m] = acc[m] + (g6 * sz1[m * NZS + qp] - g5 * sz1[m * NZS + qm]);
            }
            const unsigned off = (unsigned)(SI(i, gj, gk) * sizeof(T));
#pragma unroll
            for (int m = 0; m < 9; m++) *comp_at(a.div, m * NS, off) = acc[m];
        } else if (is_cell) {
#pragma unroll
            for (int m = 0; m < 9; m++) fx_prev[m] = fxn[m];
        }
        /* own cell of plane i+1 into the plane buffer; i+2 becomes i+1 */
        if (is_cell && i < iend) {
            st9(cbase, vn);
#pragma unroll
            for (int m = 0; m < 9; m++) vn[m] = vq[m];
        }
    }
}

/* The per-cell half of the split stage: one thread per interior cell, a
 * wavefront over 64 consecutive k of one (i, j) row (chunk `ch` of the
 * row's nch = ceil((NZ-2)/64)), four wavefronts to a block; the stage
 * input is read once (UtoV gives the primitives the sources use, as the
 * plane loads of the march do), the divergence from a.div.  blk: the
 * block's index among this disc's update blocks of the stage.  FLAT: each
 * plane's cells are dealt 64 to a wavefront in (j, k) order instead (discs
 * whose rows are short against a wavefront).  Floor counters and the Setdt
 * partial (dtpart slot `slot`) as in fused_tile. */
template <typename T, bool ALT, bool FLAT = false, int RJ = 1>
__device__ __forceinline__ void update_cell(const FArgs<T> &a, int nch, int blk, int slot)
{
    const DevGeom &g = a.g;
    const size_t NS = a.NS;
    const int lane = threadIdx.x & 63;
    const int nj = g.NY - 2;
    int i, j, k;
    bool on;
    if constexpr (FLAT) {
        /* each plane's interior cells in (j, k) order, k fastest, 64 to a
         * wave whatever NZ (short rows fill whole waves), no wave crossing
         * a plane: i is uniform in the wave, its table row a scalar load */
        const unsigned nk = (unsigned)(g.NZ - 2), P = (unsigned)nj * nk, wpp = (P + 63) / 64;
        const unsigned w = (unsigned)__builtin_amdgcn_readfirstlane((int)(((unsigned)blk * blockDim.x + threadIdx.x) >> 6));
        const unsigned ip = w / wpp;
        const unsigned q = (w - ip * wpp) * 64 + (unsigned)lane;
        const unsigned r = q / nk;
        on = ip < (unsigned)(g.NX - 2) && q < P;
        i = 1 + (int)ip;
        j = 1 + (int)r;
        k = 1 + (int)(q - r * nk);
    } else {
        /* RJ consecutive rows j of one plane per wavefront (nphi is even:
         * the axis map j <-> j + nphi/2 needs it) */
        const int njb = (nj + RJ - 1) / RJ;
        const int w = __builtin_amdgcn_readfirstlane((int)((blk * blockDim.x + threadIdx.x) >> 6));
        const int r = w / nch;
        k = 1 + (w - r * nch) * 64 + lane;
        on = r < (g.NX - 2) * njb && k <= g.NZ - 2;
        i = 1 + r / njb;
        j = 1 + (r % njb) * RJ;
    }
    int f0 = 0, f1 = 0, f2 = 0;
    double best = 0.0;
    long long bix = LLONG_MAX;
    if (on) {
        const cdp it = (cdp)a.itab + (size_t)i * IT;
        const size_t ik = (size_t)i * g.NZ + k;
        /* every load of the RJ cells issued before the first is used; the
         * (i, k) and k table rows serve all of them */
        T uin[RJ][9], unn[RJ][9], acc[RJ][9], rfl[RJ], ub[NOFF];
        bool row[RJ];
#pragma unroll
        for (int q = 0; q < RJ; q++) {
            row[q] = j + q <= g.NY - 2;
            const size_t cc = SI(i, row[q] ? j + q : j, k);
#pragma unroll
            for (int m = 0; m < 9; m++) { uin[q][m] = a.Uin[m * NS + cc]; unn[q][m] = T(0); }
            if (a.kind == 2 || a.kind == 3 || a.kind == 5) {
#pragma unroll
                for (int m = 0; m < 9; m++) unn[q][m] = a.Un[m * NS + cc];
            }
#pragma unroll
            for (int m = 0; m < 9; m++) acc[q][m] = a.div[m * NS + cc];
            rfl[q] = a.rhofloor[cc];
        }
        ub_load(a, i, k, ub);
        double ikd[4], kt[5], cabs = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) ikd[q] = a.ikd[4 * ik + q];
        const double *kr = a.ktab + (size_t)k * KT;
        kt[0] = kr[KT_SZS]; kt[1] = kr[KT_Z]; kt[2] = kr[KT_ONEPA]; kt[3] = kr[KT_OZ2S2]; kt[4] = kr[KT_ZABS];
        if (i <= g.irin) cabs = g.cabs[ik];
#pragma unroll
        for (int q = 0; q < RJ; q++) {
            if (!row[q]) continue;
            const int jq = j + q;
            const double sc[2] = {ro(g.sincos, (size_t)i * g.NY + jq), ro(g.sincos, (size_t)g.NX * g.NY + (size_t)i * g.NY + jq)};
            const size_t cc = SI(i, jq, k), ccr = C3(i, jq, k);
            T v[9];
            {
                T u[9];
#pragma unroll
                for (int m = 0; m < 9; m++) u[m] = uin[q][m];
                u[8] = T(a.c.psicoef) * u[8];
                Prec<T>::utov(a.c, u, ub, it[IT_OX], v);
            }
            cell_update<T, ALT>(a, i, jq, k, cc, ccr, it, acc[q], v, uin[q], unn[q], rfl[q], ikd, kt, sc, &cabs, ub, f0, f1,
                                f2, best, bix);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off); f1 += __shfl_xor(f1, off); f2 += __shfl_xor(f2, off);
    }
    if (lane == 0) {
        if (f0) atomicAdd(&a.counts[0], f0);
        if (f1) atomicAdd(&a.counts[1], f1);
        if (f2) atomicAdd(&a.counts[2], f2);
    }
    if (a.want_dt) {
        __shared__ double sv[4];
        __shared__ long long sx[4];
        block_dt_max<4>(best, bix, sv, sx);
        if (threadIdx.x == 0) { a.dtpart[2 * slot] = best; a.dtpart[2 * slot + 1] = (double)bix; }
    }
}

/* ---- r-boundary faces 0 and NX-2 ------------------------------------------
 * XVlrPeriodicBoundary (minmod, MC, superbee): both faces take
 * (upper(NX-2), lower(1)); XVlrFreeBoundary (vanLeer): face 0 takes
 * (lower(1), lower(1)), face NX-2 (upper(NX-2), upper(NX-2)). */
template <typename T, int LIM, bool HLL>
__device__ __forceinline__ void edge_cell(const FArgs<T> &a, T *edge, int q)
{
    const DevGeom &g = a.g;
    const int nj = g.NY - 2, nk = g.NZ - 2;
    if (q >= nj * nk) return;
    const int j = 1 + q / nk, k = 1 + q % nk;
    T vm[9], v0[9], vp[9], up[9], lo[9], f[9];
    const int L = g.NX - 2;
    cell_V(a, SI(L - 1, j, k), L - 1, k, vm); cell_V(a, SI(L, j, k), L, k, v0); cell_V(a, SI(L + 1, j, k), L + 1, k, vp);
    cell_states<0, LIM>(a.c, vm, v0, vp, T(g.oneodx[L]), T(g.oneodx[L - 1]), T(1.0), up, (T *)nullptr);
    cell_V(a, SI(0, j, k), 0, k, vm); cell_V(a, SI(1, j, k), 1, k, v0); cell_V(a, SI(2, j, k), 2, k, vp);
    cell_states<0, LIM>(a.c, vm, v0, vp, T(g.oneodx[1]), T(g.oneodx[0]), T(1.0), (T *)nullptr, lo);
    const size_t NE = (size_t)g.NY * g.NZ, e = (size_t)j * g.NZ + k;
    riemann<0, HLL>(a.c, LIM == 2 ? lo : up, lo, T(g.xm[0]), T(g.oneox[0]), f);
#pragma unroll
    for (int m = 0; m < 9; m++) edge[m * NE + e] = f[m];
    riemann<0, HLL>(a.c, up, LIM == 2 ? up : lo, T(g.xm[L]), T(g.oneox[L]), f);
#pragma unroll
    for (int m = 0; m < 9; m++) edge[9 * NE + m * NE + e] = f[m];
}

/* ---- ghost cells: boundary map of the shadow, then absorb / UtoV / Check */
template <typename T>
__device__ __forceinline__ size_t halo_cells(const DevGeom &g)
{
    const size_t NX = g.NX, NY = g.NY, NZ = g.NZ;
    return 2 * NY * NZ + (NX - 2) * 2 * NZ + (NX - 2) * (NY - 2) * 2;
}

/* ghost cell r of the list: the r-side planes i = 0, NX-1 (all j, k), then
 * the phi sides j = 0, NY-1 (interior i, all k), then the z sides k = 0,
 * NZ-1 (interior i, j); I: the integer type of the division (32-bit where
 * the count fits: a 64-bit division is a long software sequence) */
template <typename I>
__device__ __forceinline__ void halo_ijk(I r, I NX, I NY, I NZ, int &i, int &j, int &k)
{
    const I nA = 2 * NY * NZ, nB = (NX - 2) * 2 * NZ;
    if (r < nA) { i = (r / (NY * NZ)) ? (int)NX - 1 : 0; r %= NY * NZ; j = (int)(r / NZ); k = (int)(r % NZ); }
    else if ((r -= nA) < nB) { i = 1 + (int)(r / (2 * NZ)); r %= 2 * NZ; j = (r / NZ) ? (int)NY - 1 : 0; k = (int)(r % NZ); }
    else { r -= nB; i = 1 + (int)(r / ((NY - 2) * 2)); r %= (NY - 2) * 2; j = 1 + (int)(r / 2); k = (r % 2) ? (int)NZ - 1 : 0; }
}

/* ghost cells first, first + stride, ... of the composed boundary map;
 * every lane of the wave must call it (the counters are wave-reduced) */
template <typename T, bool ZGL = false>
__device__ __forceinline__ void halo_range(const FArgs<T> &a, size_t first, size_t stride)
{
    const DevGeom &g = a.g;
    const size_t NX = g.NX, NY = g.NY, NZ = g.NZ, N = a.N, NS = a.NS;
    const size_t nA = 2 * NY * NZ, nB = (NX - 2) * 2 * NZ, nC = (NX - 2) * (NY - 2) * 2;
    const size_t tot = nA + nB + nC;
    const bool narrow = tot + stride <= 0xffffffffull;
    int f0 = 0, f1 = 0, f2 = 0;
    for (size_t t = first; t < tot; t += stride) {
        int i, j, k;
        if (narrow) halo_ijk<unsigned>((unsigned)t, (unsigned)NX, (unsigned)NY, (unsigned)NZ, i, j, k);
        else halo_ijk<size_t>(t, NX, NY, NZ, i, j, k);
        int si, sj, sk;
        unsigned sign;
        bool zero3;
        ghost_source(g, a.rk3, i, j, k, si, sj, sk, sign, zero3);
        const size_t dst = SI(i, j, k), dstr = C3(i, j, k);
        /* the z-side ghosts of the interior columns read the packed planes
         * k = 1, 2, NZ-2; the others the shadow in the state layout */
        const bool zside = (k == 0 || k == (int)NZ - 1) && i >= 1 && i <= (int)NX - 2 && j >= 1 && j <= (int)NY - 2;
        const size_t P = NX * NY;
        const int zsl = sk == 1 ? 0 : (sk == 2 ? 1 : 2);
        const T *sp = zside ? a.shadow + 9 * NS + 9 * ((size_t)zsl * P + (size_t)si * NY + sj) : a.shadow + SI(si, sj, sk);
        const size_t sstride = zside ? 1 : NS;
        T u[9], v[9], ub[NOFF];
        ub_load(a, i, k, ub);
        if constexpr (sizeof(T) == 4) {
            /* the source's deviation -> its full value (offsets of (si, sk)),
             * signed, -> a deviation from the ghost's own offsets; exact
             * where the two offsets agree and no sign changes */
            T us[NOFF];
            ub_load(a, si, sk, us);
#pragma unroll
            for (int m = 0; m < 9; m++) {
                const T val = sp[m * sstride];
                if (oslot(m) >= 0) {
                    const double full = (double)val + (double)us[oslot(m)];
                    u[m] = (float)(((sign >> m) & 1u ? -full : full) - (double)ub[oslot(m)]);
                } else {
                    u[m] = (sign >> m) & 1u ? -val : val;
                }
            }
            if (zero3) u[3] = T(0.0);
        } else {
#pragma unroll
            for (int m = 0; m < 9; m++) {
                const T val = sp[m * sstride];
                u[m] = (sign >> m) & 1u ? -val : val;
            }
            if (zero3) u[3] = T(0.0);
        }
        if (a.last) {
            const bool ax = i >= g.NX - 11, ac = i <= g.irin, az = !a.rk3 && k >= g.NZ - 11;
            if (ax || ac || az) {
#pragma unroll
                for (int m = 0; m < 9; m++) {
                    double uid = a.Uinit[m * N + dstr];
                    if constexpr (sizeof(T) == 4) {
                        if (oslot(m) >= 0) uid -= (double)ub[oslot(m) >= 0 ? oslot(m) : 0];
                    }
                    const T ui = T(uid);
                    if (ax) u[m] -= (T(g.xabs[i - (g.NX - 11)]) * (u[m] - ui));
                    if (ac) u[m] -= (T(g.cabs[(size_t)i * g.NZ + k]) * (u[m] - ui));
                    if (az) u[m] -= (T(g.zabs[k - (g.NZ - 11)]) * (u[m] - ui));
                }
            }
        }
        Prec<T>::utov(a.c, u, ub, g.oneox[i], v);
        const int fl = Prec<T>::check(a.c, u, ub, v, g.x[i], a.rhofloor[dst]);
        f0 += fl & 1; f1 += (fl >> 1) & 1; f2 += (fl >> 2) & 1;
        /* A z-side ghost is the only cell of its 64-byte line (the rest is
         * row padding: KOFF cells before k = 1, NZP - NZ - KOFF after NZ-1),
         * so the whole line is written, avoiding a partial-line write. */
        constexpr int E = 64 / (int)sizeof(T);
        if (ZGL && (k == 0 || k == (int)NZ - 1)) {
            T *q = a.Uout + ZG(k != 0, i, j);        /* the batch layout's packed slot */
#pragma unroll
            for (int m = 0; m < 9; m++) q[m] = u[m];
        } else if ((k == 0 && (dst % E) == E - 1) || (k == (int)NZ - 1 && (dst % E) == 0)) {
            const size_t line = dst - dst % E;
#pragma unroll
            for (int m = 0; m < 9; m++) {
                T *q = a.Uout + m * NS + line;
#pragma unroll
                for (int e = 0; e < E; e++) q[e] = u[m];
            }
        } else {
#pragma unroll
            for (int m = 0; m < 9; m++) a.Uout[m * NS + dst] = u[m];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off); f1 += __shfl_xor(f1, off); f2 += __shfl_xor(f2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (f0) atomicAdd(&a.counts[0], f0);
        if (f1) atomicAdd(&a.counts[1], f1);
        if (f2) atomicAdd(&a.counts[2], f2);
    }
}


}  // namespace FUSED_NS
}  // namespace cyl
#endif
