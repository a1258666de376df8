/*
 * ctx.h -- device-side tables shared by the kernels (internal).
 *
 * HBM layout.  Every cell array is component-major and k-fastest,
 * A[m][i][j][k] with the ghosts included, exactly the reference layout
 * (commonfield.h:114-118), so host <-> device copies are single memcpys and
 * a wavefront of consecutive lanes walks consecutive k.  Geometry that the
 * reference stores per cell but that depends on (i,k) only -- g1..g6,
 * dvolume, minlength, Grxaxi, Grzaxi -- lives in [NX][NZ] tables.
 */
#ifndef CYLMHD_CTX_H
#define CYLMHD_CTX_H

#include <climits>

#include <hip/hip_runtime.h>

#include "../../include/cylmhd.h"
#include "device_math.h"
#include "grid.h"

namespace cyl {

/* geometry / gravity tables on the device (all FP64) */
struct DevGeom {
    int NX, NY, NZ, jy, irin;
    const double *x, *xm, *oneox, *oneodx, *oneodz, *y, *z;
    double oneody;
    const double *g1, *g2, *g3, *g5, *dvol, *minlength;         /* [NX*NZ] */
    const double *grxaxi, *grzaxi;                              /* [NX*NZ] */
    const double *sp_x2s2, *sp_o1x2s2, *sp_xs2o15, *sp_x2s2o15; /* [NX]    */
    const double *sp_szs, *sp_onepa, *sp_oz2s2;                 /* [NZ]    */
    const double *sp_q;                                         /* [NX*NZ] */
    const double *sincos;                                       /* [2][NX][NY]: sin, cos */
    double epssp, cotpitch;
    int armnum;
    const double *xabs, *zabs, *cabs;                           /* [11],[11],[(irin+1)*NZ] */
};

/* Gravity of one cell as SpiralGravityCooling uses it (source.h:36-209):
 * Grxaxi/Grzaxi for i < irin, otherwise Grx.. = axisymmetric + spiral
 * (spiralgravity.h:83-113), built from time-independent tables plus the
 * per-step sin/cos of the arm phase.  Expression order as the reference. */
__device__ __forceinline__ bool gravity_cell(const DevGeom &g, int i, int j, int k, double &gx,
                                             double &gy, double &gz, bool axi_only = false)
{
    const size_t ik = (size_t)i * g.NZ + k;
    if (i < g.irin || axi_only) {
        gx = g.grxaxi[ik];
        gy = 0.0;
        gz = g.grzaxi[ik];
        return false;
    }
    const double m1 = 10.0;
    const double epssp = g.epssp;
    const double q = g.sp_q[ik];
    const double szs = g.sp_szs[k];
    const double sinphase = g.sincos[(size_t)i * g.NY + j];
    const double cosphase = g.sincos[(size_t)g.NX * g.NY + (size_t)i * g.NY + j];
    const double x = g.x[i], z = g.z[k];
    const double spiralfr = (-epssp * m1 * q * szs * g.sp_xs2o15[i]
                             * ((x * q * q - ((2.0 - g.sp_x2s2[i]) * g.sp_o1x2s2[i])) * cosphase
                                + g.armnum * g.cotpitch * sinphase));
    const double spiralfphi = (epssp * m1 * q * szs * g.sp_xs2o15[i] * g.armnum * sinphase);
    const double spiralfz = (-epssp * m1 * q * z * szs * g.sp_x2s2o15[i]
                             * (g.sp_onepa[k] * q * q + g.sp_oz2s2[k]) * cosphase);
    gx = g.grxaxi[ik] + spiralfr;
    gy = 0.0 + spiralfphi;
    gz = g.grzaxi[ik] + spiralfz;
    return true;
}

/* Composed ghost-cell map of the cell boundary set applied after each stage
 * update.  RK3 set (firststep.h:125-157): Y periodic, X axis, X right free,
 * Z periodic; RK2 set (halfstep.h:117-148): Y periodic, X axis, X right
 * free, Z left vz-antisymmetric mirror (k=0 <- k=2, U3[k=1] = 0), Z right
 * free.  Applied in that order every ghost value is a signed copy of one
 * interior value; this returns it.  sign bit m set = negate component m;
 * zero3 = component 3 forced to 0 (RK2, k == 1). */
__device__ __forceinline__ void ghost_source(const DevGeom &g, bool rk3, int i, int j, int k, int &si,
                                             int &sj, int &sk, unsigned &sign, bool &zero3)
{
    const int NX = g.NX, NY = g.NY, NZ = g.NZ, jy = g.jy;
    sign = 0u;
    zero3 = false;
    if (rk3) {
        sk = k == 0 ? NZ - 2 : (k == NZ - 1 ? 1 : k);
    } else {
        sk = k == 0 ? 2 : (k == NZ - 1 ? NZ - 2 : k);
        if (k == 0) sign |= 1u << 3;
        if (k == 1) zero3 = true;
    }
    if (i == 0) {
        si = 1;
        sj = j < jy / 2 + 1 ? j + jy / 2 : j - jy / 2;
        sign ^= (1u << 1) | (1u << 2) | (1u << 4) | (1u << 5);
    } else {
        si = i == NX - 1 ? NX - 2 : i;
        sj = j == 0 ? NY - 2 : (j == NY - 1 ? 1 : j);
    }
}

/* ---- Setdt's maximum (setdt.h:52-69) ----------------------------------
 * The largest 1/dt and the first cell (lowest reference-layout index) that
 * attains it.  The pair order (value descending, index ascending) is a total
 * order, so the result does not depend on the reduction order. */
__device__ __forceinline__ bool dt_better(double v, long long ix, double bv, long long bix)
{
    return v > bv || (v == bv && ix < bix);
}

/* within a wavefront: xor butterfly over the 64 lanes (every lane must be
 * active), every lane ends with the wavefront's maximum */
__device__ __forceinline__ void wave_dt_max(double &v, long long &ix)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const long long oi = __shfl_xor(ix, off);
        if (dt_better(ov, oi, v, ix)) { v = ov; ix = oi; }
    }
}

/* across a block of NW wavefronts (all threads call it): wavefront maxima
 * through NW LDS slots, then wavefront 0; thread 0 returns the block's */
template <int NW>
__device__ __forceinline__ void block_dt_max(double &v, long long &ix, double *sv, long long *sx)
{
    wave_dt_max(v, ix);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sv[w] = v; sx[w] = ix; }
    __syncthreads();
    if (w == 0) {
        v = lane < NW ? sv[lane] : 0.0;
        ix = lane < NW ? sx[lane] : LLONG_MAX;
        wave_dt_max(v, ix);
    }
}

}  // namespace cyl
#endif
