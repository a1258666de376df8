/*
 * device_math.h -- pointwise device physics shared by every kernel.
 *
 * Each function restates one reference routine (file:line at the function)
 * in the reference's expression order.  The library is compiled with
 * -ffp-contract=off, so with T = double every operation rounds exactly as
 * the reference's x86-64 build does (IEEE +,-,*,/ and sqrt are correctly
 * rounded on gfx950 too); only libm calls (pow in Cooling, sincos of the
 * spiral phase) may differ in the last ulp.
 *
 * Directions: DIR 0 = r (x), 1 = phi (y), 2 = z.  Component indices follow
 * commonfield.h:10-39: 0 rho, 1..3 velocity / momentum, 4..6 B, 7 p / E,
 * 8 psi.  For the HLLD star states the tangential pair is taken in the
 * order each direction forms its dot products: x (phi, z), y (z, r),
 * z (r, phi)  (hlldx.h:286, hlldy.h:294, hlldz.h:289).
 */
#ifndef CYLMHD_DEVICE_MATH_H
#define CYLMHD_DEVICE_MATH_H

#include <climits>

#include <hip/hip_runtime.h>

namespace cyl {

/* maxmin.h:16-26 */
template <typename T> __device__ __forceinline__ T rmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T rmin(T a, T b) { return a > b ? b : a; }

struct Consts {          /* scalars of one stage, passed by value */
    double gm, gm1, oneogm1;
    double dt, ch, psicoef;
    double prcr, betacr, oneobetacr, rhocr;
    double omegasp;
    double te0;
    int    cooling;
    int    source;            /* CYLMHD_SOURCE_* (in cooling's alignment gap) */
    double cool_lambda0, cool_tmin, cool_index;
};

/* A of the Raymond fit, 0.108e-37 (te0 T)^10.26 (cooling.h:50).  Kept out
 * of line: it runs only inside the cooling band, and inlining the FP64 pow
 * would reserve its registers across the whole stage update. */
__device__ __noinline__ double raymond_A(double te0, double temperature)
{
    return 0.108e-37 * pow(te0 * temperature, 10.26);
}

/* Power-law cooling factor T^index (index != 1/2), out of line for the same
 * reason: an inlined FP64 pow keeps its polynomial constants in registers
 * across the whole r-march of the stage kernel. */
__device__ __noinline__ double powerlaw_pow(double temperature, double index)
{
    return pow(temperature, index);
}

/* Cooling() is declared int (cooling.h:22), so its double result goes
 * through a conversion to int.  C leaves that undefined for NaN (the
 * temperature exactly at a break point makes sign = 0/0) and out of range;
 * the compiled reference on x86-64 (cvttsd2si) gives INT_MIN there, and so
 * does this, where the device's own conversion would give 0 for NaN. */
template <typename T>
__device__ __forceinline__ int ref_int(T v)
{
    return (v != v || !(fabs(v) < T(2147483648.0))) ? INT_MIN : (int)v;
}

/* ---- Cooling (cooling.h:22-60) --------------------------------------- */
template <typename T>
__device__ __forceinline__ T cooling(const Consts &c, T rho, T pr)
{
    if (c.cooling == 1) {
        const T temperature = pr / rho;
        const T a = T(0.031), b = T(0.054), cc = T(0.078), d = T(0.235);
        const T tempma = temperature - a, fabstempma = fabs(tempma);
        const T tempmb = temperature - b, fabstempmb = fabs(tempmb);
        const T tempmc = temperature - cc, fabstempmc = fabs(tempmc);
        const T tempmd = temperature - d, fabstempmd = fabs(tempmd);
        const T signa = tempma / fabstempma, signb = tempmb / fabstempmb;
        const T signc = tempmc / fabstempmc, signd = tempmd / fabstempmd;
        const T fa = T(1.0) - signa * signb;
        const T B = T(0.0), C = T(0.0);
        /* outside the band a < T < b the factor of A is exactly 0 (and 0.5*
         * (A*0 + B*.. + C*..) truncates to 0), so pow is only evaluated
         * inside it; NaN factors (T equal to a break point) still flow */
        if (fa == T(0.0)) return T(0.0);
        const T A = T(raymond_A(double(T(c.te0)), double(temperature)));
        /* the reference returns through int */
        return T(ref_int(T(0.5) * (A * fa + B * (T(1.0) - signb * signc) + C * (T(1.0) - signc * signd))));
    }
    if (c.cooling == 2) {
        const T temperature = pr / rho;
        if (temperature > T(c.cool_tmin))
            return c.cool_index == 0.5 ? T(c.cool_lambda0) * sqrt(temperature)
                                       : T(c.cool_lambda0) * T(powerlaw_pow(double(temperature), c.cool_index));
        return T(0.0);
    }
    return T(0.0);
}

/* (b2 + gp)^2 - 4 gp bn^2 >= (b2 - gp)^2 >= 0.  With FP32 state the two
 * products may be fused into an FMA, which can leave it a few ulps below
 * zero when b2 = bn^2 and b2 ~ gp (a purely normal field at beta ~ 1.2);
 * that case is clamped at 0.  FP64 keeps the reference's unfused form. */
template <typename T>
__device__ __forceinline__ T fast_disc(T b2, T gmpr, T bn)
{
    const T d = (b2 + gmpr) * (b2 + gmpr) - T(4.0) * gmpr * bn * bn;
    return sizeof(T) == 4 ? fmaxf(float(d), 0.0f) : d;
}

/* ---- limiter interface states ------------------------------------------
 * upper = state at the cell's +side face (Vl of that face),
 * lower = state at the cell's -side face (Vr of that face).
 * LIM 0 Minmod: limiterx.h:171-217 / limitery.h:46-94 / limiterz.h:146-190
 *   (MUSCL-Hancock factor 0.5*(1 -/+ lambda dt/dx)); the y factor is
 *   ((lambda*dt)*oneox)*oneody, x and z pass f2 = 1 (exact).
 * LIM 1 MC: limiterx.h:335-385 / limitery.h:197-247 / limiterz.h:294-358.
 * LIM 2 vanLeer: limiterx.h:266-326 / limitery.h:129-188 / limiterz.h:232-285.
 * LIM 3 superbee: limiterx.h:394-449 / limitery.h:256-311 / limiterz.h:367-432. */
template <int DIR, int LIM, typename T>
__device__ __forceinline__ void cell_states(const Consts &c, const T *vm, const T *v0, const T *vp,
                                            T f1hi, T f1lo, T f2, T *upper, T *lower)
{
    T lambmax = T(0), lambmin = T(0);
    if (LIM == 0) {
        const int nb = 4 + DIR, nv = 1 + DIR;
        const T b2 = (v0[4] * v0[4] + v0[5] * v0[5] + v0[6] * v0[6]);
        const T gmpr = T(c.gm) * v0[7];
        const T fast = sqrt(((b2 + gmpr) + sqrt(fast_disc(b2, gmpr, v0[nb]))) / (T(2.0) * v0[0]));
        lambmax = rmax(T(0), v0[nv] + fast);
        lambmin = rmin(T(0), v0[nv] - fast);
    }
    const T dt = T(c.dt);
    if (sizeof(T) == 4 && (LIM == 0 || LIM == 1)) {
        /* FP32 state: the same slopes as medians, one v_med3 each.
         * minmod(a, b) = med3(a, b, 0): with a b <= 0 the median is 0,
         * otherwise the one of smaller magnitude; and minmod(a, b, c) =
         * minmod(minmod(a, b), c) (the MC centred slope c has the sign of a
         * and b when they agree).  Medians round nothing, so the result
         * equals the reference form except where a*b underflows (and in
         * the sign of a zero slope). */
        const T hu = T(0.5) * (T(1.0) - lambmax * dt * f1hi * f2);
        const T hl = T(0.5) * (T(1.0) + lambmin * dt * f1lo * f2);
#pragma unroll
        for (int m = 0; m < 9; m++) {
            const float a = v0[m] - vm[m];
            const float b = vp[m] - v0[m];
            const float mm = __builtin_amdgcn_fmed3f(a, b, 0.0f);
            if (LIM == 1) {
                const float cc = 0.25f * (vp[m] - vm[m]);
                const float grad = __builtin_amdgcn_fmed3f(mm, cc, 0.0f);
                if (upper) upper[m] = v0[m] + grad;
                if (lower) lower[m] = v0[m] - grad;
            } else {
                if (upper) upper[m] = v0[m] + hu * mm;
                if (lower) lower[m] = v0[m] - hl * mm;
            }
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < 9; m++) {
        const T a = v0[m] - vm[m];
        const T b = vp[m] - v0[m];
        T grad;
        if (LIM == 1) {
            const T cc = T(0.25) * (vp[m] - vm[m]);
            if (a * b <= T(0.0)) grad = T(0.0);
            else if (a > T(0.0)) grad = rmin(a, rmin(b, cc));
            else grad = rmax(a, rmax(b, cc));
            if (upper) upper[m] = v0[m] + grad;
            if (lower) lower[m] = v0[m] - grad;
        } else if (LIM == 2) {
            if (a * b <= T(0.0)) grad = T(0.0);
            else grad = a * b / (a + b);
            if (upper) upper[m] = v0[m] + grad;
            if (lower) lower[m] = v0[m] - grad;
        } else if (LIM == 3) {
            if (a * b <= T(0.0)) grad = T(0.0);
            else if (fabs(a) >= fabs(b)) grad = a > T(0.0) ? rmin(a, T(2.0) * b) : rmax(a, T(2.0) * b);
            else grad = a > T(0.0) ? rmin(T(2.0) * a, b) : rmax(T(2.0) * a, b);
            if (upper) upper[m] = v0[m] + T(0.5) * grad;
            if (lower) lower[m] = v0[m] - T(0.5) * grad;
        } else {
            if (a * b <= T(0.0)) grad = T(0.0);
            else if (a > T(0.0)) grad = rmin(a, b);
            else grad = rmax(a, b);
            if (upper) upper[m] = (v0[m] + T(0.5) * (T(1.0) - lambmax * dt * f1hi * f2) * grad);
            if (lower) lower[m] = (v0[m] - T(0.5) * (T(1.0) + lambmin * dt * f1lo * f2) * grad);
        }
    }
}

/* ---- VlrtoUlr (vlrtoulr.h:25-43) --------------------------------------- */
template <typename T>
__device__ __forceinline__ void vlr_to_ulr(const Consts &c, T v2, T b2, T r, const T *V, T *U)
{
    U[0] = V[0];
    U[1] = V[0] * V[1];
    U[2] = r * V[0] * V[2];
    U[3] = V[0] * V[3];
    U[4] = V[4];
    U[5] = V[5];
    U[6] = V[6];
    U[7] = T(0.5) * (V[0] * v2 + b2) + T(c.oneogm1) * V[7];
}

/* ---- physical flux from (V,U): vlrulrtoflr.h:24-51, vlrulrtoglr.h:25-53,
 *      vlrulrtohlr.h:24-52.  geo = xm[i] (x) or x[i] (y, z); oneox for y. */
template <int DIR, typename T>
__device__ __forceinline__ void phys_flux(T b2, T vdotb, const T *V, const T *U, T geo, T oneox, T *F)
{
    if (DIR == 0) {
        F[0] = U[1];
        F[1] = U[1] * V[1] + (V[7] + T(0.5) * b2) - V[4] * V[4];
        F[2] = U[2] * V[1] - geo * V[5] * V[4];
        F[3] = U[3] * V[1] - V[6] * V[4];
        F[4] = T(0.0);
        F[5] = V[5] * V[1] - V[2] * V[4];
        F[6] = V[6] * V[1] - V[3] * V[4];
        F[7] = (U[7] + V[7] + T(0.5) * b2) * V[1] - vdotb * V[4];
    } else if (DIR == 1) {
        F[0] = U[2] * oneox;
        F[1] = U[1] * V[2] - V[4] * V[5];
        F[2] = U[2] * V[2] + geo * ((V[7] + T(0.5) * b2) - V[5] * V[5]);
        F[3] = U[3] * V[2] - V[6] * V[5];
        F[4] = V[4] * V[2] - V[1] * V[5];
        F[5] = T(0.0);
        F[6] = V[6] * V[2] - V[3] * V[5];
        F[7] = (U[7] + V[7] + T(0.5) * b2) * V[2] - vdotb * V[5];
    } else {
        F[0] = U[3];
        F[1] = U[1] * V[3] - V[4] * V[6];
        F[2] = U[2] * V[3] - geo * V[5] * V[6];
        F[3] = U[3] * V[3] + (V[7] + T(0.5) * b2) - V[6] * V[6];
        F[4] = V[4] * V[3] - V[1] * V[6];
        F[5] = V[5] * V[3] - V[2] * V[6];
        F[6] = T(0.0);
        F[7] = (U[7] + V[7] + T(0.5) * b2) * V[3] - vdotb * V[6];
    }
}

/* ---- flux from star-state primitives: primitivetof.h:22-41,
 *      primitivetog.h, primitivetoh.h (physical r, phi, z components) --- */
template <int DIR, typename T>
__device__ __forceinline__ void prim_flux(T vdotb, T rho, T vx, T vy, T vz, T bx, T by, T bz, T en,
                                          T ptot, T geo, T *F)
{
    if (DIR == 0) {
        F[0] = rho * vx;
        F[1] = rho * vx * vx + ptot - bx * bx;
        F[2] = geo * (rho * vy * vx - by * bx);
        F[3] = rho * vz * vx - bz * bx;
        F[4] = T(0.0);
        F[5] = by * vx - vy * bx;
        F[6] = bz * vx - vz * bx;
        F[7] = (en + ptot) * vx - vdotb * bx;
    } else if (DIR == 1) {
        F[0] = rho * vy;
        F[1] = rho * vx * vy - bx * by;
        F[2] = geo * (rho * vy * vy + ptot - by * by);
        F[3] = rho * vz * vy - bz * by;
        F[4] = bx * vy - vx * by;
        F[5] = T(0.0);
        F[6] = bz * vy - vz * by;
        F[7] = (en + ptot) * vy - vdotb * by;
    } else {
        F[0] = rho * vz;
        F[1] = rho * vx * vz - bx * bz;
        F[2] = geo * (rho * vy * vz - by * bz);
        F[3] = rho * vz * vz + ptot - bz * bz;
        F[4] = bx * vz - vx * bz;
        F[5] = by * vz - vy * bz;
        F[6] = T(0.0);
        F[7] = (en + ptot) * vz - vdotb * bz;
    }
}

/* ---- GLM flux part (hlldx.h:482-494 and its y/z twins) ---------------- */
template <int DIR, typename T>
__device__ __forceinline__ void glm_flux(T ch, const T *vl, const T *vr, T *F)
{
    const int nb = 4 + DIR;
    F[nb] += T(0.5) * (vl[8] + vr[8] + ch * (vl[nb] - vr[nb]));
    F[8] = T(0.5) * ch * (ch * (vl[nb] + vr[nb]) + (vl[8] - vr[8]));
}

template <typename T>
__device__ __forceinline__ T fast_speed(T gm, const T *V, int nb, T b2)
{
    const T gmpr = gm * V[7];
    return sqrt(((b2 + gmpr) + sqrt(fast_disc(b2, gmpr, V[nb]))) / (T(2.0) * V[0]));
}

/* ---- HLLD (hlldx.h:53-477, hlldy.h:55-485, hlldz.h:53-481) ------------- */
template <int DIR, typename T>
__device__ __forceinline__ void hlld(const Consts &c, const T *Vl, const T *Vr, T geo, T oneox, T *F)
{
    constexpr int n = 1 + DIR, nb = 4 + DIR;
    constexpr int t1 = DIR == 0 ? 2 : (DIR == 1 ? 3 : 1);
    constexpr int t2 = DIR == 0 ? 3 : (DIR == 1 ? 1 : 2);
    const T gm = T(c.gm);
    T Ul[8], Ur[8];
    const T vl2 = Vl[1] * Vl[1] + Vl[2] * Vl[2] + Vl[3] * Vl[3];
    const T bl2 = Vl[4] * Vl[4] + Vl[5] * Vl[5] + Vl[6] * Vl[6];
    const T vldotbl = Vl[1] * Vl[4] + Vl[2] * Vl[5] + Vl[3] * Vl[6];
    const T ptl = Vl[7] + T(0.5) * bl2;
    const T vfl = fast_speed(gm, Vl, nb, bl2);
    const T vr2 = Vr[1] * Vr[1] + Vr[2] * Vr[2] + Vr[3] * Vr[3];
    const T br2 = Vr[4] * Vr[4] + Vr[5] * Vr[5] + Vr[6] * Vr[6];
    const T vrdotbr = Vr[1] * Vr[4] + Vr[2] * Vr[5] + Vr[3] * Vr[6];
    const T ptr = Vr[7] + T(0.5) * br2;
    const T vfr = fast_speed(gm, Vr, nb, br2);
    const T sl = rmin(Vl[n], Vr[n]) - rmax(vfl, vfr);
    const T sr = rmax(Vl[n], Vr[n]) + rmax(vfl, vfr);
    if (sl >= T(0.0)) {
        vlr_to_ulr(c, vl2, bl2, geo, Vl, Ul);
        phys_flux<DIR>(bl2, vldotbl, Vl, Ul, geo, oneox, F);
    } else if (sr <= T(0.0)) {
        vlr_to_ulr(c, vr2, br2, geo, Vr, Ur);
        phys_flux<DIR>(br2, vrdotbr, Vr, Ur, geo, oneox, F);
    } else {
        vlr_to_ulr(c, vl2, bl2, geo, Vl, Ul);
        vlr_to_ulr(c, vr2, br2, geo, Vr, Ur);
        const T slvl = sl - Vl[n];
        const T srvr = sr - Vr[n];
        T sm;
        if (DIR == 1)
            sm = ((srvr * Vr[0] * Vr[2] - slvl * Vl[0] * Vl[2]) - ptr + ptl) / (srvr * Ur[0] - slvl * Ul[0]);
        else
            sm = (srvr * Ur[n] - slvl * Ul[n] - ptr + ptl) / (srvr * Ur[0] - slvl * Ul[0]);
        const T oneoslsm = T(1.0) / (sl - sm);
        const T oneosrsm = T(1.0) / (sr - sm);
        const T ptint = (srvr * Ur[0] * ptl - slvl * Ul[0] * ptr + Ur[0] * Ul[0] * srvr * slvl * (Vr[n] - Vl[n]))
                        / (srvr * Ur[0] - slvl * Ul[0]);
        const T bn = (sr * Ur[nb] - sl * Ul[nb]) / (sr - sl);
        const T rholl = Ul[0] * slvl * oneoslsm;
        const T rhorr = Ur[0] * srvr * oneosrsm;
        T D, sgn, den, num;
        /* left-left state, branchless degeneracy switch (hlldx.h:222-247) */
        D = Ul[0] * slvl * (sl - sm) - bn * bn;
        sgn = (fabs(fabs(D) - T(1.0e-8) * bn * bn) / (fabs(D) - T(1.0e-8) * bn * bn));
        den = ((T(1.0) / D + T(1.0)) + (T(1.0) / D - T(1.0)) * sgn) * T(0.5);
        num = ((((Ul[0] * slvl * slvl - bn * bn) + T(1.0)) + ((Ul[0] * slvl * slvl - bn * bn) - T(1.0)) * sgn) * T(0.5));
        const T vt1ll = Vl[t1] - (bn * Ul[t1 + 3] * (sm - Vl[n]) * ((den + den * sgn) * T(0.5)));
        const T vt2ll = Vl[t2] - (bn * Ul[t2 + 3] * (sm - Vl[n]) * ((den + den * sgn) * T(0.5)));
        const T bt1ll = Ul[t1 + 3] * num * den;
        const T bt2ll = Ul[t2 + 3] * num * den;
        const T vlldotbll = sm * bn + vt1ll * bt1ll + vt2ll * bt2ll;
        /* right-right state (hlldx.h:295-320) */
        D = Ur[0] * srvr * (sr - sm) - bn * bn;
        sgn = (fabs(fabs(D) - T(1.0e-8) * bn * bn) / (fabs(D) - T(1.0e-8) * bn * bn));
        den = ((T(1.0) / D + T(1.0)) + (T(1.0) / D - T(1.0)) * sgn) * T(0.5);
        num = ((((Ur[0] * srvr * srvr - bn * bn) + T(1.0)) + ((Ur[0] * srvr * srvr - bn * bn) - T(1.0)) * sgn) * T(0.5));
        const T vt1rr = Vr[t1] - (bn * Ur[t1 + 3] * (sm - Vr[n]) * ((den + den * sgn) * T(0.5)));
        const T vt2rr = Vr[t2] - (bn * Ur[t2 + 3] * (sm - Vr[n]) * ((den + den * sgn) * T(0.5)));
        const T bt1rr = Ur[t1 + 3] * num * den;
        const T bt2rr = Ur[t2 + 3] * num * den;
        const T vrrdotbrr = sm * bn + vt1rr * bt1rr + vt2rr * bt2rr;
        const T enll = (slvl * Ul[7] - ptl * Vl[n] + ptint * sm + bn * (vldotbl - vlldotbll)) * oneoslsm;
        const T enrr = (srvr * Ur[7] - ptr * Vr[n] + ptint * sm + bn * (vrdotbr - vrrdotbrr)) * oneosrsm;
        const T sqrtrholl = sqrt(rholl);
        const T sqrtrhorr = sqrt(rhorr);
        const T sll = sm - fabs(bn) / sqrtrholl;
        const T srr = sm + fabs(bn) / sqrtrhorr;
        T v[4], b[4], vdotb, rho, en;
        if (sm >= T(0.0) && sll >= T(0.0)) {
            v[t1] = vt1ll; v[t2] = vt2ll; b[t1] = bt1ll; b[t2] = bt2ll;
            vdotb = vlldotbll; rho = rholl; en = enll;
        } else if (sm < T(0.0) && srr <= T(0.0)) {
            v[t1] = vt1rr; v[t2] = vt2rr; b[t1] = bt1rr; b[t2] = bt2rr;
            vdotb = vrrdotbrr; rho = rhorr; en = enrr;
        } else {
            const T sg = fabs(bn) / bn;
            const T oneosqrtrho = T(1.0) / (sqrtrholl + sqrtrhorr);
            const T vt1int = (sqrtrholl * vt1ll + sqrtrhorr * vt1rr + (bt1rr - bt1ll) * sg) * oneosqrtrho;
            const T vt2int = (sqrtrholl * vt2ll + sqrtrhorr * vt2rr + (bt2rr - bt2ll) * sg) * oneosqrtrho;
            const T bt1int = (sqrtrholl * bt1rr + sqrtrhorr * bt1ll + sqrtrholl * sqrtrhorr * (vt1rr - vt1ll) * sg) * oneosqrtrho;
            const T bt2int = (sqrtrholl * bt2rr + sqrtrhorr * bt2ll + sqrtrholl * sqrtrhorr * (vt2rr - vt2ll) * sg) * oneosqrtrho;
            const T vintdotbint = sm * bn + vt1int * bt1int + vt2int * bt2int;
            v[t1] = vt1int; v[t2] = vt2int; b[t1] = bt1int; b[t2] = bt2int;
            vdotb = vintdotbint;
            if (sm >= T(0.0)) { rho = rholl; en = enll - sqrtrholl * (vlldotbll - vintdotbint) * sg; }
            else              { rho = rhorr; en = enrr + sqrtrhorr * (vrrdotbrr - vintdotbint) * sg; }
        }
        v[n] = sm;
        b[n] = bn;
        prim_flux<DIR>(vdotb, rho, v[1], v[2], v[3], b[1], b[2], b[3], en, ptint, geo, F);
    }
    glm_flux<DIR>(T(c.ch), Vl, Vr, F);
}

/* ---- HLL (hllx.h:43-179, hlly.h, hllz.h) ------------------------------ */
template <int DIR, typename T>
__device__ __forceinline__ void hll(const Consts &c, const T *Vl, const T *Vr, T geo, T oneox, T *F)
{
    constexpr int n = 1 + DIR, nb = 4 + DIR;
    const T gm = T(c.gm);
    T Ul[8], Ur[8], Fl[8], Fr[8];
    T v2 = Vl[1] * Vl[1] + Vl[2] * Vl[2] + Vl[3] * Vl[3];
    T b2 = Vl[4] * Vl[4] + Vl[5] * Vl[5] + Vl[6] * Vl[6];
    T vdotb = Vl[1] * Vl[4] + Vl[2] * Vl[5] + Vl[3] * Vl[6];
    vlr_to_ulr(c, v2, b2, geo, Vl, Ul);
    phys_flux<DIR>(b2, vdotb, Vl, Ul, geo, oneox, Fl);
    const T vfl = fast_speed(gm, Vl, nb, b2);
    v2 = Vr[1] * Vr[1] + Vr[2] * Vr[2] + Vr[3] * Vr[3];
    b2 = Vr[4] * Vr[4] + Vr[5] * Vr[5] + Vr[6] * Vr[6];
    vdotb = Vr[1] * Vr[4] + Vr[2] * Vr[5] + Vr[3] * Vr[6];
    vlr_to_ulr(c, v2, b2, geo, Vr, Ur);
    phys_flux<DIR>(b2, vdotb, Vr, Ur, geo, oneox, Fr);
    const T vfr = fast_speed(gm, Vr, nb, b2);
    const T sl = rmin(Vl[n], Vr[n]) - rmax(vfl, vfr);
    const T sr = rmax(Vl[n], Vr[n]) + rmax(vfl, vfr);
    if (sl >= T(0.0)) {
#pragma unroll
        for (int m = 0; m < 8; m++) F[m] = Fl[m];
    } else if (sr >= T(0.0)) {
#pragma unroll
        for (int m = 0; m < 8; m++) F[m] = (sr * Fl[m] - sl * Fr[m] + sr * sl * (Ur[m] - Ul[m])) / (sr - sl);
    } else {
#pragma unroll
        for (int m = 0; m < 8; m++) F[m] = Fr[m];
    }
    glm_flux<DIR>(T(c.ch), Vl, Vr, F);
}

template <int DIR, bool HLL, typename T>
__device__ __forceinline__ void riemann(const Consts &c, const T *Vl, const T *Vr, T geo, T oneox, T *F)
{
    if (HLL) hll<DIR>(c, Vl, Vr, geo, oneox, F);
    else hlld<DIR>(c, Vl, Vr, geo, oneox, F);
}

/* ---- source terms of one cell -------------------------------------------
 * SpiralGravityCooling (source.h:26-134, the default), AxiGravityCooling
 * (source.h:143-200) and AxiGravity (source.h:207-261) share one form: gx,
 * gy, gz are Grx.. in the spiral region, else Grxaxi, 0, Grzaxi (everywhere
 * for the two axisymmetric ones, whose callers pass spiral = false); AxiGravity
 * has no cooling term.  UniformGz (source.h:269-300) and NoSource
 * (source.h:308-349) use none of the gravity arguments; NoSource leaves S8
 * as it was, which is 0 in every run.  ALT = false compiles the default
 * (SpiralGravityCooling) only, for kernels launched with it.               */
template <typename T, bool ALT = true>
__device__ __forceinline__ void sources(const Consts &c, const T *V, T x, T oneox, T gx, T gy, T gz,
                                        bool spiral, T *S)
{
    const T V0 = V[0], V1 = V[1], V2 = V[2], V3 = V[3], V4 = V[4], V5 = V[5], V6 = V[6], V7 = V[7];
    if (ALT && c.source == 3) {
#pragma unroll
        for (int m = 0; m < 9; m++) S[m] = T(0.0);
        S[3] = -V0 * T(1.0);
        S[7] = -V3 * V0 * T(1.0);
        return;
    }
    if (ALT && c.source == 4) {
#pragma unroll
        for (int m = 0; m < 9; m++) S[m] = T(0.0);
        S[1] = (V0 * V2 * V2 - V5 * V5 + (V7 + T(0.5) * (V4 * V4 + V5 * V5 + V6 * V6))) * oneox;
        S[5] = (V5 * V1 - V2 * V4) * oneox;
        return;
    }
    const T om = T(c.omegasp);
    if (ALT && c.source == 2) {
        S[0] = T(0.0);
        S[1] = ((V0 * V2 * V2 - V5 * V5 + (V7 + T(0.5) * (V4 * V4 + V5 * V5 + V6 * V6))) * oneox
                + V0 * gx + V0 * om * om * x + (T(2.0) * V0 * V2 * om));
        S[2] = x * (T(-2.0) * V0 * V1 * om);
        S[3] = V0 * gz;
        S[4] = T(0.0);
        S[5] = (V5 * V1 - V2 * V4) * oneox;
        S[6] = T(0.0);
        S[7] = (V0 * (V1 * gx + V3 * gz) + V0 * x * V1 * om * om);
        S[8] = T(0.0);
        return;
    }
    const T lam = cooling(c, V0, V7);
    S[0] = T(0.0);
    S[1] = ((V0 * V2 * V2 - V5 * V5 + (V7 + T(0.5) * (V4 * V4 + V5 * V5 + V6 * V6))) * oneox
            + V0 * gx + V0 * om * om * x + (T(2.0) * V0 * V2 * om));
    if (spiral) S[2] = x * (V0 * gy - (T(2.0) * V0 * V1 * om));
    else        S[2] = x * (T(-2.0) * V0 * V1 * om);
    S[3] = V0 * gz;
    S[4] = T(0.0);
    S[5] = (V5 * V1 - V2 * V4) * oneox;
    S[6] = T(0.0);
    if (spiral) S[7] = (V0 * (V1 * gx + V2 * gy + V3 * gz) - (V0 * V0 * lam) + V0 * x * V1 * om * om);
    else        S[7] = (V0 * (V1 * gx + V3 * gz) - (V0 * V0 * lam) + V0 * x * V1 * om * om);
    S[8] = T(0.0);
}

/* ---- UtoV (utov.h:24-68) ------------------------------------------------ */
template <typename T>
__device__ __forceinline__ void utov(const Consts &c, const T *U, T oneox, T *V)
{
    const T oneorho = T(1.0) / U[0];
    const T b2 = U[4] * U[4] + U[5] * U[5] + U[6] * U[6];
    const T r2v2 = U[1] * U[1] + U[2] * U[2] * oneox * oneox + U[3] * U[3];
    V[0] = U[0];
    V[1] = U[1] * oneorho;
    V[2] = U[2] * oneorho * oneox;
    V[3] = U[3] * oneorho;
    V[4] = U[4];
    V[5] = U[5];
    V[6] = U[6];
    V[7] = T(c.gm1) * (U[7] - T(0.5) * (r2v2 * oneorho + b2));
    V[8] = U[8];
}

/* ---- Check (check.h:23-188) on one cell; returns flags bit0 rho, bit1 p,
 *      bit2 beta ------------------------------------------------------------ */
template <typename T>
__device__ __forceinline__ int check(const Consts &c, T *U, T *V, T x, T rhofloor)
{
    int flags = 0;
    const T v2 = V[2] * V[2];
    const T b2 = (V[4] * V[4] + V[5] * V[5] + V[6] * V[6]);
    const T oneogm1 = T(c.oneogm1);
    if (V[0] < rhofloor) {
        flags |= 1;
        V[0] = rhofloor;
        V[1] = T(0.0);
        V[3] = T(0.0);
        U[0] = V[0];
        U[1] = T(0.0);
        U[2] = x * V[0] * V[2];
        U[3] = T(0.0);
        U[7] = T(0.5) * (V[0] * v2 + b2) + oneogm1 * V[7];
    }
    const T prfloor = V[0] * T(c.prcr);
    if (V[7] < prfloor) {
        flags |= 2;
        V[1] = T(0.0);
        V[3] = T(0.0);
        V[7] = prfloor;
        U[1] = T(0.0);
        U[3] = T(0.0);
        U[7] = (T(0.5) * (V[0] * v2 + b2) + oneogm1 * V[7]);
    }
    const T oneobeta = T(0.5) * b2 / V[7];
    if (oneobeta > T(c.oneobetacr)) {
        flags |= 4;
        V[0] = T(c.betacr) * oneobeta * V[0];
        V[1] = T(0.0);
        V[3] = T(0.0);
        V[7] = T(c.betacr) * oneobeta * V[7];
        U[0] = V[0];
        U[1] = T(0.0);
        U[2] = x * V[0] * V[2];
        U[3] = T(0.0);
        U[7] = (T(0.5) * (V[0] * v2 + b2) + oneogm1 * V[7]);
    }
    return flags;
}

}  // namespace cyl
#endif
