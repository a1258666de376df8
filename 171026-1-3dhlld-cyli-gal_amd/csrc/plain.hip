/*
 * plain.hip -- the plain FP64 path: the reference's data flow, one pass per
 * reference routine, every pass a gfx950 kernel.
 *
 *   k_glm          GLMSourceSolver   (glmsourcesolver.h:19-47)
 *   k_flux<DIR>    Minmod/MC/vanLeer/superbee limiter + interface BCs + HLLD/HLL + GLM flux,
 *                  one thread per face (limiterx.h / hlldx.h and y/z twins)
 *   k_update<KIND> sources (source.h:26-134) + stage update
 *                  (firststep.h:97-118 ...)
 *   k_ghost        the stage's cell boundary set as one composed map
 *   k_absorb       XRight/Central/ZRight absorbing layers
 *   k_utov_check   UtoV (utov.h:24-68) + Check (check.h:23-188)
 *
 * Only the faces the update reads are computed (the reference also fills
 * ghost-row faces that nothing reads).  V is stored between stages as in the
 * reference, so results match its FP64 path bit for bit except where libm
 * differs (pow in Cooling, sincos of the spiral phase).
 */
#include "paths.h"

namespace cyl {

#define C3(i, j, k) ((((size_t)(i)) * g.NY + (size_t)(j)) * g.NZ + (size_t)(k))
#define F3(i, j, k) ((((size_t)(i)) * (g.NY - 1) + (size_t)(j)) * (g.NZ - 1) + (size_t)(k))

__global__ void k_glm(double *__restrict__ U8, double *__restrict__ V8, size_t N, double psicoef)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
        U8[c] = psicoef * U8[c];
        V8[c] = psicoef * V8[c];
    }
}

__device__ __forceinline__ void load9(const double *__restrict__ A, size_t N, size_t c, double *v)
{
#pragma unroll
    for (int m = 0; m < 9; m++) v[m] = A[m * N + c];
}

/* upper (+face) or lower (-face) interface state of cell (ci,cj,ck) */
template <int DIR, int LIM, bool UPPER>
__device__ __forceinline__ void face_state(const DevGeom &g, const Consts &c, const double *__restrict__ V,
                                           size_t N, int ci, int cj, int ck, double *out)
{
    const size_t st = DIR == 0 ? (size_t)g.NY * g.NZ : (DIR == 1 ? (size_t)g.NZ : 1);
    const size_t cc = C3(ci, cj, ck);
    double vm[9], v0[9], vp[9];
    load9(V, N, cc - st, vm);
    load9(V, N, cc, v0);
    load9(V, N, cc + st, vp);
    double f1, f2 = 1.0;
    if (DIR == 0) f1 = UPPER ? g.oneodx[ci] : g.oneodx[ci - 1];
    else if (DIR == 1) { f1 = g.oneox[ci]; f2 = g.oneody; }
    else f1 = UPPER ? g.oneodz[ck] : g.oneodz[ck - 1];
    if (UPPER) cell_states<DIR, LIM>(c, vm, v0, vp, f1, f1, f2, out, (double *)nullptr);
    else cell_states<DIR, LIM>(c, vm, v0, vp, f1, f1, f2, (double *)nullptr, out);
}

/* One thread per face: left/right states from the limiter of the two cells
 * that share the face, with the interface-state boundary the limiter applies
 * on the first and last face (xvlrboundary.h, yvlrboundary.h, zvlrboundary.h
 * as called from limiter{x,y,z}.h), then the Riemann flux. */
template <int DIR, int LIM, bool HLL>
__global__ __launch_bounds__(256) void k_flux(DevGeom g, Consts c, const double *__restrict__ V,
                                              const double *__restrict__ Vinit_z, const double *__restrict__ Vinit_y,
                                              double *__restrict__ Fo, size_t N, size_t NF)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + (DIR == 2 ? 0 : 1);
    const int j = blockIdx.y * blockDim.y + threadIdx.y + (DIR == 1 ? 0 : 1);
    const int i = blockIdx.z + (DIR == 0 ? 0 : 1);
    const int kmax = g.NZ - 2;   /* inclusive upper bounds */
    const int jmax = g.NY - 2;
    if (k > kmax || j > jmax) return;
    double vl[9], vr[9], fx[9];
    if (DIR == 0) {
        const int last = g.NX - 2;
        if (LIM == 2) {
            /* XVlrFreeBoundary: Vl[0] = Vr[0], Vr[NX-2] = Vl[NX-2] */
            if (i == 0) face_state<0, LIM, false>(g, c, V, N, 1, j, k, vl);
            else face_state<0, LIM, true>(g, c, V, N, i, j, k, vl);
            if (i == last) face_state<0, LIM, true>(g, c, V, N, last, j, k, vr);
            else face_state<0, LIM, false>(g, c, V, N, i + 1, j, k, vr);
        } else {
            face_state<0, LIM, true>(g, c, V, N, i == 0 ? last : i, j, k, vl);       /* XVlrPeriodic */
            face_state<0, LIM, false>(g, c, V, N, i == last ? 1 : i + 1, j, k, vr);
        }
    } else if (DIR == 1) {
        const int last = g.NY - 2;
        if (LIM == 2 && j == 0) {
            /* YVlrInitializeBoundary after the periodic copy: Vinitial at j = 0 */
            for (int m = 0; m < 9; m++) vl[m] = Vinit_y[(size_t)m * 2 * g.NX * g.NZ + (size_t)i * g.NZ + k];
        } else {
            face_state<1, LIM, true>(g, c, V, N, i, j == 0 ? last : j, k, vl);       /* YVlrPeriodic */
        }
        if (LIM == 2 && j == last) {
            for (int m = 0; m < 9; m++)
                vr[m] = Vinit_y[(size_t)m * 2 * g.NX * g.NZ + (size_t)g.NX * g.NZ + (size_t)i * g.NZ + k];
        } else {
            face_state<1, LIM, false>(g, c, V, N, i, j == last ? 1 : j + 1, k, vr);
        }
    } else {
        const int last = g.NZ - 2;
        const size_t NP = (size_t)g.NX * g.NY;
        if (LIM == 1) {
            /* ZVlrFixedBoundary: Vinitial at k = 0 and k = NZ-1 */
            if (k == 0) { for (int m = 0; m < 9; m++) vl[m] = Vinit_z[(size_t)m * 2 * NP + (size_t)i * g.NY + j]; }
            else face_state<2, LIM, true>(g, c, V, N, i, j, k, vl);
            if (k == last) { for (int m = 0; m < 9; m++) vr[m] = Vinit_z[(size_t)m * 2 * NP + NP + (size_t)i * g.NY + j]; }
            else face_state<2, LIM, false>(g, c, V, N, i, j, k + 1, vr);
        } else if (LIM == 2) {
            /* ZVlrFreeBoundary: Vl[0] = Vr[0], Vr[NZ-2] = Vl[NZ-2] */
            if (k == 0) face_state<2, LIM, false>(g, c, V, N, i, j, 1, vl);
            else face_state<2, LIM, true>(g, c, V, N, i, j, k, vl);
            if (k == last) face_state<2, LIM, true>(g, c, V, N, i, j, last, vr);
            else face_state<2, LIM, false>(g, c, V, N, i, j, k + 1, vr);
        } else if (LIM == 3) {
            /* ZVlrPeriodicBoundary: Vl[0] = Vl[NZ-2], Vr[NZ-2] = Vr[0] */
            face_state<2, LIM, true>(g, c, V, N, i, j, k == 0 ? last : k, vl);
            face_state<2, LIM, false>(g, c, V, N, i, j, k == last ? 1 : k + 1, vr);
        } else {
            /* ZLeftVlVzAntisymBzsym: Vl[0] = Vr[1] (the lower state of cell 2),
             * vz negated; ZRightVrFree: Vr[NZ-2] = Vl[NZ-2] */
            if (k == 0) {
                face_state<2, LIM, false>(g, c, V, N, i, j, 2, vl);
                vl[3] = -vl[3];
            } else {
                face_state<2, LIM, true>(g, c, V, N, i, j, k, vl);
            }
            if (k == last) { for (int m = 0; m < 9; m++) vr[m] = vl[m]; }
            else face_state<2, LIM, false>(g, c, V, N, i, j, k + 1, vr);
        }
    }
    const double geo = DIR == 0 ? g.xm[i] : g.x[i];
    riemann<DIR, HLL>(c, vl, vr, geo, g.oneox[i], fx);
    const size_t f = F3(i, j, k);
#pragma unroll
    for (int m = 0; m < 9; m++) Fo[m * NF + f] = fx[m];
}

/* KIND: 1 FirstStep, 2 SecondStep, 3 ThirdStep, 4 Halfstep, 5 Fullstep */
template <int KIND>
__global__ __launch_bounds__(256) void k_update(DevGeom g, Consts c, const double *__restrict__ V,
                                                const double *__restrict__ F, const double *__restrict__ G,
                                                const double *__restrict__ H, const double *U,
                                                const double *U1, double *Uout, size_t N, size_t NF)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 1;
    if (k > g.NZ - 2 || j > g.NY - 2) return;
    const size_t cc = C3(i, j, k);
    const size_t ik = (size_t)i * g.NZ + k;
    double v[9], S[9];
    load9(V, N, cc, v);
    double gx, gy, gz;
    const bool spiral = gravity_cell(g, i, j, k, gx, gy, gz, c.source == 1 || c.source == 2);
    sources(c, v, g.x[i], g.oneox[i], gx, gy, gz, spiral, S);
    const double g1 = g.g1[ik], g2 = g.g2[ik], g3 = g.g3[ik], g4 = g3, g5 = g.g5[ik], g6 = g5;
    const size_t f = F3(i, j, k), fi = F3(i - 1, j, k), fj = F3(i, j - 1, k), fk = F3(i, j, k - 1);
    const double dt = c.dt, oneo3 = 1.0 / 3.0;
#pragma unroll
    for (int m = 0; m < 9; m++) {
        const double div = ((g2 * F[m * NF + f] - g1 * F[m * NF + fi]) + (g4 * G[m * NF + f] - g3 * G[m * NF + fj])
                            + (g6 * H[m * NF + f] - g5 * H[m * NF + fk]));
        const double u = U[m * N + cc];
        double r;
        if (KIND == 1) r = u - dt * div + dt * S[m];
        else if (KIND == 2) r = 0.25 * (3.0 * u + U1[m * N + cc] - dt * div + dt * S[m]);
        else if (KIND == 3) r = oneo3 * (u + 2.0 * (U1[m * N + cc] - dt * div + dt * S[m]));
        else if (KIND == 4) r = (u - dt * div + dt * S[m]);
        else r = 0.5 * (u + U1[m * N + cc] - dt * div + dt * S[m]);
        Uout[m * N + cc] = r;
    }
}

/* ghost cells of the box, enumerated as i-faces, j-faces, k-faces; RK2 adds
 * the interior k = 1 plane whose vz momentum ZLeftVzAntisymBSym zeroes */
__global__ void k_ghost(DevGeom g, double *__restrict__ U, size_t N, bool rk3)
{
    const size_t NX = g.NX, NY = g.NY, NZ = g.NZ;
    const size_t nA = 2 * NY * NZ, nB = (NX - 2) * 2 * NZ, nC = (NX - 2) * (NY - 2) * 2;
    const size_t nD = rk3 ? 0 : (NX - 2) * (NY - 2);
    const size_t tot = nA + nB + nC + nD;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        size_t r = t;
        if (r < nA) { i = (r / (NY * NZ)) ? (int)NX - 1 : 0; r %= NY * NZ; j = (int)(r / NZ); k = (int)(r % NZ); }
        else if ((r -= nA) < nB) { i = 1 + (int)(r / (2 * NZ)); r %= 2 * NZ; j = (r / NZ) ? (int)NY - 1 : 0; k = (int)(r % NZ); }
        else if ((r -= nB) < nC) { i = 1 + (int)(r / ((NY - 2) * 2)); r %= (NY - 2) * 2; j = 1 + (int)(r / 2); k = (r % 2) ? (int)NZ - 1 : 0; }
        else {
            r -= nC;
            i = 1 + (int)(r / (NY - 2));
            j = 1 + (int)(r % (NY - 2));
            U[3 * N + C3(i, j, 1)] = 0.0;
            continue;
        }
        int si, sj, sk;
        unsigned sign;
        bool zero3;
        ghost_source(g, rk3, i, j, k, si, sj, sk, sign, zero3);
        const size_t src = C3(si, sj, sk), dst = C3(i, j, k);
#pragma unroll
        for (int m = 0; m < 9; m++) {
            if (m == 3 && zero3) { U[3 * N + dst] = 0.0; continue; }
            const double val = U[m * N + src];
            U[m * N + dst] = (sign >> m) & 1u ? -val : val;
        }
    }
}

/* XRightAbsorbBoundary (xboundary.h:124-149), CentralAbsorbBoundary
 * (centralboundary.h:27-52), ZRightAbsorbBoundary (zboundary.h:33-58);
 * applied per cell in that order */
__global__ void k_absorb(DevGeom g, double *__restrict__ U, const double *__restrict__ Uinit, size_t N,
                         const int *__restrict__ ilist, int ni, const int *__restrict__ olist, int no, bool zright)
{
    const size_t NY = g.NY, NZ = g.NZ;
    const size_t n1 = (size_t)ni * NY * NZ;
    const size_t n2 = zright ? (size_t)no * NY * 11 : 0;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n1 + n2; t += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (t < n1) {
            i = ilist[t / (NY * NZ)];
            j = (int)((t / NZ) % NY);
            k = (int)(t % NZ);
        } else {
            const size_t r = t - n1;
            i = olist[r / (NY * 11)];
            j = (int)((r / 11) % NY);
            k = (int)(NZ - 11 + r % 11);
        }
        const size_t cc = C3(i, j, k);
        const bool ax = i >= g.NX - 11, ac = i <= g.irin, az = zright && k >= g.NZ - 11;
        for (int m = 0; m < 9; m++) {
            double u = U[m * N + cc];
            const double ui = Uinit[m * N + cc];
            if (ax) u -= (g.xabs[i - (g.NX - 11)] * (u - ui));
            if (ac) u -= (g.cabs[(size_t)i * g.NZ + k] * (u - ui));
            if (az) u -= (g.zabs[k - (g.NZ - 11)] * (u - ui));
            U[m * N + cc] = u;
        }
    }
}

__global__ void k_utov_check(DevGeom g, Consts c, double *__restrict__ U, double *__restrict__ V,
                             const double *__restrict__ Uinit, size_t N, int *__restrict__ counts)
{
    const size_t NYZ = (size_t)g.NY * g.NZ;
    int f0 = 0, f1 = 0, f2 = 0;
    for (size_t cc = blockIdx.x * (size_t)blockDim.x + threadIdx.x; cc < N; cc += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(cc / NYZ);
        double u[9], v[9];
        load9(U, N, cc, u);
        utov(c, u, g.oneox[i], v);
        const int fl = check(c, u, v, g.x[i], Uinit[cc] * c.rhocr);
        if (fl) {
            U[cc] = u[0]; U[N + cc] = u[1]; U[2 * N + cc] = u[2]; U[3 * N + cc] = u[3]; U[7 * N + cc] = u[7];
        }
#pragma unroll
        for (int m = 0; m < 9; m++) V[m * N + cc] = v[m];
        f0 += fl & 1; f1 += (fl >> 1) & 1; f2 += (fl >> 2) & 1;
    }
    /* wave-level sums, one atomic per wave and counter */
    for (int off = 32; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off); f1 += __shfl_xor(f1, off); f2 += __shfl_xor(f2, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (f0) atomicAdd(&counts[0], f0);
        if (f1) atomicAdd(&counts[1], f1);
        if (f2) atomicAdd(&counts[2], f2);
    }
}

static unsigned grid1d(size_t n, unsigned bs, unsigned cap = 65536u)
{
    size_t b = (n + bs - 1) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

template <int DIR, int LIM, bool HLL>
static void launch_flux(hipStream_t s, const DevGeom &g, const Consts &c, const DevState &d, double *Fo)
{
    const int nk = DIR == 2 ? g.NZ - 1 : g.NZ - 2;
    const int nj = DIR == 1 ? g.NY - 1 : g.NY - 2;
    const int ni = DIR == 0 ? g.NX - 1 : g.NX - 2;
    dim3 bs(64, 4, 1), gs((nk + 63) / 64, (nj + 3) / 4, ni);
    hipLaunchKernelGGL((k_flux<DIR, LIM, HLL>), gs, bs, 0, s, g, c, d.V, d.Vinit_z, d.Vinit_y, Fo, d.N, d.NF);
}

template <int LIM, bool HLL>
static void launch_fluxes(hipStream_t s, const DevGeom &g, const Consts &c, const DevState &d)
{
    launch_flux<0, LIM, HLL>(s, g, c, d, d.F);
    launch_flux<1, LIM, HLL>(s, g, c, d, d.G);
    launch_flux<2, LIM, HLL>(s, g, c, d, d.H);
}

void plain_stage(hipStream_t s, const DevGeom &g, const DevState &d, const Consts &c, const StageArgs &a,
                 hipEvent_t ev0, hipEvent_t ev1)
{
    const size_t N = d.N;
    const bool last = a.rk3 ? a.stage == 3 : a.stage == 2;
    double *Uin = a.stage == 1 ? d.U : d.U1;
    double *Uout = last ? d.U : d.U1;
    hipLaunchKernelGGL(k_glm, dim3(grid1d(N, 256, 8192)), dim3(256), 0, s, Uin + 8 * N, d.V + 8 * N, N, c.psicoef);
    if (ev0) hipEventRecord(ev0, s);
    switch (a.limiter * 2 + (a.hll ? 1 : 0)) {
    case 0: launch_fluxes<0, false>(s, g, c, d); break;
    case 1: launch_fluxes<0, true>(s, g, c, d); break;
    case 2: launch_fluxes<1, false>(s, g, c, d); break;
    case 3: launch_fluxes<1, true>(s, g, c, d); break;
    case 4: launch_fluxes<2, false>(s, g, c, d); break;
    case 5: launch_fluxes<2, true>(s, g, c, d); break;
    case 6: launch_fluxes<3, false>(s, g, c, d); break;
    default: launch_fluxes<3, true>(s, g, c, d); break;
    }
    dim3 bs(64, 4, 1), gs((g.NZ - 2 + 63) / 64, (g.NY - 2 + 3) / 4, g.NX - 2);
    const int kind = a.rk3 ? a.stage : 3 + a.stage;
    switch (kind) {
    case 1: hipLaunchKernelGGL(k_update<1>, gs, bs, 0, s, g, c, d.V, d.F, d.G, d.H, d.U, d.U1, Uout, N, d.NF); break;
    case 2: hipLaunchKernelGGL(k_update<2>, gs, bs, 0, s, g, c, d.V, d.F, d.G, d.H, d.U, d.U1, Uout, N, d.NF); break;
    case 3: hipLaunchKernelGGL(k_update<3>, gs, bs, 0, s, g, c, d.V, d.F, d.G, d.H, d.U, d.U1, Uout, N, d.NF); break;
    case 4: hipLaunchKernelGGL(k_update<4>, gs, bs, 0, s, g, c, d.V, d.F, d.G, d.H, d.U, d.U1, Uout, N, d.NF); break;
    default: hipLaunchKernelGGL(k_update<5>, gs, bs, 0, s, g, c, d.V, d.F, d.G, d.H, d.U, d.U1, Uout, N, d.NF); break;
    }
    if (ev1) hipEventRecord(ev1, s);
    const size_t nghost = 2 * (size_t)g.NY * g.NZ + (size_t)(g.NX - 2) * 2 * g.NZ + (size_t)(g.NX - 2) * (g.NY - 2) * 2
                          + (a.rk3 ? 0 : (size_t)(g.NX - 2) * (g.NY - 2));
    hipLaunchKernelGGL(k_ghost, dim3(grid1d(nghost, 256, 4096)), dim3(256), 0, s, g, Uout, N, a.rk3);
    if (last) {
        const size_t nab = (size_t)d.n_absorb_i * g.NY * g.NZ + (a.rk3 ? 0 : (size_t)d.n_nonabsorb_i * g.NY * 11);
        hipLaunchKernelGGL(k_absorb, dim3(grid1d(nab, 256, 4096)), dim3(256), 0, s, g, Uout, d.Uinit, N,
                           d.absorb_i, d.n_absorb_i, d.nonabsorb_i, d.n_nonabsorb_i, !a.rk3);
    }
    hipLaunchKernelGGL(k_utov_check, dim3(grid1d(N, 256, 16384)), dim3(256), 0, s, g, c, Uout, d.V, d.Uinit, N,
                       d.counts + 3 * (a.stage - 1));
}

/* ---- per-step sin/cos of the spiral phase (spiralgravity.h:63-65) ------- */
__global__ void k_sincos(DevGeom g, double *__restrict__ out, const double *__restrict__ logxs, double t,
                         double domega)
{
    const size_t n = (size_t)g.NX * g.NY;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q / g.NY), j = (int)(q % g.NY);
        const double phase = g.armnum * (g.y[j] - domega * t + g.cotpitch * logxs[i]);
        double sp, cp;
        sincos(phase, &sp, &cp);
        out[q] = sp;
        out[n + q] = cp;
    }
}

void launch_sincos(hipStream_t s, const DevGeom &g, double *sc, const double *logxs, double t, double omegapattern,
                   double omegasp)
{
    const size_t n = (size_t)g.NX * g.NY;
    hipLaunchKernelGGL(k_sincos, dim3(grid1d(n, 256, 4096)), dim3(256), 0, s, g, sc, logxs, t, omegapattern - omegasp);
}

/* ---- Setdt (setdt.h:24-94): max of 1/dt with the first (i,j,k) attaining
 * it; FP64, deterministic (block partials, then one block) --------------- */
__device__ __forceinline__ bool better(double v, long long ix, double bv, long long bix)
{
    return dt_better(v, ix, bv, bix);
}

/* 256-thread blocks: lane shuffles within each wavefront, then 4 LDS slots */
__device__ void block_argmax(double &v, long long &ix)
{
    __shared__ double sv[4];
    __shared__ long long si[4];
    block_dt_max<4>(v, ix, sv, si);
}

__global__ __launch_bounds__(256) void k_setdt(DevGeom g, Consts c, const double *__restrict__ V, size_t N,
                                               double *__restrict__ red)
{
    const size_t ni = (size_t)(g.NX - 12), nj = (size_t)(g.NY - 2), nk = (size_t)(g.NZ - 2);
    const size_t M = ni * nj * nk;
    double best = 0.0;
    long long bix = LLONG_MAX;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < M; q += (size_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(q / (nj * nk)), j = 1 + (int)((q / nk) % nj), k = 1 + (int)(q % nk);
        const size_t cc = C3(i, j, k);
        const double V0 = V[cc], V1 = V[N + cc], V2 = V[2 * N + cc], V3 = V[3 * N + cc];
        const double V4 = V[4 * N + cc], V5 = V[5 * N + cc], V6 = V[6 * N + cc], V7 = V[7 * N + cc];
        const double v2 = V1 * V1 + V2 * V2 + V3 * V3;
        const double b2 = V4 * V4 + V5 * V5 + V6 * V6;
        const double t1 = sqrt(v2 + (c.gm * V7 + b2) / V0) / g.minlength[(size_t)i * g.NZ + k];
        const double tc = (2.0 * ((V0 * V0 * cooling(c, V0, V7)) / (1.5 * V7)));
        const double tt = rmax(t1, tc);
        if (better(tt, (long long)cc, best, bix) && tt > 0.0) { best = tt; bix = (long long)cc; }
    }
    block_argmax(best, bix);
    if (threadIdx.x == 0) { red[2 * blockIdx.x] = best; red[2 * blockIdx.x + 1] = (double)bix; }
}

__global__ __launch_bounds__(256) void k_argmax_final(double *__restrict__ red, int nb)
{
    double best = 0.0;
    long long bix = LLONG_MAX;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const double v = red[2 * b];
        const long long ix = red[2 * b + 1] >= 9.0e18 ? LLONG_MAX : (long long)red[2 * b + 1];
        if (better(v, ix, best, bix) && v > 0.0) { best = v; bix = ix; }
    }
    block_argmax(best, bix);
    __syncthreads();
    if (threadIdx.x == 0) { red[0] = best; red[1] = bix == LLONG_MAX ? -1.0 : (double)bix; }
}

void launch_setdt(hipStream_t s, const DevGeom &g, const Consts &c, const double *V, double *red, int nred)
{
    const size_t M = (size_t)(g.NX - 12) * (g.NY - 2) * (g.NZ - 2);
    const int nb = (int)grid1d(M, 256, (unsigned)(nred / 2));
    hipLaunchKernelGGL(k_setdt, dim3(nb), dim3(256), 0, s, g, c, V, (size_t)g.NX * g.NY * g.NZ, red);
    hipLaunchKernelGGL(k_argmax_final, dim3(1), dim3(256), 0, s, red, nb);
}

/* ---- conserved totals and L1 div B (FP64, deterministic) ---------------- */
__device__ void block_sum4(double *v)
{
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    for (int q = 0; q < 4; q++) sh[q][t] = v[q];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < 4; q++) sh[q][t] += sh[q][t + off];
        __syncthreads();
    }
    for (int q = 0; q < 4; q++) v[q] = sh[q][0];
}

__global__ __launch_bounds__(256) void k_diag(DevGeom g, const double *__restrict__ U, size_t N, double *__restrict__ red)
{
    const size_t ni = g.NX - 2, nj = g.NY - 2, nk = g.NZ - 2, M = ni * nj * nk;
    const size_t si = (size_t)g.NY * g.NZ, sj = g.NZ;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < M; q += (size_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(q / (nj * nk)), j = 1 + (int)((q / nk) % nj), k = 1 + (int)(q % nk);
        const size_t cc = C3(i, j, k), ik = (size_t)i * g.NZ + k;
        const double dv = g.dvol[ik];
        const double brp = 0.5 * (U[4 * N + cc] + U[4 * N + cc + si]), brm = 0.5 * (U[4 * N + cc] + U[4 * N + cc - si]);
        const double bpp = 0.5 * (U[5 * N + cc] + U[5 * N + cc + sj]), bpm = 0.5 * (U[5 * N + cc] + U[5 * N + cc - sj]);
        const double bzp = 0.5 * (U[6 * N + cc] + U[6 * N + cc + 1]), bzm = 0.5 * (U[6 * N + cc] + U[6 * N + cc - 1]);
        const double db = (g.g2[ik] * brp - g.g1[ik] * brm) + (g.g3[ik] * bpp - g.g3[ik] * bpm)
                          + (g.g5[ik] * bzp - g.g5[ik] * bzm);
        acc[0] += U[cc] * dv;
        acc[1] += U[7 * N + cc] * dv;
        acc[2] += fabs(db) * dv;
        acc[3] += fabs(U[8 * N + cc]) * dv;
    }
    block_sum4(acc);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) red[4 * blockIdx.x + q] = acc[q];
}

__global__ __launch_bounds__(256) void k_sum4_final(double *__restrict__ red, int nb)
{
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int q = 0; q < 4; q++) acc[q] += red[4 * b + q];
    block_sum4(acc);
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) red[q] = acc[q];
}

void launch_diag(hipStream_t s, const DevGeom &g, const double *U, double *red, int nred)
{
    const size_t M = (size_t)(g.NX - 2) * (g.NY - 2) * (g.NZ - 2);
    const int nb = (int)grid1d(M, 256, (unsigned)(nred / 4));
    hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, s, g, U, (size_t)g.NX * g.NY * g.NZ, red);
    hipLaunchKernelGGL(k_sum4_final, dim3(1), dim3(256), 0, s, red, nb);
}

}  // namespace cyl
