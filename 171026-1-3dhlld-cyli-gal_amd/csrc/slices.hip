/*
 * slices.hip -- planes and azimuthal lines of primitive and derived
 * quantities, cut out of the FP64 primitive state V[9][NX][NY][NZ] on the
 * device for the text writers (GDLOutputCylinder, output.h:280-384, and the
 * slices the reference's IDL scripts read).
 *
 * One thread per output value; the reductions (means over k or j, the
 * column integral over k) run sequentially in that thread, in increasing
 * index order, so every value is one deterministic FP64 expression.  The
 * planes are small (NX*NY or 2*NX*NZ values): the cut is a strided gather,
 * bound by the latency of one pass, not by bandwidth.
 */
#include "paths.h"

namespace cyl {

#define C3(i, j, k) ((((size_t)(i)) * g.NY + (size_t)(j)) * g.NZ + (size_t)(k))

/* quantity q of cell (i, j, k) */
__device__ double slice_value(const DevGeom &g, const Consts &c, const double *__restrict__ V, size_t N, int q,
                              int i, int j, int k)
{
    const size_t cc = C3(i, j, k);
    if (q < 9) return V[(size_t)q * N + cc];
    const double rho = V[cc], pr = V[7 * N + cc];
    const double b4 = V[4 * N + cc], b5 = V[5 * N + cc], b6 = V[6 * N + cc];
    switch (q) {
    case CYLMHD_Q_BENERGY:   /* output.h:353-355 */
        return 0.5 * (b4 * b4 + b5 * b5 + b6 * b6);
    case CYLMHD_Q_BETA:
        return pr / (0.5 * (b4 * b4 + b5 * b5 + b6 * b6));
    case CYLMHD_Q_TEMP:      /* cooling.h:25 */
        return pr / rho;
    case CYLMHD_Q_ONEOTCOOL: /* setdt.h:56-58 */
        return 2.0 * ((rho * rho * cooling<double>(c, rho, pr)) / (1.5 * pr));
    case CYLMHD_Q_VDOTB:
        return V[1 * N + cc] * b4 + V[2 * N + cc] * b5 + V[3 * N + cc] * b6;
    case CYLMHD_Q_DIVB: {
        /* the finite-volume divergence of cylmhd_diagnostics: face values
         * as the mean of the two cells, metric factors g1, g2, g3, g5 */
        if (i < 1 || i > g.NX - 2 || j < 1 || j > g.NY - 2 || k < 1 || k > g.NZ - 2) return 0.0;
        const size_t si = (size_t)g.NY * g.NZ, sj = g.NZ, ik = (size_t)i * g.NZ + k;
        const double *Br = V + 4 * N, *Bp = V + 5 * N, *Bz = V + 6 * N;
        const double brp = 0.5 * (Br[cc] + Br[cc + si]), brm = 0.5 * (Br[cc] + Br[cc - si]);
        const double bpp = 0.5 * (Bp[cc] + Bp[cc + sj]), bpm = 0.5 * (Bp[cc] + Bp[cc - sj]);
        const double bzp = 0.5 * (Bz[cc] + Bz[cc + 1]), bzm = 0.5 * (Bz[cc] + Bz[cc - 1]);
        return (g.g2[ik] * brp - g.g1[ik] * brm) + (g.g3[ik] * bpp - g.g3[ik] * bpm)
               + (g.g5[ik] * bzp - g.g5[ik] * bzm);
    }
    case CYLMHD_Q_RHOBZ:
        return rho * b6;
    case CYLMHD_Q_SPIRALPOT: {
        /* E cos(phase) with E / r = the amplitude of the arm's phi force
         * eps m1 q s(z) (x/sr^2) (1+x^2/sr^2)^-1.5 m sin(phase)
         * (spiralgravity.h:93-98, m1 = 10): F_phi = -(1/r) dPhi/dphi */
        const size_t ik = (size_t)i * g.NZ + k;
        const double cosphase = g.sincos[(size_t)g.NX * g.NY + (size_t)i * g.NY + j];
        return g.epssp * 10.0 * g.sp_q[ik] * g.sp_szs[k] * g.sp_x2s2o15[i] * cosphase;
    }
    case CYLMHD_Q_VPARA:
    case CYLMHD_Q_VPERP: {
        /* arm direction: phase = m (phi + cot(p) ln r) + const is constant
         * along (sin p, -cos p) in (r, phi); oriented with the rotation */
        const double sp = 1.0 / sqrt(1.0 + g.cotpitch * g.cotpitch), cp = g.cotpitch * sp;
        const double vr = V[1 * N + cc], vp = V[2 * N + cc];
        return q == CYLMHD_Q_VPARA ? vp * cp - vr * sp : vr * cp + vp * sp;
    }
    }
    return 0.0;
}

__global__ __launch_bounds__(256) void k_slice(DevGeom g, Consts c, const double *__restrict__ V,
                                               const double *__restrict__ dzm, int q, int plane, int a, int b,
                                               double *__restrict__ out, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t N = (size_t)g.NX * g.NY * g.NZ;
    const int ni = g.NX - 2, nj = g.NY - 2, nk = g.NZ - 2;
    double v = 0.0;
    switch (plane) {
    case CYLMHD_PLANE_RPHI:
        v = slice_value(g, c, V, N, q, 1 + t / nj, 1 + t % nj, a);
        break;
    case CYLMHD_PLANE_RPHI_ZMEAN:
    case CYLMHD_PLANE_RPHI_ZINT: {
        const int i = 1 + t / nj, j = 1 + t % nj;
        const bool integ = plane == CYLMHD_PLANE_RPHI_ZINT;
        double s = 0.0;
        for (int k = 1; k <= nk; k++) {
            const double w = slice_value(g, c, V, N, q, i, j, k);
            s += integ ? w * dzm[k - 1] : w;
        }
        v = integ ? s : s / (double)nk;
        break;
    }
    case CYLMHD_PLANE_RZ:
        v = slice_value(g, c, V, N, q, 1 + t % ni, a, 1 + t / ni);
        break;
    case CYLMHD_PLANE_ZR:
        v = slice_value(g, c, V, N, q, 1 + t / nk, a, 1 + t % nk);
        break;
    case CYLMHD_PLANE_XZ:
    case CYLMHD_PLANE_XZ_PHIMEAN: {
        const int col = t % (2 * g.NX), k = 1 + t / (2 * g.NX);
        const bool opp = col < g.NX;
        const int i = opp ? g.NX - 1 - col : col - g.NX;
        if (plane == CYLMHD_PLANE_XZ) {
            const int j = opp ? 1 + (a - 1 + nj / 2) % nj : a;
            v = slice_value(g, c, V, N, q, i, j, k);
        } else {
            double s = 0.0;
            for (int j = 1; j <= nj; j++) s += slice_value(g, c, V, N, q, i, j, k);
            v = s / (double)nj;
        }
        break;
    }
    case CYLMHD_PLANE_PHI:
        v = slice_value(g, c, V, N, q, a, 1 + t, b);
        break;
    }
    out[t] = v;
}

void launch_slice(hipStream_t s, const DevGeom &g, const Consts &c, const double *V, const double *dzm, int q,
                  int plane, int a, int b, double *out, int n)
{
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_slice, dim3(nb), dim3(256), 0, s, g, c, V, dzm, q, plane, a, b, out, n);
}

}  // namespace cyl
