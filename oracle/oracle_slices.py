"""CPU restatement of the slice cuts and the text output format.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
cylmhd_slice, cylmhd_gdl_output(_ctx) and cylmhd_idl_output; nothing in the
library uses it.

Quantities (per cell, FP64, the expression order of the reference where it
has one):
  0..8     the primitives V[m]
  BENERGY  0.5 (b4 b4 + b5 b5 + b6 b6)           output.h:353-355
  BETA     pr / (0.5 (b4 b4 + b5 b5 + b6 b6))
  TEMP     pr / rho                              cooling.h:25
  ONEOTCOOL 2 ((rho rho Lambda(rho, pr)) / (1.5 pr))   setdt.h:56-58
  VDOTB    v1 b4 + v2 b5 + v3 b6
  DIVB     finite-volume divergence with face means and g1, g2, g3, g5
           (cylindricalgrid.h:114-125), 0 outside the interior
  RHOBZ    rho b6
  SPIRALPOT eps m1 q s(z) x^2/sr^2 (1+x^2/sr^2)^-1.5 cos(phase): the potential
           whose -(1/r) d/dphi is the arm's phi force (spiralgravity.h:37-98)
  VPARA, VPERP  velocity along / across the arms (phase = const), pitch p:
           vphi cos p - vr sin p,  vr cos p + vphi sin p
Planes (file order):
  RPHI     k = a, i outer, j inner, interior      output.h:318-330
  RZ       j = a, k outer, i inner, interior      output.h:366-380
  ZR       j = a, i outer, k inner: fltarr(nz, nx) + readf (density-rz.pro:27-61)
  XZ       k outer; the opposite half-plane from i = NX-1 down to 0, then
           j = a from i = 0 up: fltarr(2 NX, nz) (density-xz.pro:21-61)
  *_ZMEAN / *_ZINT / *_PHIMEAN  sequential sums in increasing index order
  PHI      i = a, k = b, interior j
Text: one "%e\\n" per value (output.h:288-380).
"""
import os

import numpy as np

Q_BENERGY, Q_BETA, Q_TEMP, Q_ONEOTCOOL, Q_VDOTB, Q_DIVB, Q_RHOBZ, Q_SPIRALPOT, Q_VPARA, Q_VPERP = range(9, 19)


def spiral_tables(p, x, y, z, t):
    """The arm factors of SpiralGravity (spiralgravity.h:37-77) at time t:
    q[i,k] = 1/sqrt(x^2 + (a + sqrt(z^2 + b^2))^2), s[k] = scalez/sqrt(z^2 +
    scalez^2), x2s2o15[i] = x^2/sr^2 (1 + x^2/sr^2)^-1.5, cos(phase)[i,j] with
    phase = m (y - (Omega_p - Omega_sp) t + cot(pitch) ln(x/scalephase)) and
    eps(t) (spiralgravity.h:41-42)."""
    a1, b1 = 0.65, 0.026
    xs = x * (1.0 / p.scaler)
    x2s2 = xs * xs
    x2s2o15 = x2s2 * (1.0 / (1.0 + x2s2)) ** 1.5
    q = 1.0 / np.sqrt((x * x)[:, None] + ((a1 + np.sqrt(z * z + b1 * b1)) * (a1 + np.sqrt(z * z + b1 * b1)))[None, :])
    szs = p.scalez / np.sqrt(z * z + p.scalez * p.scalez)
    cot = 1.0 / np.tan(p.pitch)
    phase = p.armnum * (y[None, :] - (p.omegapattern - p.omegasp) * t + cot * np.log(x * (1.0 / p.scalephase))[:, None])
    sw = (t - p.tausp) / abs(t - p.tausp)
    eps = 0.5 * p.epssp0 * (t / p.tausp * (1.0 - sw) + (1.0 + sw))
    return dict(q=q, szs=szs, x2s2o15=x2s2o15, cosphase=np.cos(phase), eps=eps, cot=cot)


(PLANE_RPHI, PLANE_RPHI_ZMEAN, PLANE_RPHI_ZINT, PLANE_RZ, PLANE_ZR, PLANE_XZ,
 PLANE_XZ_PHIMEAN, PLANE_PHI) = range(8)


def quantity(V, q, geo=None, cooling=None, spiral=None):
    """q over the whole grid, shape (NX, NY, NZ).  geo = the oracle's
    geometry tables [6][NX][NZ] (DIVB); cooling(rho, pr) -> Lambda (ONEOTCOOL);
    spiral = spiral_tables(...) (SPIRALPOT, VPARA, VPERP)."""
    if q < 9:
        return V[q]
    rho, pr = V[0], V[7]
    b4, b5, b6 = V[4], V[5], V[6]
    if q == Q_BENERGY:
        return 0.5 * (b4 * b4 + b5 * b5 + b6 * b6)
    if q == Q_BETA:
        return pr / (0.5 * (b4 * b4 + b5 * b5 + b6 * b6))
    if q == Q_TEMP:
        return pr / rho
    if q == Q_ONEOTCOOL:
        lam = np.vectorize(cooling, otypes=[np.float64])(rho, pr)
        return 2.0 * ((rho * rho * lam) / (1.5 * pr))
    if q == Q_VDOTB:
        return V[1] * b4 + V[2] * b5 + V[3] * b6
    if q == Q_DIVB:
        g1, g2, g3, g5 = (geo[n][:, None, :] for n in range(4))
        out = np.zeros_like(rho)
        c = (slice(1, -1),) * 3
        brp = 0.5 * (b4[1:-1, 1:-1, 1:-1] + b4[2:, 1:-1, 1:-1])
        brm = 0.5 * (b4[1:-1, 1:-1, 1:-1] + b4[:-2, 1:-1, 1:-1])
        bpp = 0.5 * (b5[1:-1, 1:-1, 1:-1] + b5[1:-1, 2:, 1:-1])
        bpm = 0.5 * (b5[1:-1, 1:-1, 1:-1] + b5[1:-1, :-2, 1:-1])
        bzp = 0.5 * (b6[1:-1, 1:-1, 1:-1] + b6[1:-1, 1:-1, 2:])
        bzm = 0.5 * (b6[1:-1, 1:-1, 1:-1] + b6[1:-1, 1:-1, :-2])
        G1, G2, G3, G5 = (t[1:-1, :, 1:-1] for t in (g1, g2, g3, g5))
        out[c] = (G2 * brp - G1 * brm) + (G3 * bpp - G3 * bpm) + (G5 * bzp - G5 * bzm)
        return out
    if q == Q_RHOBZ:
        return rho * b6
    if q == Q_SPIRALPOT:
        sp = spiral
        return (sp["eps"] * 10.0 * sp["q"][:, None, :] * sp["szs"][None, None, :] * sp["x2s2o15"][:, None, None]
                * sp["cosphase"][:, :, None]) * np.ones_like(rho)
    if q in (Q_VPARA, Q_VPERP):
        s = 1.0 / np.sqrt(1.0 + spiral["cot"] * spiral["cot"])
        c = spiral["cot"] * s
        return V[2] * c - V[1] * s if q == Q_VPARA else V[1] * c + V[2] * s
    raise ValueError(q)


def plane(Q, kind, a=0, b=0, dzm=None):
    """The values of one plane of Q (NX, NY, NZ) in file order."""
    NX, NY, NZ = Q.shape
    nj, nk = NY - 2, NZ - 2
    if kind == PLANE_RPHI:
        return Q[1:-1, 1:-1, a].reshape(-1).copy()
    if kind in (PLANE_RPHI_ZMEAN, PLANE_RPHI_ZINT):
        s = np.zeros((NX - 2, nj))
        for k in range(1, nk + 1):
            s = s + (Q[1:-1, 1:-1, k] * dzm[k - 1] if kind == PLANE_RPHI_ZINT else Q[1:-1, 1:-1, k])
        return (s if kind == PLANE_RPHI_ZINT else s / float(nk)).reshape(-1)
    if kind == PLANE_RZ:
        return Q[1:-1, a, 1:-1].T.reshape(-1).copy()
    if kind == PLANE_ZR:
        return Q[1:-1, a, 1:-1].reshape(-1).copy()
    if kind == PLANE_XZ:
        jo = 1 + (a - 1 + nj // 2) % nj
        rows = [np.concatenate([Q[::-1, jo, k], Q[:, a, k]]) for k in range(1, nk + 1)]
        return np.concatenate(rows)
    if kind == PLANE_XZ_PHIMEAN:
        s = np.zeros((NX, NZ))
        for j in range(1, nj + 1):
            s = s + Q[:, j, :]
        m = s / float(nj)
        return np.concatenate([np.concatenate([m[::-1, k], m[:, k]]) for k in range(1, nk + 1)])
    if kind == PLANE_PHI:
        return Q[a, 1:-1, b].copy()
    raise ValueError(kind)


def text(values):
    """The bytes of a text slice: "%e\\n" per value, as C's printf writes it."""
    return "".join("%e\n" % v for v in values).encode()


def write_text(path, values):
    with open(path, "wb") as f:
        f.write(text(values))


# The files cylmhd_idl_output writes (name, quantity, plane, which index):
# each name is the one an IDL script opens (density.pro:53 ... vz-xz.pro:55),
# each plane the shape of the fltarr it reads into.
IDL_FILES = [
    ("zplanerho", 0, PLANE_RPHI, "k"), ("zplanebeta", Q_BETA, PLANE_RPHI, "k"),
    ("zplanebenergy", Q_BENERGY, PLANE_RPHI, "k"), ("z0bphi", 5, PLANE_RPHI, "k1"),
    ("z100bphi", 5, PLANE_RPHI, "k100"), ("zplanemeanbr", 4, PLANE_RPHI_ZMEAN, None),
    ("divb0", Q_DIVB, PLANE_RPHI, "k"), ("faradaydepth-rphi", Q_RHOBZ, PLANE_RPHI_ZINT, None),
    ("vdotb", Q_VDOTB, PLANE_RPHI, "k"), ("temp-rphi", Q_TEMP, PLANE_RPHI, "k"),
    ("rho-rz", 0, PLANE_ZR, "j"), ("vphi-rz", 2, PLANE_ZR, "j"), ("bphi-rz", 5, PLANE_ZR, "j"),
    ("temp-rz", Q_TEMP, PLANE_ZR, "j"),
    ("rho-xz", 0, PLANE_XZ, "j"), ("xzbeta", Q_BETA, PLANE_XZ, "j"),
    ("xzmeanbphi", 5, PLANE_XZ_PHIMEAN, None), ("divb-xz", Q_DIVB, PLANE_XZ, "j"),
    ("temp-xz", Q_TEMP, PLANE_XZ, "j"), ("xzvz", 3, PLANE_XZ, "j"),
    ("oneocoolingtime-xz", Q_ONEOTCOOL, PLANE_XZ, "j"),
]


def idl_files(V, x, y, z, geo, dzm, cooling, nout, kplane=1, k100=100, jplane=1, spiral=None,
              azim_i=(0, 0, 0, 0), azim_k=0):
    """{relative path: bytes} of the files cylmhd_idl_output writes."""
    NX, NY, NZ = V.shape[1:]
    k100 = min(k100, NZ - 2)
    idx = {"k": kplane, "k1": 1, "k100": k100, "j": jplane, None: 0}
    d = "idl%d" % nout
    xs = np.concatenate([-x[::-1], x])
    out = {}
    for dd in sorted({"idl0", d}):
        out[os.path.join(dd, "r.txt")] = text(x[1:-1])
        out[os.path.join(dd, "phi.txt")] = text(y[1:-1])
        out[os.path.join(dd, "z.txt")] = text(z[1:-1])
        out[os.path.join(dd, "x.txt")] = text(xs)
    cache = {}
    for name, q, kind, which in IDL_FILES:
        if q not in cache:
            cache[q] = quantity(V, q, geo, cooling)
        vals = plane(cache[q], kind, idx[which], 0, dzm)
        out[os.path.join(d, "%s-%d.dat" % (name, nout))] = text(vals)
    # azimuthal profiles at (azim_i[K], azim_k): phiout0..3.pro, fourieramp.pro
    ak = azim_k or 1
    for K in range(4):
        ai = azim_i[K] or 1 + (K + 1) * (NX - 2) // 5
        ad = "azimuth-%d" % K
        out[os.path.join(ad, "phi.dat")] = text(y[1:-1])
        for name, q in (("potential", Q_SPIRALPOT), ("rho", 0), ("te", Q_TEMP), ("vperp", Q_VPERP),
                        ("vpara", Q_VPARA), ("velocitypara", Q_VPARA)):
            Q = quantity(V, q, geo, cooling, spiral)
            out[os.path.join(ad, "%s-%d.dat" % (name, nout))] = text(plane(Q, PLANE_PHI, ai, ak))
    return out


def gdl_files(V, x, y, z, nout):
    """{name: bytes} of GDLOutputCylinder (output.h:280-384)."""
    be = quantity(V, Q_BENERGY)
    return {
        "r.dat": text(x[1:-1]), "phi.dat": text(y[1:-1]), "z.dat": text(z[1:-1]),
        "2d-density-rphi-%d.dat" % nout: text(plane(V[0], PLANE_RPHI, 1)),
        "2d-pressure-rphi-%d.dat" % nout: text(plane(V[7], PLANE_RPHI, 1)),
        "2d-benergy-rphi-%d.dat" % nout: text(plane(be, PLANE_RPHI, 1)),
        "2d-density-rz-%d.dat" % nout: text(plane(V[0], PLANE_RZ, 0)),
    }
