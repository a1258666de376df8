"""Precision report: the FP32-state path against an FP64 path on one disc.

Both paths advance the same initial state with the SAME dt/ch sequence
(taken from the FP64 run, as main.c:324-348 would compute it), so the
difference measured is that of the arithmetic, not of the time steps.
Reported per conserved variable: relative L1 and L-infinity error over the
interior, and over the cells outside the absorbing layers (the evolved
solution), with the components that vanish in the initial disc (rho vr,
rho vz, Br, Bz, psi) measured against the vector they belong to (rho v,
B); the plain per-component L1 (every component against its own sum);
plus L1 div B of both runs and the floor hits of the FP32 run.
"""
import numpy as np

import cylmhd as cm

NAMES = ["rho", "rho_vr", "r_rho_vphi", "rho_vz", "br", "bphi", "bz", "e", "psi"]


def _scales(b, x, red):
    b = b[:, 1:-1, 1:-1, 1:-1]
    x = x[1:-1]
    mom = np.sqrt(b[1] ** 2 + (b[2] / x[:, None, None]) ** 2 + b[3] ** 2)
    bmag = np.sqrt(b[4] ** 2 + b[5] ** 2 + b[6] ** 2)
    own = np.array([red(np.abs(b[m])) for m in range(9)])
    vec = {1: red(mom), 3: red(mom), 4: red(bmag), 6: red(bmag), 8: red(bmag)}
    return np.array([max(own[m], vec.get(m, 0.0)) for m in range(9)])


def errors(a, b, x, ilo=1, ihi=None):
    """(L1, Linf) relative error per component of a against b over the
    interior cells with ilo <= i <= ihi (default: all interior cells)."""
    ihi = a.shape[1] - 2 if ihi is None else ihi
    sl = (slice(None), slice(ilo, ihi + 1), slice(1, -1), slice(1, -1))
    d = np.abs(a[sl] - b[sl])
    bb = b[:, ilo - 1:ihi + 2]           # _scales strips one ghost layer
    xx = x[ilo - 1:ihi + 2]
    l1 = np.array([d[m].sum() for m in range(9)]) / _scales(bb, xx, np.sum)
    linf = np.array([d[m].max() for m in range(9)]) / _scales(bb, xx, np.max)
    return l1, linf


def physical_range(p):
    """i-range outside the absorbing layers: the central region i <= irin,
    irin = (int)(rin/dx0) (input.h:72, centralboundary.h:27-52), and the
    outer 11 r-cells (xboundary.h:322-346) are relaxed towards the initial
    state every step, so they are not part of the evolved solution."""
    irin = int(p.rin / p.dx0)
    return irin + 1, p.ix + 2 - 12


def compare(p, steps, ref_path=cm.PATH_PLAIN64, test_path=cm.PATH_FUSED32, log=None):
    """Advance p's disc `steps` steps on both paths; return the report dict."""
    U0, V0 = cm.disc_state(p)
    x = cm.grid(p)[0]
    q = p.copy()
    q.path = ref_path
    hist = []
    with cm.Context(q) as g:
        g.set_initial(U0, V0)
        t = 0.0
        for n in range(steps):
            dt, ch, _ = g.setdt()
            if not np.isfinite(dt) or dt < p.dtmin:
                raise cm.CylmhdError(f"dt collapsed at step {n} (dt={dt})")
            g.gravity(t)
            g.step(dt, ch)
            hist.append((t, dt, ch))
            t += dt
        d64 = g.diagnostics()
        U64, _ = g.download()
    if log:
        log(f"[precision] {cm.PATH_NAMES[ref_path]}: {steps} steps to t={t:.4e}")
    q = p.copy()
    q.path = test_path
    floors = np.zeros((3, 3), dtype=np.int64)
    with cm.Context(q) as g:
        g.set_initial(U0, V0)
        del U0, V0
        for (tn, dt, ch) in hist:
            g.gravity(tn)
            g.step(dt, ch)
            floors += g.floor_counts()
        d32 = g.diagnostics()
        U32, _ = g.download()
    l1, linf = errors(U32, U64, x)
    sl = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    l1_plain = np.abs(U32[sl] - U64[sl]).sum(axis=(1, 2, 3)) / np.maximum(np.abs(U64[sl]).sum(axis=(1, 2, 3)), 1e-300)
    ilo, ihi = physical_range(p)
    pl1, plinf = errors(U32, U64, x, ilo, ihi)
    del U32, U64
    return dict(
        steps=steps, t=t, grid=[p.ix, p.jy, p.kz],
        ref=cm.PATH_NAMES[ref_path], test=cm.PATH_NAMES[test_path],
        l1={k: float(v) for k, v in zip(NAMES, l1)},
        linf={k: float(v) for k, v in zip(NAMES, linf)},
        l1_max=float(l1.max()),
        l1_plain={k: float(v) for k, v in zip(NAMES, l1_plain)},
        physical={"i_range": [ilo, ihi],
                  "l1": {k: float(v) for k, v in zip(NAMES, pl1)},
                  "linf": {k: float(v) for k, v in zip(NAMES, plinf)},
                  "l1_max": float(pl1.max())},
        divb_l1={"ref": d64["divb_l1"], "test": d32["divb_l1"]},
        mass_rel=abs(d32["mass"] - d64["mass"]) / abs(d64["mass"]),
        energy_rel=abs(d32["energy"] - d64["energy"]) / abs(d64["energy"]),
        floor_hits={"rho": int(floors[:, 0].sum()), "pressure": int(floors[:, 1].sum()),
                    "beta": int(floors[:, 2].sum())},
    )
