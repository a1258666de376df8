"""Shared test cases: parameter sets and seeded initial states.

Two families:
  * disc  -- the benchmark discs (cylmhd_params_config) at small sizes;
  * rand  -- a random MHD state on the same geometry, in the reference's
             rotating frame with its galactic potential, built to reach every
             branch: supersonic faces (HLLD sl>=0 / sr<=0), all Alfven-fan
             branches, the degenerate-HLLD switch (purely normal field with
             equal states), the Raymond cooling band, and all three Check
             floors (density via a larger initial density, pressure via
             cold cells, beta via strong-field cells).
"""
import numpy as np

import cylmhd as cm

REF_OMEGASP = 10.0e+5 / (1.0 / 4.7058e+15) * (1.0 / 3.08568e+21)


# The reference's compile-time floors and its (only) cooling law
# (input.h:78-80, cooling.h:22-78): what the reference harness is built with.
REF_FIXED = dict(rhocr=1.0e-4, prcr=0.015, betacr=0.1, cooling=cm.COOL_RAYMOND)


def ref_disc_case(config=1, nr=16, nphi=16, nz=10, **over):
    """A benchmark disc run under the reference's own floors and cooling,
    for comparisons with the compiled reference."""
    return disc_case(config, nr, nphi, nz, **{**REF_FIXED, **over})


def disc_params(config=1, nr=16, nphi=16, nz=10, **over):
    p = cm.params_config(config, nr, nphi, nz)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def disc_case(config=1, nr=16, nphi=16, nz=10, **over):
    p = disc_params(config, nr, nphi, nz, **over)
    U, V = cm.disc_state(p)
    return p, U, V, U.copy(), V.copy()


def rand_params(nr=16, nphi=8, nz=10, **over):
    p = cm.params_config(1, nr, nphi, nz)
    p.omegasp = REF_OMEGASP
    p.omegapattern = REF_OMEGASP
    p.gravity = cm.GRAV_GALACTIC
    p.epssp0 = 0.0
    p.tausp = 200.0 * 1.0e+6 * 3.1536e+7 / 4.7058e+15
    for k, v in REF_FIXED.items():
        setattr(p, k, v)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def rand_V(p, seed=7, consistent=False, degenerate="near", cooling_band=False):
    """degenerate: "near" -- the equal-state patches carry a 1e-4 tangential
    field, so the HLLD switch takes its degenerate side with D small but
    nonzero; "exact" -- purely normal field, D == 0 exactly, where the
    reference's branchless switch (hlldx.h:222-234) forms inf - inf = NaN.
    cooling_band: a patch at T = 0.045, inside the band where the Raymond fit
    is nonzero (cooling.h:33-57); its cooling rate then sets dt and, through
    ch = minminlength/dt (setdt.h:89), a very large GLM speed."""
    NX, NY, NZ = p.dims
    eps = 0.0 if degenerate == "exact" else 1.0e-4
    rng = np.random.default_rng(seed)
    sh = (NX, NY, NZ)
    V = np.zeros((9,) + sh)
    V[0] = rng.uniform(0.5, 2.0, sh)
    V[7] = V[0] * rng.uniform(0.005, 0.028, sh)
    if cooling_band:
        V[7, 2:4, 6:8, 7:9] = 0.045 * V[0, 2:4, 6:8, 7:9]
    for m in (1, 2, 3):
        V[m] = 1.0 * rng.uniform(-1.0, 1.0, sh)
    for m in (4, 5, 6):
        V[m] = rng.uniform(-0.2, 0.2, sh)
    V[8] = rng.uniform(-0.1, 0.1, sh)
    # uniform strong-field block (low beta -> beta floor), all components
    V[4:7, 2:9, 1:6, 2:8] = 0.6
    # degenerate HLLD patch: equal states, purely radial strong field
    V[:, 8:11, 4:6, 4:7] = np.array([1.0, 0.1, 0.2, 0.0, 0.5, eps, eps, 0.03, 0.0])[:, None, None, None]
    # same with a vertical field (z faces) and an azimuthal one (phi faces)
    V[:, 4:7, 5:7, 2:4] = np.array([1.0, 0.0, 0.1, 0.2, eps, eps, 0.5, 0.03, 0.0])[:, None, None, None]
    V[:, 10:13, 1:3, 6:8] = np.array([1.0, 0.1, 0.0, 0.0, eps, 0.5, eps, 0.03, 0.0])[:, None, None, None]
    if consistent:
        U = cm.vtou(p, V)
        cm.fill_ghosts(p, U, V)
    return V


def rand_case(seed=7, consistent=False, degenerate="near", cooling_band=False, **over):
    """(p, U_init, V_init, U_cur, V_cur): the initial state used for
    SaveInitial differs from the current one in a few cells whose density it
    raises 1e5-fold, so the density floor fires there."""
    p = rand_params(**over)
    Vc = rand_V(p, seed, consistent, degenerate, cooling_band)
    Vi = Vc.copy()
    Vi[0, 5:8, 3:5, 5:8] *= 1.0e5
    Ui = cm.vtou(p, Vi)
    Uc = cm.vtou(p, Vc)
    if consistent:
        cm.fill_ghosts(p, Ui, Vi)
        cm.fill_ghosts(p, Uc, Vc)
    return p, Ui, Vi, Uc, Vc


def break_case(cell=(6, 3, 3), temperature=0.031, **over):
    """The random state (Raymond cooling on) with one cell at exactly a
    break temperature of Cooling (cooling.h:31: a = 0.031): rho = 1 and
    pr = 0.031, so T - a = 0 and sign = 0/0 = NaN (cooling.h:45); Cooling
    returns that NaN through int (the x86-64 conversion gives INT_MIN)."""
    p, Ui, Vi, Uc, Vc = rand_case(**over)
    Vc[0][cell] = 1.0
    Vc[7][cell] = temperature
    Uc = cm.vtou(p, Vc)
    return p, Ui, Vi, Uc, Vc


def rel_err(a, b):
    """max |a-b| / max |b| per component (L-infinity relative)."""
    out = []
    for m in range(a.shape[0]):
        den = np.max(np.abs(b[m]))
        num = np.max(np.abs(a[m] - b[m]))
        out.append(num / den if den > 0 else num)
    return np.array(out)


def group_scales(b, x, l1=False, interior=True):
    """Per-component error scale of a conserved state b[9][NX][NY][NZ].

    Components that vanish in the initial disc (rho vr, rho vz, Br, Bz, psi)
    hold only what one step creates, so an error relative to their own
    magnitude measures rounding against a near-zero quantity.  Each
    component is therefore measured against the larger of its own size and
    the size of the vector it belongs to: rho v (with U2 / r for the phi
    part) for 1-3, B for 4-6 and for psi (the field it cleans).  Density,
    angular momentum, B_phi and energy keep their own scale.
    With l1=True the scales are sums over the (interior) cells, else maxima."""
    if interior:
        sl = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        b = b[sl]
        x = x[1:-1]
    mom = np.sqrt(b[1] ** 2 + (b[2] / x[:, None, None]) ** 2 + b[3] ** 2)
    bmag = np.sqrt(b[4] ** 2 + b[5] ** 2 + b[6] ** 2)
    red = np.sum if l1 else np.max
    own = np.array([red(np.abs(b[m])) for m in range(9)])
    vec = {1: red(mom), 3: red(mom), 4: red(bmag), 6: red(bmag), 8: red(bmag)}
    return np.array([max(own[m], vec.get(m, 0.0)) for m in range(9)])


def rel_err_scaled(a, b, x):
    """max |a-b| per component over all cells / group_scales (max)."""
    num = np.array([np.max(np.abs(a[m] - b[m])) for m in range(9)])
    return num / group_scales(b, x, l1=False, interior=False)


def l1_rel_scaled(a, b, x):
    """sum |a-b| per component over interior cells / group_scales (sum)."""
    sl = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    num = np.array([np.sum(np.abs(a[sl][m] - b[sl][m])) for m in range(9)])
    return num / group_scales(b, x, l1=True, interior=True)


def l1_rel(a, b, interior=True):
    """sum |a-b| / sum |b| per component over interior cells."""
    sl = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1)) if interior else (slice(None),)
    a, b = a[sl], b[sl]
    out = []
    for m in range(a.shape[0]):
        den = np.sum(np.abs(b[m]))
        num = np.sum(np.abs(a[m] - b[m]))
        out.append(num / den if den > 0 else num)
    return np.array(out)
