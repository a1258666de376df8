"""GPU: the reference's edge cases on the HIP path, against the oracle.

* Exact-degenerate HLLD: equal states with a purely normal field give D = 0
  in the branchless degeneracy switch and (1/D + 1) + (1/D - 1) sgn =
  inf - inf = NaN (hlldx.h:222-234).
* A cell exactly at a break temperature of Cooling (cooling.h:45-48): the
  sign factor is 0/0 = NaN, returned through int (INT_MIN on the compiled
  reference, x86-64).
* The default source, SpiralGravityCooling (source.h:26-134), with the
  Raymond cooling band live (cooling.h:33-57).
The oracle is pinned to the compiled reference on the same cases
(test_oracle_ref).  Bar: the NaN masks are equal and the finite values agree
to 1e-12 relative per conserved variable.
"""
import numpy as np
import pytest

import cylmhd as cm
from cases import break_case, rand_case, rel_err
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

TOL = 1.0e-12


def step_both(p, Ui, Vi, Uc, Vc, nsteps, path):
    q = p.copy()
    q.path = path
    o = Oracle(q)
    o.set_initial(Ui, Vi)
    o.set_state(Uc, Vc)
    out = []
    with cm.Context(q) as g:
        g.set_initial(Ui, Vi)
        g.upload(Uc, Vc)
        t = 0.0
        for _ in range(nsteps):
            dto, cho, _ = o.setdt()
            dtg, chg, _ = g.setdt()
            o.gravity(t)
            g.gravity(t)
            o.step(dto, cho)
            g.step(dto, cho)
            t += dto
            Ug, Vg = g.download()
            Uo, Vo = o.state()
            out.append((Ug, Uo, dtg, dto, g.floor_counts(), o.floor_counts()))
    o.close()
    return out


def nan_aware_err(a, b):
    """NaN masks must be equal; relative max error of the finite values."""
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    err = []
    for m in range(a.shape[0]):
        f = np.isfinite(b[m])
        np.testing.assert_array_equal(np.isfinite(a[m]), f)
        den = np.max(np.abs(b[m][f])) if f.any() else 0.0
        num = np.max(np.abs(a[m][f] - b[m][f])) if f.any() else 0.0
        err.append(num / den if den > 0 else num)
    return np.array(err)


def test_exact_degenerate_hlld_nan_pattern_plain64(gpu):
    p, Ui, Vi, Uc, Vc = rand_case(degenerate="exact")
    Ug, Uo, _, _, _, _ = step_both(p, Ui, Vi, Uc, Vc, 1, cm.PATH_PLAIN64)[0]
    assert np.isnan(Uo).any()
    e = nan_aware_err(Ug, Uo)
    print("NaN cells", int(np.isnan(Uo).sum()), "finite err", e)
    assert e.max() <= TOL, e


def test_cooling_break_temperature_plain64(gpu):
    p, Ui, Vi, Uc, Vc = break_case()
    Ug, Uo, _, _, _, _ = step_both(p, Ui, Vi, Uc, Vc, 1, cm.PATH_PLAIN64)[0]
    e = nan_aware_err(Ug, Uo)
    print("non-finite cells", int((~np.isfinite(Uo)).sum()), "finite err", e)
    assert e.max() <= TOL, e
    # the break path was taken (the oracle's own check of it is in test_oracle_ref)
    assert not np.isfinite(Uo[7][6, 3, 3]) or abs(Uo[7][6, 3, 3]) > 1e3


@pytest.mark.parametrize("path,nsteps", [(cm.PATH_PLAIN64, 3), (cm.PATH_FUSED64, 1)], ids=["plain64", "fused64"])
def test_raymond_band_default_source(gpu, path, nsteps):
    """SpiralGravityCooling with a patch in the Raymond band: its cooling
    rate sets dt (setdt.h:56-58) and, through ch, the GLM speed."""
    p, Ui, Vi, Uc, Vc = rand_case(cooling_band=True, consistent=path == cm.PATH_FUSED64)
    assert p.source == cm.SOURCE_SPIRAL_COOLING and p.cooling == cm.COOL_RAYMOND
    for n, (Ug, Uo, dtg, dto, cg, co) in enumerate(step_both(p, Ui, Vi, Uc, Vc, nsteps, path)):
        e = rel_err(Ug, Uo)
        print(f"step {n}: eU={e.max():.2e} dt {dtg:.17g} {dto:.17g}")
        assert e.max() <= TOL, e
        assert abs(dtg - dto) <= 1e-13 * dto
        if path == cm.PATH_PLAIN64:
            np.testing.assert_array_equal(cg, co)
