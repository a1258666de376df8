"""The oracle against the reference itself (CPU, no GPU).

oracle/_ref holds the reference's own routines compiled from /root/reference
(oracle/build_ref.py); these tests run them and the oracle on identical
inputs and require bit-identical results: state U and V after several RK
steps, every face flux of the last stage, dt/ch per step, the "dt factor
mesh" cell and the Check floor counters (single-threaded reference: its
OpenMP counters race).  The golden fixtures in tests/golden carry the same
comparison to machines without the reference build.
"""
import os

import numpy as np
import pytest

import cylmhd as cm
from cases import break_case, rand_case, ref_disc_case
from oracle_py import Oracle, parse_counts, parse_dt_mesh, ref_harness, run_ref

SCHEMES = {"mm_hlld": (cm.LIMITER_MINMOD, cm.SOLVER_HLLD), "mc_hlld": (cm.LIMITER_MC, cm.SOLVER_HLLD),
           "mm_hll": (cm.LIMITER_MINMOD, cm.SOLVER_HLL), "mc_hll": (cm.LIMITER_MC, cm.SOLVER_HLL),
           "vl_hlld": (cm.LIMITER_VANLEER, cm.SOLVER_HLLD), "sb_hlld": (cm.LIMITER_SUPERBEE, cm.SOLVER_HLLD)}


def run_oracle(p, Ui, Vi, Uc, Vc, nsteps):
    o = Oracle(p)
    o.set_initial(Ui, Vi)
    o.set_state(Uc, Vc)
    t, hist = o.run_steps(nsteps)
    U, V = o.state()
    return o, U, V, hist


def compare(name, p, Ui, Vi, Uc, Vc, nsteps, galactic):
    if ref_harness(name) is None:
        pytest.skip(f"reference harness {name} not built (needs /root/reference)")
    o, U, V, hist = run_oracle(p, Ui, Vi, Uc, Vc, nsteps)
    gx, gz = o.gravity_axi()
    r, out = run_ref(name, Vi, Vc, gx, gz, nsteps, p.rk, galactic=galactic, U_init=Ui, U_cur=Uc)
    # both stop at the same step when dt falls under dtmin (main.c:331-334)
    assert len(r["dt"]) == len(hist) >= 1
    nsteps = len(hist)
    np.testing.assert_array_equal(r["dt"], [h[0] for h in hist])
    np.testing.assert_array_equal(r["ch"], [h[1] for h in hist])
    assert parse_dt_mesh(out)[:nsteps] == [h[2] for h in hist]
    np.testing.assert_array_equal(r["U"], U)
    np.testing.assert_array_equal(r["V"], V)
    fl = o.fluxes()
    np.testing.assert_array_equal(r["F"], fl[0])
    np.testing.assert_array_equal(r["G"], fl[1])
    np.testing.assert_array_equal(r["H"], fl[2])
    nst = 3 if p.rk == 3 else 2
    np.testing.assert_array_equal(parse_counts(out)[-nst:], o.floor_counts()[:nst])
    return o, r, out


@pytest.mark.parametrize("rk", [3, 2])
@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_disc_matches_reference(scheme, rk):
    lim, sol = SCHEMES[scheme]
    p, Ui, Vi, Uc, Vc = ref_disc_case(1, 16, 16, 10, limiter=lim, solver=sol, rk=rk, epssp0=0.0)
    compare(f"disc16_{scheme}", p, Ui, Vi, Uc, Vc, 4, galactic=False)


@pytest.mark.parametrize("rk", [3, 2])
@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_random_state_matches_reference(scheme, rk):
    lim, sol = SCHEMES[scheme]
    p, Ui, Vi, Uc, Vc = rand_case(limiter=lim, solver=sol, rk=rk)
    o, r, out = compare(f"rand16_{scheme}", p, Ui, Vi, Uc, Vc, 3, galactic=True)
    counts = parse_counts(out)
    assert (counts.sum(axis=0) > 0).all(), counts    # all three floors fired


SOURCES = {"axic": cm.SOURCE_AXI_COOLING, "axi": cm.SOURCE_AXI, "ugz": cm.SOURCE_UNIFORM_GZ,
           "none": cm.SOURCE_NONE}


@pytest.mark.parametrize("src", list(SOURCES))
def test_source_variants_match_reference(src):
    """The reference's other source functions (source.h:143-349), selected in
    the harness the way its author swaps the SpiralGravityCooling call: the
    random state with a patch in the Raymond cooling band, so the two
    axisymmetric forms differ by their cooling term."""
    p, Ui, Vi, Uc, Vc = rand_case(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD, cooling_band=True,
                                  source=SOURCES[src])
    compare(f"rand16_mc_hlld_{src}", p, Ui, Vi, Uc, Vc, 3, galactic=True)


@pytest.mark.parametrize("src", list(SOURCES))
def test_source_variants_differ(src):
    """Each variant changes the state against SpiralGravityCooling with a 5 %
    spiral at full amplitude after the first step (so the selection reaches
    the stage update) and against every other variant."""
    outs = {}
    for name, mode in [("spiral", cm.SOURCE_SPIRAL_COOLING)] + list(SOURCES.items()):
        p, Ui, Vi, Uc, Vc = rand_case(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD, cooling_band=True, source=mode,
                                      epssp0=0.05, tausp=1.0e-30)
        outs[name] = run_oracle(p, Ui, Vi, Uc, Vc, 2)[1]     # the spiral ramp is 0 at t = 0
    for other, U in outs.items():
        if other != src:
            assert not np.array_equal(outs[src], U), (src, other)


@pytest.mark.parametrize("scheme", ["mm_hlld", "mc_hlld"])
def test_cooling_band_matches_reference(scheme):
    lim, sol = SCHEMES[scheme]
    p, Ui, Vi, Uc, Vc = rand_case(limiter=lim, solver=sol, cooling_band=True)
    compare(f"rand16_{scheme}", p, Ui, Vi, Uc, Vc, 3, galactic=True)


def test_exact_degenerate_hlld_is_nan_in_both():
    """Equal states with a purely normal field give D == 0 in the HLLD
    degeneracy switch, and (1/D + 1) + (1/D - 1)*sgn = inf - inf = NaN
    (hlldx.h:227-230).  The oracle reproduces the reference NaN for NaN."""
    name = "rand16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = rand_case(degenerate="exact")
    o, U, V, hist = run_oracle(p, Ui, Vi, Uc, Vc, 1)
    gx, gz = o.gravity_axi()
    r, out = run_ref(name, Vi, Vc, gx, gz, 1, p.rk, galactic=True, U_init=Ui, U_cur=Uc)
    assert np.isnan(r["U"]).any()
    np.testing.assert_array_equal(np.isnan(r["U"]), np.isnan(U))
    np.testing.assert_array_equal(r["U"], U)


def test_cooling_break_temperature_matches_reference():
    """A cell exactly at Cooling's break temperature a = 0.031: the sign
    factors are 0/0 = NaN and Cooling returns NaN through int, which the
    compiled reference (x86-64) turns into INT_MIN -- a huge heating term in
    SpiralGravityCooling's energy source.  The oracle gives the same bits."""
    name = "rand16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = break_case()
    assert cm.lib() and Vc[7][6, 3, 3] / Vc[0][6, 3, 3] == 0.031
    o, U, V, hist = run_oracle(p, Ui, Vi, Uc, Vc, 1)
    gx, gz = o.gravity_axi()
    r, out = run_ref(name, Vi, Vc, gx, gz, 1, p.rk, galactic=True, U_init=Ui, U_cur=Uc)
    np.testing.assert_array_equal(r["U"], U)
    np.testing.assert_array_equal(r["V"], V)
    # the break path was taken: the energy of that cell differs from the same
    # step with the temperature one ulp away from the break point
    q, Ui2, Vi2, Uc2, Vc2 = break_case(temperature=np.nextafter(0.031, 1.0))
    _, U2, _, _ = run_oracle(q, Ui2, Vi2, Uc2, Vc2, 1)
    e, e2 = U[7][6, 3, 3], U2[7][6, 3, 3]
    assert np.isfinite(e2) and (not np.isfinite(e) or abs(e) > 1e3 * abs(e2)), (e, e2)


def test_geometry_and_tables_match_reference():
    name = "disc16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = ref_disc_case(1, 16, 16, 10, epssp0=0.0)
    o = Oracle(p)
    gx, gz = o.gravity_axi()
    r, _ = run_ref(name, Vi, Vc, gx, gz, 0, 3)
    g = o.grid()
    for key in ("x", "xm", "y", "z"):
        np.testing.assert_array_equal(r[key], g[key])
    assert r["minminlength"] == g["minminlength"]
    geo = o.geometry()
    np.testing.assert_array_equal(r["geom"], geo)
    xabs, zabs, cabs, irin = o.absorb_factors()
    np.testing.assert_array_equal(r["xabs"], xabs)
    np.testing.assert_array_equal(r["cabs"][: cabs.size].reshape(cabs.shape), cabs)
    # the product's own host grid is the same grid
    x, y, z = cm.grid(p)
    np.testing.assert_array_equal(x, g["x"])
    np.testing.assert_array_equal(y, g["y"])
    np.testing.assert_array_equal(z, g["z"])


def test_galactic_gravity_matches_initialgravity():
    """InitialGravity (initialgravity.h:22-79) compiled from the reference vs
    the oracle's and the product's tables."""
    name = "rand16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = rand_case()
    o = Oracle(p)
    gx, gz = o.gravity_axi()
    r, _ = run_ref(name, Vi, Vc, gx * 0, gz * 0, 0, 3, galactic=True)
    np.testing.assert_array_equal(r["grx"], gx)
    np.testing.assert_array_equal(r["grz"], gz)


def test_binary_dump_is_byte_identical(tmp_path):
    """BinaryOutput (output.h:1202-1414) of the reference vs
    cylmhd_binary_output on the same state."""
    name = "disc16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = ref_disc_case(1, 16, 16, 10, epssp0=0.0)
    o, U, V, hist = run_oracle(p, Ui, Vi, Uc, Vc, 2)
    gx, gz = o.gravity_axi()
    refdir = tmp_path / "ref"
    refdir.mkdir()
    run_ref(name, Vi, Vc, gx, gz, 2, 3, dumpdir=str(refdir), U_init=Ui, U_cur=Uc)
    t = sum(h[0] for h in hist)
    x, y, z = cm.grid(p)
    mine = tmp_path / "mine"
    cm.binary_output(str(mine), t, x, y, z, V)
    files = sorted(os.listdir(refdir / "0"))
    assert files == sorted(os.listdir(mine))
    for f in files:
        assert (refdir / "0" / f).read_bytes() == (mine / f).read_bytes(), f
    # RestartRead (restartread.h:9-143) reads the nine primitive files back
    V2 = cm.restart_read(str(mine), V.shape[1:])
    np.testing.assert_array_equal(V2, V)


def test_galactic_disc3_matches_reference_model():
    """GalacticDisc3 (model.h:149-327) compiled from the reference vs the
    host restatement (cylmhd_model_galactic_disc3): primitives bit for bit,
    U = VtoU(V) with the k = 0, 1 planes of its z boundary pair
    (ZLeftVzAntisymBSymBoundary then ZLeftFreeBoundary), and its gravity
    equal to InitialGravity's (the oracle's tables)."""
    name = "rand16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = rand_case()
    V = cm.model_galactic_disc3(p)
    r, _ = run_ref(name, Vi, Vc, np.zeros(Vi.shape[1::2]), np.zeros(Vi.shape[1::2]), 0, 3, model=True)
    np.testing.assert_array_equal(r["V"], V)
    U = cm.vtou(p, V)
    np.testing.assert_array_equal(r["U"][..., 2:], U[..., 2:])
    np.testing.assert_array_equal(r["U"][3, :, :, 1], 0.0)
    keep = [m for m in range(9) if m != 3]
    np.testing.assert_array_equal(r["U"][keep][..., 1], U[keep][..., 1])
    np.testing.assert_array_equal(r["U"][..., 0], r["U"][..., 1])
    o = Oracle(p)
    gx, gz = o.gravity_axi()
    o.close()
    np.testing.assert_array_equal(r["grx"], gx)
    np.testing.assert_array_equal(r["grz"], gz)
