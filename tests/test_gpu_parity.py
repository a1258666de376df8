"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bars (north star): the plain FP64 path matches the reference's FP64 path to
1e-12 relative per conserved variable after one RK step; the FP64 paths track
total mass, total energy and L1 div B to 1e-10 relative over 100 steps; the
FP32 path (FP64 recovery) stays within 1e-5 relative L1 of FP64.  The oracle
is bit-identical to the compiled reference on these cases (test_oracle_ref).
"""
import numpy as np
import pytest

import cylmhd as cm
from cases import disc_case, l1_rel_scaled, rand_case, rel_err, rel_err_scaled
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

TOL_STEP = 1.0e-12   # per conserved variable, after one RK step (FP64 paths)
TOL_INT = 1.0e-10    # totals over 100 steps (FP64 paths)


def run_pair(p, Ui, Vi, Uc, Vc, nsteps, path):
    """Run the oracle and the GPU path side by side, step by step, feeding
    both the oracle's dt/ch (so a dt difference cannot compound).  Returns
    per-step (rel_err_U, rel_err_V, dt_gpu, dt_orc, counts_gpu, counts_orc)."""
    q = p.copy()
    q.path = path
    o = Oracle(q)
    o.set_initial(Ui, Vi)
    o.set_state(Uc, Vc)
    g = cm.Context(q)
    g.set_initial(Ui, Vi)
    g.upload(Uc, Vc)
    hist = []
    t = 0.0
    for n in range(nsteps):
        dto, cho, ijko = o.setdt()
        dtg, chg, ijkg = g.setdt()
        o.gravity(t)
        g.gravity(t)
        o.step(dto, cho)
        g.step(dto, cho)
        t += dto
        Ug, Vg = g.download()
        Uo, Vo = o.state()
        hist.append(dict(eU=rel_err(Ug, Uo), eV=rel_err(Vg, Vo), dtg=dtg, dto=dto, chg=chg, cho=cho,
                         ijkg=ijkg, ijko=ijko, cg=g.floor_counts(), co=o.floor_counts(), Ug=Ug, Uo=Uo))
    g.close()
    o.close()
    return hist


@pytest.mark.parametrize("limiter", [cm.LIMITER_MINMOD, cm.LIMITER_MC, cm.LIMITER_VANLEER, cm.LIMITER_SUPERBEE])
@pytest.mark.parametrize("solver", [cm.SOLVER_HLLD, cm.SOLVER_HLL])
def test_plain64_disc(gpu, limiter, solver):
    p, Ui, Vi, Uc, Vc = disc_case(2, 16, 16, 10, limiter=limiter, solver=solver)
    hist = run_pair(p, Ui, Vi, Uc, Vc, 3, cm.PATH_PLAIN64)
    for n, h in enumerate(hist):
        print(f"step {n}: eU={h['eU'].max():.2e} eV={h['eV'].max():.2e} counts {h['cg'].ravel()} {h['co'].ravel()}")
        assert h["eU"].max() <= TOL_STEP, h["eU"]
        assert abs(h["dtg"] - h["dto"]) <= 1e-14 * h["dto"]
        assert h["ijkg"] == h["ijko"]
        np.testing.assert_array_equal(h["cg"], h["co"])


@pytest.mark.parametrize("rk", [3, 2])
@pytest.mark.parametrize("limiter,solver", [(cm.LIMITER_MINMOD, cm.SOLVER_HLLD), (cm.LIMITER_MC, cm.SOLVER_HLL),
                                             (cm.LIMITER_VANLEER, cm.SOLVER_HLLD), (cm.LIMITER_SUPERBEE, cm.SOLVER_HLLD)])
def test_plain64_random_state(gpu, rk, limiter, solver):
    p, Ui, Vi, Uc, Vc = rand_case(rk=rk, limiter=limiter, solver=solver)
    hist = run_pair(p, Ui, Vi, Uc, Vc, 3, cm.PATH_PLAIN64)
    for n, h in enumerate(hist):
        print(f"step {n}: eU={h['eU'].max():.2e} counts {h['cg'].ravel()} {h['co'].ravel()}")
        assert h["eU"].max() <= TOL_STEP, h["eU"]
        np.testing.assert_array_equal(h["cg"], h["co"])
    assert hist[0]["co"].sum() > 0          # the floors were exercised


@pytest.mark.parametrize("limiter", [cm.LIMITER_MINMOD, cm.LIMITER_MC])
def test_fused64_disc(gpu, limiter):
    """The disc starts with vr = vz = Br = Bz = psi = 0: those components are
    measured against the vector they belong to (tests/cases.group_scales)."""
    p, Ui, Vi, Uc, Vc = disc_case(2, 32, 32, 20, limiter=limiter)
    x = cm.grid(p)[0]
    hist = run_pair(p, Ui, Vi, Uc, Vc, 2, cm.PATH_FUSED64)
    for n, h in enumerate(hist):
        e = rel_err_scaled(h["Ug"], h["Uo"], x)
        print(f"step {n}: scaled eU={e} own-scale eU={h['eU']}")
        assert e.max() <= TOL_STEP, e
        assert abs(h["dtg"] - h["dto"]) <= 1e-12 * h["dto"]


@pytest.mark.parametrize("rk", [3, 2])
def test_fused64_random_state(gpu, rk):
    p, Ui, Vi, Uc, Vc = rand_case(rk=rk, consistent=True, nr=40, nphi=32, nz=34)
    hist = run_pair(p, Ui, Vi, Uc, Vc, 1, cm.PATH_FUSED64)
    h = hist[0]
    print(f"eU={h['eU']}")
    assert h["eU"].max() <= TOL_STEP, h["eU"]


def test_fused32_disc(gpu):
    p, Ui, Vi, Uc, Vc = disc_case(3, 32, 32, 20)
    x = cm.grid(p)[0]
    hist = run_pair(p, Ui, Vi, Uc, Vc, 5, cm.PATH_FUSED32)
    for n, h in enumerate(hist):
        e = l1_rel_scaled(h["Ug"], h["Uo"], x)
        print(f"step {n}: L1 rel {e}")
        assert e.max() <= 1.0e-5, e


SOURCES = [cm.SOURCE_AXI_COOLING, cm.SOURCE_AXI, cm.SOURCE_UNIFORM_GZ, cm.SOURCE_NONE]


@pytest.mark.parametrize("source", SOURCES)
def test_plain64_source_variants(gpu, source):
    """The reference's other source functions (source.h:143-349) on the
    random state with a patch in the cooling band: bit-level agreement as for
    the default source (the oracle matches the reference on these cases)."""
    p, Ui, Vi, Uc, Vc = rand_case(limiter=cm.LIMITER_MC, cooling_band=True, source=source)
    hist = run_pair(p, Ui, Vi, Uc, Vc, 3, cm.PATH_PLAIN64)
    for n, h in enumerate(hist):
        print(f"step {n}: eU={h['eU'].max():.2e}")
        assert h["eU"].max() <= TOL_STEP, h["eU"]
        np.testing.assert_array_equal(h["cg"], h["co"])


@pytest.mark.parametrize("source", SOURCES)
def test_fused64_source_variants(gpu, source):
    p, Ui, Vi, Uc, Vc = rand_case(limiter=cm.LIMITER_MC, consistent=True, nr=40, nphi=32, nz=34, source=source)
    hist = run_pair(p, Ui, Vi, Uc, Vc, 1, cm.PATH_FUSED64)
    h = hist[0]
    print(f"eU={h['eU']}")
    assert h["eU"].max() <= TOL_STEP, h["eU"]


@pytest.mark.parametrize("source", SOURCES)
def test_fused32_source_variants(gpu, source):
    p, Ui, Vi, Uc, Vc = disc_case(3, 32, 32, 20, source=source)
    x = cm.grid(p)[0]
    hist = run_pair(p, Ui, Vi, Uc, Vc, 3, cm.PATH_FUSED32)
    for n, h in enumerate(hist):
        e = l1_rel_scaled(h["Ug"], h["Uo"], x)
        print(f"step {n}: L1 rel {e}")
        assert e.max() <= 1.0e-5, e


@pytest.mark.parametrize("path", [cm.PATH_PLAIN64, cm.PATH_FUSED64], ids=["plain64", "fused64"])
def test_full_size_256_one_step(gpu, path):
    """The 256x256x128 disc (MC + HLLD, spiral m = 2 at 5 %) at full size:
    one RK3 step against the oracle (16 threads on the host), 1e-12 per
    conserved variable, dt and its cell equal."""
    p, Ui, Vi, Uc, Vc = disc_case(2, 256, 256, 128)
    x = cm.grid(p)[0]
    hist = run_pair(p, Ui, Vi, Uc, Vc, 1, path)
    h = hist[0]
    e = rel_err_scaled(h["Ug"], h["Uo"], x)
    print(f"{cm.PATH_NAMES[path]}: scaled eU={e.max():.2e} dt {h['dtg']:.17g} {h['dto']:.17g}")
    assert e.max() <= TOL_STEP, e
    assert abs(h["dtg"] - h["dto"]) <= 1e-14 * h["dto"] and h["ijkg"] == h["ijko"]


@pytest.mark.parametrize("path", [cm.PATH_PLAIN64, cm.PATH_FUSED64])
def test_totals_100_steps(gpu, path):
    """Total mass, energy and L1 div B tracked to 1e-10 relative over 100
    steps, each side with its own dt (the full main.c loop)."""
    p, Ui, Vi, Uc, Vc = disc_case(2, 24, 24, 16)
    q = p.copy()
    q.path = path
    o = Oracle(q)
    o.set_initial(Ui, Vi)
    g = cm.Context(q)
    g.set_initial(Ui, Vi)
    t_o, _ = o.run_steps(100)
    t_g, dts = g.run(100)
    do, dg = o.diagnostics(), g.diagnostics()
    print(do, dg, t_o, t_g)
    assert abs(t_g - t_o) <= TOL_INT * t_o
    for key in ("mass", "energy", "divb_l1"):
        assert abs(dg[key] - do[key]) <= TOL_INT * abs(do[key]) + 1e-300, (key, dg[key], do[key])
    g.close()


VL_SB = [cm.LIMITER_VANLEER, cm.LIMITER_SUPERBEE]


@pytest.mark.parametrize("limiter", VL_SB, ids=["vanleer", "superbee"])
@pytest.mark.parametrize("solver", [cm.SOLVER_HLLD, cm.SOLVER_HLL], ids=["hlld", "hll"])
def test_fused64_vanleer_superbee_disc(gpu, limiter, solver):
    """vanLeer (XVlrFree, YVlrInitialize, ZVlrFree) and superbee (ZVlrPeriodic)
    interface boundaries on the fused FP64 path, z tiles split (nz = 34)."""
    p, Ui, Vi, Uc, Vc = disc_case(2, 32, 32, 34, limiter=limiter, solver=solver)
    x = cm.grid(p)[0]
    hist = run_pair(p, Ui, Vi, Uc, Vc, 2, cm.PATH_FUSED64)
    for n, h in enumerate(hist):
        e = rel_err_scaled(h["Ug"], h["Uo"], x)
        print(f"step {n}: scaled eU={e.max():.2e}")
        assert e.max() <= TOL_STEP, e
        assert abs(h["dtg"] - h["dto"]) <= 1e-12 * h["dto"]


@pytest.mark.parametrize("rk", [3, 2])
@pytest.mark.parametrize("limiter", VL_SB, ids=["vanleer", "superbee"])
def test_fused64_vanleer_superbee_random_state(gpu, limiter, rk):
    """A random state puts different values on every boundary plane, so a
    wrong interface boundary (periodic instead of free, the wrong Vinitial
    plane, a shifted wrap) shows at the 1e-12 level."""
    p, Ui, Vi, Uc, Vc = rand_case(rk=rk, limiter=limiter, consistent=True, nr=40, nphi=32, nz=34)
    hist = run_pair(p, Ui, Vi, Uc, Vc, 1, cm.PATH_FUSED64)
    h = hist[0]
    print(f"eU={h['eU']}")
    assert h["eU"].max() <= TOL_STEP, h["eU"]


@pytest.mark.parametrize("limiter", VL_SB, ids=["vanleer", "superbee"])
def test_fused32_vanleer_superbee_disc(gpu, limiter):
    p, Ui, Vi, Uc, Vc = disc_case(3, 32, 32, 34, limiter=limiter)
    x = cm.grid(p)[0]
    hist = run_pair(p, Ui, Vi, Uc, Vc, 3, cm.PATH_FUSED32)
    for n, h in enumerate(hist):
        e = l1_rel_scaled(h["Ug"], h["Uo"], x)
        print(f"step {n}: L1 rel {e.max():.2e}")
        assert e.max() <= 1.0e-5, e
