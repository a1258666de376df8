"""Batching form: many discs in one launch per stage, per-disc dt on the
device, the step replayed from a hipGraph.

Each cell's arithmetic in the batch kernels is the single-disc FP32 stage's
(same work-item function), and the per-disc dt, ch and spiral amplitude are
formed by the same expressions; only the psi damping factor is evaluated by
the device's exp instead of the host's.  So a disc advanced in a batch must
agree with the same disc advanced alone through the single-disc ABI to FP32
rounding, whatever the tile shape, segment length and launch grid.
"""
import numpy as np
import pytest

import cylmhd as cm
from cases import group_scales, l1_rel_scaled

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 32), (48, 64, 32), (32, 48, 32), (64, 32, 48)]


def batch_params(dtmin=None):
    ps = []
    for d, (nr, nphi, nz) in enumerate(SHAPES):
        p = cm.params_config(4, nr, nphi, nz)
        p.path = cm.PATH_FUSED32
        p.seed = 1000 + d
        p.epssp0 = 0.025 * d
        if dtmin is not None and d == 1:
            p.dtmin = dtmin
        ps.append(p)
    return ps


def run_single(p, nsteps):
    U, V = cm.disc_state(p)
    with cm.Context(p) as g:
        g.set_initial(U, V)
        t = 0.0
        for _ in range(nsteps):
            dt, ch, _ = g.setdt()
            g.gravity(t)
            g.step(dt, ch)
            t += dt
        Uo, _ = g.download()
    return Uo, t


def start_batch(ps):
    b = cm.Batch(ps)
    for d, p in enumerate(ps):
        U, V = cm.disc_state(p)
        b.disc(d).set_initial(U, V)
    return b


def test_batch_matches_single_discs(gpu):
    ps = batch_params()
    nsteps = 5
    with start_batch(ps) as b:
        ms = b.run(nsteps)
        assert ms > 0
        for d, p in enumerate(ps):
            st = b.status(d)
            Ub, _ = b.disc(d).download()
            Us, ts = run_single(p, nsteps)
            e = l1_rel_scaled(Ub, Us, cm.grid(p)[0])
            print(d, (p.ix, p.jy, p.kz), st, ts, e.max())
            assert st["alive"] and st["steps"] == nsteps
            # dt comes from the FP32 states; a last-ulp psi damping factor
            # moves the states, and with them dt, at the 1e-8 level
            assert abs(st["t"] - ts) <= 1e-6 * ts
            assert e.max() <= 1e-6, e


def test_batch_tuner_keeps_the_states(gpu):
    ps = batch_params()
    with start_batch(ps) as b:
        choice = b.tune(2)
        print("tuned", choice)
        assert choice[0] in (0, 1, 2, 3) and choice[1] in (0, 8, 16, 32) and choice[2] in (0, 512, 1024)
        assert choice[3] in (1, 2)
        assert all(b.disc_tile(d)[0] in (0, 1, 2, 3) for d in range(len(ps)))
        b.run(3)
        tuned = [b.disc(d).download()[0] for d in range(len(ps))]
        times = [b.status(d)["t"] for d in range(len(ps))]
    with start_batch(ps) as b:
        b.run(3)
        for d, p in enumerate(ps):
            U = b.disc(d).download()[0]
            assert b.status(d)["t"] == pytest.approx(times[d], rel=1e-12)
            e = l1_rel_scaled(tuned[d], U, cm.grid(p)[0])
            assert e.max() <= 1e-6, (d, e)


def test_batch_tuner_mid_run_keeps_time_steps_and_stopped_discs(gpu):
    """Tuning after some steps: each disc resumes from its own time, step
    count and alive flag (the spiral ramp and the arm phase follow t), so
    run(3), tune(2), run(2) ends where run(5) does; a disc that stopped
    below dtmin stays stopped.  After the tune the first dt comes from the
    stored state instead of the stage kernel's partials, which can move it
    in the last bits, so the states agree to FP32 rounding."""
    ps = batch_params(dtmin=1.0)           # disc 1 stops at once
    with start_batch(ps) as b:
        b.run(3)
        b.tune(2)
        b.run(2)
        tuned = [(b.status(d), b.disc(d).download()[0]) for d in range(len(ps))]
    with start_batch(ps) as b:
        b.run(5)
        for d, p in enumerate(ps):
            st = b.status(d)
            st_t, U_t = tuned[d]
            print(d, st, st_t)
            assert st_t["alive"] == st["alive"] and st_t["steps"] == st["steps"]
            assert st["steps"] == (0 if d == 1 else 5)
            assert st_t["t"] == pytest.approx(st["t"], rel=1e-6, abs=0.0)
            e = l1_rel_scaled(U_t, b.disc(d).download()[0], cm.grid(p)[0])
            assert e.max() <= 1e-6, (d, e)


def compare_with_single(ps, nsteps=2):
    """Two steps: the floors of Check are discontinuous in the state, and
    with the less diffusive limiters a last-ulp difference (the device's exp
    in the psi damping) flips one pressure-floor decision in the third step
    of one of these discs (then rho vr is zeroed in one run and not in the
    other: a local jump, not a drift)."""
    with start_batch(ps) as b:
        b.run(nsteps)
        for d, p in enumerate(ps):
            st = b.status(d)
            Ub, _ = b.disc(d).download()
            Us, ts = run_single(p, nsteps)
            e = l1_rel_scaled(Ub, Us, cm.grid(p)[0])
            print(d, (p.ix, p.jy, p.kz), p.limiter, p.source, st, ts, e.max())
            assert st["alive"] and st["steps"] == nsteps
            assert abs(st["t"] - ts) <= 1e-6 * ts
            assert e.max() <= 1e-6, e
            # the z-side ghost planes, which a batch keeps packed apart, come
            # back through the download like the single disc's: the same
            # relative error as the interior's (pointwise, both differ alike),
            # measured on the scales of the L1 bar (the components that start
            # at zero -- rho vz, Br, Bz, psi -- against their vector: on their
            # own noise-level magnitude the two compiled kernels' roundings
            # differ by 1e-4 relative after two steps)
            scale = group_scales(Us, cm.grid(p)[0], l1=False, interior=False)[:, None, None] + 1e-30
            inner = (np.abs(Ub - Us)[:, 1:-1, 1:-1, 1:-1] / scale[:, :, :, None]).max()
            for kg in (0, -1):
                dz = np.abs(Ub[:, :, :, kg] - Us[:, :, :, kg]) / scale
                print("ghost plane", kg, "mean", dz.mean(), "max", dz.max(), "interior max", inner)
                assert dz.mean() <= 1e-6, (kg, dz.mean())


def test_batch_mixed_sources_match_single_discs(gpu):
    """Each disc of a batch may run its own source function (source.h)."""
    ps = batch_params()
    for p, src in zip(ps, [cm.SOURCE_SPIRAL_COOLING, cm.SOURCE_AXI_COOLING, cm.SOURCE_UNIFORM_GZ, cm.SOURCE_NONE]):
        p.source = src
    compare_with_single(ps)


def test_batch_stops_only_the_disc_below_dtmin(gpu):
    ps = batch_params(dtmin=1.0)          # disc 1: every dt is below dtmin
    with start_batch(ps) as b:
        b.run(2)
        for d in range(len(ps)):
            st = b.status(d)
            if d == 1:
                assert not st["alive"] and st["steps"] == 0
            else:
                assert st["alive"] and st["steps"] == 2


@pytest.mark.parametrize("limiter", [cm.LIMITER_MC, cm.LIMITER_VANLEER, cm.LIMITER_SUPERBEE],
                         ids=["mc", "vanleer", "superbee"])
def test_batch_other_limiters_match_single_discs(gpu, limiter):
    """The batch runs the limiter every disc shares, with its interface
    boundaries (vanLeer's YVlrInitialize reads each disc's own Vinitial)."""
    ps = batch_params()
    for p in ps:
        p.limiter = limiter
    compare_with_single(ps)


@pytest.mark.parametrize("solver,rk", [(cm.SOLVER_HLL, 3), (cm.SOLVER_HLLD, 2), (cm.SOLVER_HLL, 2)],
                         ids=["hll-rk3", "hlld-rk2", "hll-rk2"])
def test_batch_hll_and_rk2_match_single_discs(gpu, solver, rk):
    """The batch with the HLL solver (its kernels are built for two tile
    shapes) and with RK2 (two stages, the RK2 z boundary and z absorbing
    layer in the ghost pass) runs what each disc runs alone."""
    ps = batch_params()
    for p in ps:
        p.solver = solver
        p.rk = rk
    compare_with_single(ps)


def test_batch_rejects_mixed_limiters(gpu):
    ps = batch_params()
    ps[1].limiter = cm.LIMITER_MC
    with pytest.raises(cm.CylmhdError) as e:
        cm.Batch(ps)
    assert e.value.code == cm.EINVAL


def test_batch_disc_is_not_stepped_alone(gpu):
    """A batch's disc keeps the batch's state layout; only the batch steps it."""
    ps = batch_params()
    with start_batch(ps) as b:
        g = b.disc(0)
        dt, ch, _ = g.setdt()
        with pytest.raises(cm.CylmhdError) as e:
            g.step(dt, ch)
        assert e.value.code == cm.EINVAL
        U, V = g.download()                  # set-up, dumps and diagnostics still work
        assert np.isfinite(U).all() and g.diagnostics()["mass"] > 0


def test_batch_stage_forms_agree(gpu, monkeypatch):
    """The stage as one kernel per work item and as two (the r-march per
    work item, then the per-cell update over every disc) run the same
    arithmetic: states, times and floor counts equal bit for bit."""
    ps = batch_params()
    out = {}
    for form in ("1", "2"):
        monkeypatch.setenv("CYLMHD_STAGE_KERNELS", form)
        with start_batch(ps) as b:
            b.run(4)
            out[form] = [(b.status(d), b.disc(d).download()[0]) for d in range(len(ps))]
    for d in range(len(ps)):
        (s1, U1), (s2, U2) = out["1"][d], out["2"][d]
        print(d, s1, s2, int(np.count_nonzero(U1 != U2)))
        assert s1 == s2
        np.testing.assert_array_equal(U1, U2)
