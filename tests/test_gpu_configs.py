"""GPU: each benchmark configuration on the HIP path, at its own size and
step count, against the oracle (the CPU FP64 restatement, bit-identical to
the compiled reference) or against the FP64 GPU path where the oracle cannot
run the configuration in the time of a test.

  1. 64x64x32, minmod + HLLD, no spiral, 10 RK3 steps: plain64 and fused64
     within 1e-12 per conserved variable of the oracle after every step
     (both fed the oracle's dt), fused32 within 1e-5 relative L1.
  2. 256x256x128, MC + HLLD, spiral m = 2 at 5 %: after 1 and 100 steps,
     each path with its own dt -- mass, energy and L1 div B within 1e-10 of
     the oracle's, the time within 1e-10, and plain64's state within 1e-12.
  4. 48 discs of the batch (cylmhd_params_batch(48, 20171026)), 200 graph-
     replayed steps, against the same discs on the FP64 fused path with
     their own dt: every disc within 1e-5 relative L1 after 3 steps, and
     after 200 within the tolerance the test states (the FP32 path's
     pressure-floor decisions at the discs' inner edge); the smallest,
     median and largest disc within 1e-5 of the oracle after 3 steps.
  5. (test_gpu_large) 768x768x384, every source, 50 steps against FP64.
"""
import numpy as np
import pytest

import cylmhd as cm
from cases import disc_case, l1_rel_scaled, rel_err, rel_err_scaled
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def test_config1_every_step(gpu):
    p, Ui, Vi, Uc, Vc = disc_case(1, 64, 64, 32)
    assert p.limiter == cm.LIMITER_MINMOD and p.solver == cm.SOLVER_HLLD and p.epssp0 == 0.0
    x = cm.grid(p)[0]
    o = Oracle(p)
    o.set_initial(Ui, Vi)
    ctx = {}
    for path in (cm.PATH_PLAIN64, cm.PATH_FUSED64, cm.PATH_FUSED32):
        q = p.copy()
        q.path = path
        ctx[path] = cm.Context(q)
        ctx[path].set_initial(Ui, Vi)
    t = 0.0
    for n in range(10):
        dto, cho, ijko = o.setdt()
        for path, g in ctx.items():
            dtg, _, ijkg = g.setdt()
            if path != cm.PATH_FUSED32:
                assert abs(dtg - dto) <= 1e-13 * dto and ijkg == ijko, (path, dtg, dto)
            g.gravity(t)
            g.step(dto, cho)
        o.gravity(t)
        o.step(dto, cho)
        t += dto
        Uo, _ = o.state()
        e = {}
        for path, g in ctx.items():
            Ug, _ = g.download()
            e[path] = l1_rel_scaled(Ug, Uo, x) if path == cm.PATH_FUSED32 else (
                rel_err(Ug, Uo) if path == cm.PATH_PLAIN64 else rel_err_scaled(Ug, Uo, x))
        print(f"step {n + 1}: plain64 {e[0].max():.2e} fused64 {e[1].max():.2e} fused32 L1 {e[2].max():.2e}")
        assert e[cm.PATH_PLAIN64].max() <= 1e-12, e[cm.PATH_PLAIN64]
        assert e[cm.PATH_FUSED64].max() <= 1e-12, e[cm.PATH_FUSED64]
        assert e[cm.PATH_FUSED32].max() <= 1e-5, e[cm.PATH_FUSED32]
    for g in ctx.values():
        g.close()


@pytest.mark.timeout(1200)
def test_config2_1_and_100_steps(gpu):
    """The oracle advances the full 256x256x128 disc on the host's cores."""
    p, Ui, Vi, Uc, Vc = disc_case(2, 256, 256, 128)
    assert p.limiter == cm.LIMITER_MC and p.epssp0 == 0.05 and p.armnum == 2
    x = cm.grid(p)[0]
    o = Oracle(p)
    o.set_initial(Ui, Vi)
    t_o, _ = o.run_steps(1)
    U1o, _ = o.state()
    d1o = o.diagnostics()
    for n in range(11):                         # 99 more steps, a progress line each 9
        t_o, _ = o.run_steps(9, t_o)
        print(f"oracle step {1 + 9 * (n + 1)} t={t_o:.6e}", flush=True)
    U100o, _ = o.state()
    do = o.diagnostics()
    o.close()
    print("oracle", t_o, do)
    for path in (cm.PATH_PLAIN64, cm.PATH_FUSED64):
        q = p.copy()
        q.path = path
        with cm.Context(q) as g:
            g.set_initial(Ui, Vi)
            t_g, _ = g.run(1)
            U1g, _ = g.download()
            d1g = g.diagnostics()
            t_g, _ = g.run(99, t_g)
            U100g, _ = g.download()
            dg = g.diagnostics()
        e1 = rel_err_scaled(U1g, U1o, x)
        e100 = rel_err_scaled(U100g, U100o, x)
        print(cm.PATH_NAMES[path], "step 1", e1.max(), "step 100", e100.max(), t_g, dg)
        assert e1.max() <= 1e-12, e1
        assert abs(t_g - t_o) <= 1e-10 * t_o
        for dd, ref in ((d1g, d1o), (dg, do)):
            for key in ("mass", "energy", "divb_l1"):
                assert abs(dd[key] - ref[key]) <= 1e-10 * abs(ref[key]), (path, key, dd[key], ref[key])
        if path == cm.PATH_PLAIN64:
            assert e100.max() <= 1e-12, e100


def _by_size(ps):
    n = [p.ix * p.jy * p.kz for p in ps]
    order = np.argsort(n, kind="stable")
    return [int(order[0]), int(order[len(order) // 2]), int(order[-1])]


@pytest.mark.timeout(900)
def test_config4_batch_200_steps(gpu):
    """All 48 discs within 1e-5 of FP64 after 3 steps; after the 200 steps
    every disc is alive at step 200 with the FP64 run's time (to 1e-2: dt
    follows the state through the CFL maximum), and the FP32
    state is within the stated tolerance of FP64 (median disc 1e-3, every
    disc 1e-1).  The tolerance is not the 1e-5 bar: the innermost cells of
    these discs sit on Check's pressure floor (p = prcr rho, check.h:96-120)
    at E/p ~ 5000 (v_phi ~ 3 at r ~ 0.1), where FP32 rounding of the energy
    flux puts ~1e-4 relative noise on the recovered p; a floor decision that
    goes the other way zeroes vr and vz in one run and not in the other, and
    these local jumps spread (measured, 200 steps: median disc 4e-4, worst
    3e-2).  The FP64 path from a state rounded once to FP32 stays at 2e-7."""
    ps = cm.params_batch(48, 20171026)
    assert len({(p.ix, p.jy, p.kz) for p in ps}) > 10
    states = [cm.disc_state(p) for p in ps]
    fp64 = []
    for d, p in enumerate(ps):
        q = p.copy()
        q.path = cm.PATH_FUSED64
        g = cm.Context(q)
        g.set_initial(*states[d])
        fp64.append(g)
    with cm.Batch(ps) as b:
        for d, (U, V) in enumerate(states):
            b.disc(d).set_initial(U, V)
        b.run(3)
        for d in _by_size(ps):
            p = ps[d]
            Ub, _ = b.disc(d).download()
            o = Oracle(p)
            o.set_initial(*states[d])
            t_o, _ = o.run_steps(3)
            Uo, _ = o.state()
            o.close()
            e = l1_rel_scaled(Ub, Uo, cm.grid(p)[0])
            st = b.status(d)
            print("oracle disc", d, (p.ix, p.jy, p.kz), e.max(), st["t"], t_o, flush=True)
            assert e.max() <= 1e-5, e
            assert abs(st["t"] - t_o) <= 1e-6 * t_o
        done, fdone, t64 = 3, 0, [0.0] * len(ps)
        for upto in (3, 200):
            b.run(upto - done)
            errs = []
            for d, p in enumerate(ps):
                st = b.status(d)
                assert st["alive"] and st["steps"] == upto, (d, st)
                t64[d], _ = fp64[d].run(upto - fdone, t64[d])
                Ub, _ = b.disc(d).download()
                U64, _ = fp64[d].download()
                e = l1_rel_scaled(Ub, U64, cm.grid(p)[0])
                errs.append(e.max())
                print("step", upto, "disc", d, (p.ix, p.jy, p.kz), "L1", e.max(), "t", st["t"], t64[d], flush=True)
                # the dt of each step follows the state (the CFL maximum)
                assert abs(st["t"] - t64[d]) <= (1e-6 if upto == 3 else 1e-2) * t64[d]
                if upto == 3:
                    assert e.max() <= 1e-5, (d, e)
            errs = np.array(errs)
            print("step", upto, "L1 median", np.median(errs), "worst", errs.max(), flush=True)
            if upto == 200:
                assert np.median(errs) <= 1e-3 and errs.max() <= 1e-1, errs
            done = fdone = upto
    for g in fp64:
        g.close()
