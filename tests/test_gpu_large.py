"""The large disc with every source active (768x768x384, ~226 M cells):
power-law cooling Lambda ~ T^0.5 above T = 1e-2, spiral gravity m=2 at 10 %,
GLM alpha = 0.1, MC + HLLD.  The FP32 path against the FP64 fused path on
the same grid and dt sequence (the plain pass-by-pass FP64 path needs more
memory than one GPU holds at this size).

Checks at the full size: the two precisions agree to the FP32 bar (relative
L1 <= 1e-5 per variable over the whole interior) after the configuration's
50 steps, totals agree and div B stays at the FP64 level.  At a small size: the cooling source is active (the run loses energy
against the same run without cooling, in both precisions alike).
"""
import json

import numpy as np
import pytest

import cylmhd as cm
import precision

pytestmark = pytest.mark.gpu

GRID = (768, 768, 384)


@pytest.mark.timeout(900)
def test_large_disc_fp32_against_fp64(gpu):
    p = cm.params_config(5, *GRID)
    assert p.cooling == cm.COOL_POWERLAW and p.limiter == cm.LIMITER_MC
    r = precision.compare(p, 50, cm.PATH_FUSED64, cm.PATH_FUSED32, log=print)
    print(json.dumps({k: r[k] for k in ("l1", "l1_max", "physical", "mass_rel", "energy_rel", "divb_l1")}))
    assert np.isfinite(r["l1_max"])
    assert r["l1_max"] <= 1e-5, r["l1"]
    assert r["mass_rel"] <= 1e-7 and r["energy_rel"] <= 1e-7
    assert r["divb_l1"]["test"] <= 1.5 * r["divb_l1"]["ref"] + 1e-12


def _energy_after(p, steps):
    U, V = cm.disc_state(p)
    with cm.Context(p) as g:
        g.set_initial(U, V)
        t = 0.0
        for _ in range(steps):
            dt, ch, _ = g.setdt()
            g.gravity(t)
            g.step(dt, ch)
            t += dt
        return g.diagnostics()["energy"]


@pytest.mark.parametrize("path", [cm.PATH_FUSED32, cm.PATH_FUSED64])
def test_large_disc_cooling_is_active(gpu, path):
    p = cm.params_config(5, 64, 64, 32)
    p.path = path
    e_cool = _energy_after(p, 5)
    q = cm.params_config(5, 64, 64, 32)
    q.path = path
    q.cooling = cm.COOL_NONE
    e_none = _energy_after(q, 5)
    print(cm.PATH_NAMES[path], e_cool, e_none)
    assert e_cool < e_none
