"""Precision form: FP32 state (FP64 pressure recovery, Check energy, CFL and
totals) against the plain FP64 path, on the same dt sequence.

Bar: relative L1 error <= 1e-5 per conserved variable after 100 steps on
the 512x512x256 disc with beta log-uniform in [0.1, 100] per annulus; the
plain FP64 path itself matches the oracle to 1e-12 (test_gpu_parity).
Components that vanish initially are measured against their vector
(precision.errors).
"""
import json

import pytest

import cylmhd as cm
import precision

pytestmark = pytest.mark.gpu

TOL_L1 = 1.0e-5


@pytest.mark.parametrize("grid,steps", [((64, 64, 32), 100), ((512, 512, 256), 100)],
                         ids=["64x64x32", "512x512x256"])
def test_fp32_l1_error_after_100_steps(gpu, grid, steps):
    p = cm.params_config(3, *grid)
    r = precision.compare(p, steps, cm.PATH_PLAIN64, cm.PATH_FUSED32, log=print)
    print(json.dumps(r))
    assert r["l1_max"] <= TOL_L1, r["l1"]
    assert r["mass_rel"] <= TOL_L1 and r["energy_rel"] <= TOL_L1
    # div B of the FP32 run stays at the FP64 run's level
    assert r["divb_l1"]["test"] <= 1.5 * r["divb_l1"]["ref"] + 1e-12
