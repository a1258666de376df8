"""Precision form: FP32 state (FP64 pressure recovery, Check energy, CFL and
totals) against the plain FP64 path, on the same dt sequence.

Bar: relative L1 error <= 1e-5 per conserved variable after 100 steps on
the 512x512x256 disc with beta log-uniform in [0.1, 100] per annulus; the
plain FP64 path itself matches the oracle to 1e-12 (test_gpu_parity).
The FP32 path's short-horizon agreement (5 steps, 32x32x20) is in
test_gpu_parity.test_fused32_disc.
Components that vanish initially are measured against their vector
(precision.errors).
"""
import json

import pytest

import cylmhd as cm
import precision

pytestmark = pytest.mark.gpu

TOL_L1 = 1.0e-5


def test_fp32_l1_error_after_100_steps(gpu):
    """The bar over the whole interior, every variable: relative L1 <= 1e-5.
    rho, r rho vphi, B_phi and E are measured against their own magnitude;
    rho vr, rho vz, Br, Bz and psi (zero in the initial disc) against the
    vector they belong to.  The error over the cells outside the absorbing
    layers and the plain per-component L1 are reported beside it."""
    p = cm.params_config(3, 512, 512, 256)
    r = precision.compare(p, 100, cm.PATH_PLAIN64, cm.PATH_FUSED32, log=print)
    print(json.dumps(r))
    assert r["l1_max"] <= TOL_L1, r["l1"]
    for k in ("rho", "r_rho_vphi", "bphi", "e"):
        assert r["l1_plain"][k] <= TOL_L1, r["l1_plain"]
    assert r["mass_rel"] <= 1e-7 and r["energy_rel"] <= 1e-7
    # div B of the FP32 run stays at the FP64 run's level
    assert r["divb_l1"]["test"] <= 1.5 * r["divb_l1"]["ref"] + 1e-12
