"""The oracle against golden vectors produced by the compiled reference
(tests/golden/make_golden.py).  Runs without /root/reference and without a
GPU; requires bit-identical U, V, dt, ch and floor counters."""
import glob
import json
import os

import numpy as np
import pytest

from golden.make_golden import build_inputs, digest
from oracle_py import Oracle

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(g)[:-4] for g in GOLDEN])
def test_oracle_reproduces_reference_vectors(path):
    f = np.load(path, allow_pickle=False)
    meta = json.loads(str(f["meta"]))
    p, Ui, Vi, Uc, Vc = build_inputs(meta["kind"], meta["kwargs"])
    assert digest(Ui, Vi, Uc, Vc) == meta["input_sha256"], "input generator changed: regenerate fixtures"
    o = Oracle(p)
    o.set_initial(Ui, Vi)
    o.set_state(Uc, Vc)
    t, hist = o.run_steps(meta["nsteps"])
    U, V = o.state()
    np.testing.assert_array_equal([h[0] for h in hist], f["dt"])
    np.testing.assert_array_equal([h[1] for h in hist], f["ch"])
    np.testing.assert_array_equal(U, f["U"])
    np.testing.assert_array_equal(V, f["V"])
    nst = f["counts"].shape[0]
    np.testing.assert_array_equal(o.floor_counts()[:nst], f["counts"])


def test_fixtures_present():
    assert len(GOLDEN) >= 4
