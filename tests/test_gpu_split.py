"""GPU: the fused stage as two kernels (the r-march writing each cell's
divergence, then the per-cell update; the default) against the one-kernel
stage (CYLMHD_STAGE_KERNELS=1 at context creation), FP32 and FP64 state.
Both run the same arithmetic in the same order, so the states, dt, its
argmax cell and the floor counts must agree bit for bit after every step,
for every limiter, solver, RK order and source function, on tiles cut short
at the phi and z edges and on r-marches split into segments.  The parity of
the fused paths with the oracle itself is covered by test_gpu_parity and
test_gpu_precision.
"""
import os

import numpy as np
import pytest

import cylmhd as cm
from cases import disc_case, rand_case

pytestmark = pytest.mark.gpu


def make_ctx(p, one_kernel):
    old = os.environ.get("CYLMHD_STAGE_KERNELS")
    os.environ["CYLMHD_STAGE_KERNELS"] = "1" if one_kernel else "0"
    try:
        return cm.Context(p)
    finally:
        if old is None:
            del os.environ["CYLMHD_STAGE_KERNELS"]
        else:
            os.environ["CYLMHD_STAGE_KERNELS"] = old


def run(p, Ui, Vi, Uc, Vc, nsteps, one_kernel, path):
    q = p.copy()
    q.path = path
    g = make_ctx(q, one_kernel)
    g.set_initial(Ui, Vi)
    g.upload(Uc, Vc)
    hist = []
    t = 0.0
    for _ in range(nsteps):
        dt, ch, ijk = g.setdt()
        g.gravity(t)
        g.step(dt, ch)
        t += dt
        U, _ = g.download()
        hist.append(dict(dt=dt, ch=ch, ijk=ijk, counts=np.array(g.floor_counts()), U=U))
    g.close()
    return hist


def assert_same(p, Ui, Vi, Uc, Vc, nsteps, path=cm.PATH_FUSED32):
    split = run(p, Ui, Vi, Uc, Vc, nsteps, False, path)
    one = run(p, Ui, Vi, Uc, Vc, nsteps, True, path)
    for n, (a, b) in enumerate(zip(split, one)):
        nd = int(np.count_nonzero(a["U"] != b["U"]))
        print(f"step {n}: dt {a['dt']:.17g} / {b['dt']:.17g}, cells differing {nd}, counts {a['counts'].ravel()}")
        assert a["dt"] == b["dt"] and a["ch"] == b["ch"] and a["ijk"] == b["ijk"]
        np.testing.assert_array_equal(a["counts"], b["counts"])
        np.testing.assert_array_equal(a["U"], b["U"])
    return split


@pytest.mark.parametrize("limiter", [cm.LIMITER_MINMOD, cm.LIMITER_MC, cm.LIMITER_VANLEER, cm.LIMITER_SUPERBEE])
@pytest.mark.parametrize("solver", [cm.SOLVER_HLLD, cm.SOLVER_HLL])
@pytest.mark.parametrize("path", [cm.PATH_FUSED32, cm.PATH_FUSED64], ids=["fused32", "fused64"])
def test_split_disc(gpu, limiter, solver, path):
    # 30 x 22 x 19 interior cells: partial tiles in phi (14 / 6) and z (16)
    p, Ui, Vi, Uc, Vc = disc_case(2, 30, 22, 19, limiter=limiter, solver=solver)
    assert_same(p, Ui, Vi, Uc, Vc, 3, path)


@pytest.mark.parametrize("rk", [3, 2])
@pytest.mark.parametrize("limiter,solver", [(cm.LIMITER_MINMOD, cm.SOLVER_HLLD), (cm.LIMITER_MC, cm.SOLVER_HLLD),
                                             (cm.LIMITER_VANLEER, cm.SOLVER_HLL), (cm.LIMITER_SUPERBEE, cm.SOLVER_HLLD)])
@pytest.mark.parametrize("path", [cm.PATH_FUSED32, cm.PATH_FUSED64], ids=["fused32", "fused64"])
def test_split_random_state(gpu, rk, limiter, solver, path):
    # every HLLD branch, the cooling band and all three floors
    p, Ui, Vi, Uc, Vc = rand_case(rk=rk, limiter=limiter, solver=solver, cooling_band=True)
    hist = assert_same(p, Ui, Vi, Uc, Vc, 3, path)
    assert hist[0]["counts"].sum() > 0


@pytest.mark.parametrize("source", [1, 2, 3, 4])
@pytest.mark.parametrize("path", [cm.PATH_FUSED32, cm.PATH_FUSED64], ids=["fused32", "fused64"])
def test_split_sources(gpu, source, path):
    p, Ui, Vi, Uc, Vc = disc_case(2, 20, 16, 12, source=source)
    assert_same(p, Ui, Vi, Uc, Vc, 2, path)


@pytest.mark.parametrize("path", [cm.PATH_FUSED32, cm.PATH_FUSED64], ids=["fused32", "fused64"])
def test_split_segments(gpu, path):
    # 200 r-planes: the segment timing on the first stage picks among
    # several segmentations; every choice must give the same states
    p, Ui, Vi, Uc, Vc = disc_case(3, 200, 30, 34)
    assert_same(p, Ui, Vi, Uc, Vc, 2, path)


@pytest.mark.parametrize("path", [cm.PATH_FUSED32, cm.PATH_FUSED64, cm.PATH_PLAIN64], ids=["fused32", "fused64", "plain64"])
def test_kernel_timing_split(gpu, path):
    """cylmhd_kernel_timing: which = 1 / 2 time the r-march and the update of
    the two-kernel stage apart (their sum is the whole stage, which = 0);
    the plain path has no such split (no launches reported)."""
    p, Ui, Vi, Uc, Vc = disc_case(2, 64, 32, 32)
    q = p.copy()
    q.path = path
    with cm.Context(q) as g:
        g.set_initial(Ui, Vi)
        g.kernel_timing(0, reset=True)
        t = 0.0
        for _ in range(3):
            dt, ch, _ = g.setdt()
            g.gravity(t)
            g.step(dt, ch)
            t += dt
        m_ms, m_n = g.kernel_timing(1)
        u_ms, u_n = g.kernel_timing(2)
        s_ms, s_n = g.kernel_timing(0, reset=True)
        print(path, s_ms, s_n, m_ms, m_n, u_ms, u_n)
        assert s_n == 9 and s_ms > 0
        if path == cm.PATH_PLAIN64:
            assert m_n == 0 and u_n == 0
        else:
            assert m_n == 9 and u_n == 9 and m_ms > 0 and u_ms > 0
            assert abs((m_ms + u_ms) - s_ms) <= 0.05 * s_ms + 0.01
        assert g.kernel_timing(0)[1] == 0          # reset
