"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol of include/cylmhd.h, parameter defaults equal the reference's
input.h/constant.h, the product's grid equals the oracle's, initial states
are consistent, the dump format round-trips, and bad input is rejected with
an error code (never an exit)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import cylmhd as cm
from cases import disc_case
from oracle_py import Oracle


def test_library_exports_every_header_symbol():
    syms = cm.exported_symbols()
    assert len(syms) >= 20
    assert cm.missing_symbols() == []


def test_params_default_are_reference_values():
    p = cm.params_default()
    assert (p.ix, p.jy, p.kz) == (240, 64, 200)                   # input.h:45-47
    assert p.dy0 == (2.0 * math.pi) / 64.0                         # input.h:69
    assert p.dxmax == 50.0 * 0.01 and p.dzmax == 50.0 * 0.0003      # input.h:75-76
    assert (p.cfl, p.dtmin, p.rhocr, p.prcr, p.betacr) == (0.4, 1.0e-9, 1.0e-4, 0.015, 0.1)
    assert (p.gm, p.gm1, p.oneogm1, p.cr) == (5.0 / 3.0, 2.0 / 3.0, 1.5, 0.18)   # constant.h:62-78
    assert p.te0 == (0.619047619 * 1.673e-24) * (1.0 / 1.3806504e-16) * 6.558e+6 * 6.558e+6
    assert p.omegasp == 10.0e+5 / (1.0 / 4.7058e+15) * (1.0 / 3.08568e+21)
    assert p.tausp == 200.0 * 1.0e+6 * 3.1536e+7 / 4.7058e+15
    assert p.pitch == 15.0 * math.pi / 180.0 and p.armnum == 2 and p.epssp0 == 0.05
    assert p.rk == 3 and p.limiter == cm.LIMITER_MINMOD and p.solver == cm.SOLVER_HLLD


@pytest.mark.parametrize("config", [1, 2, 3, 5])
def test_config_grids(config):
    p = cm.params_config(config, 24, 16, 12)
    x, y, z = cm.grid(p)
    dr = 0.9 / 24
    # faces at r = 0.1 and 1, z = -0.5 and 0.5 (centres of the first and
    # last interior cells half a cell inside)
    assert abs(x[1] - dr / 2 - 0.1) < 1e-14 and abs(x[-2] + dr / 2 - 1.0) < 1e-12
    assert abs(z[1] - 1.0 / 24 + 0.5) < 1e-14 and abs(z[-2] + 1.0 / 24 - 0.5) < 1e-12
    assert y[1] == 0.0 and abs(y[-1] - 2 * math.pi) < 1e-12
    o = Oracle(p)
    g = o.grid()
    np.testing.assert_array_equal(x, g["x"])
    np.testing.assert_array_equal(y, g["y"])
    np.testing.assert_array_equal(z, g["z"])


def test_disc_state_ghosts_follow_the_boundary_map():
    p, U, V, _, _ = disc_case(2, 16, 16, 10)
    NX, NY, NZ = p.dims
    jy = p.jy
    # Y periodic
    np.testing.assert_array_equal(U[:, 1:-1, 0, 1:-1], U[:, 1:-1, NY - 2, 1:-1])
    np.testing.assert_array_equal(U[:, 1:-1, NY - 1, 1:-1], U[:, 1:-1, 1, 1:-1])
    # Z periodic
    np.testing.assert_array_equal(U[:, 1:-1, 1:-1, 0], U[:, 1:-1, 1:-1, NZ - 2])
    # X axis: across the axis with sign flips of r and phi components
    sign = np.array([1, -1, -1, 1, -1, -1, 1, 1, 1.0])[:, None]
    for j in range(1, NY - 1):
        js = j + jy // 2 if j < jy // 2 + 1 else j - jy // 2
        np.testing.assert_array_equal(U[:, 0, j, 1:-1], sign * U[:, 1, js, 1:-1])
    # X right free
    np.testing.assert_array_equal(U[:, NX - 1, 1:-1, 1:-1], U[:, NX - 2, 1:-1, 1:-1])
    # interior V is the disc, interior U = VtoU(V)
    np.testing.assert_array_equal(U[:, 1:-1, 1:-1, 1:-1], cm.vtou(p, V)[:, 1:-1, 1:-1, 1:-1])


def test_disc_matches_its_definition():
    p = cm.params_config(2, 16, 16, 10)
    V = cm.model_disc(p)
    x, y, z = cm.grid(p)
    rho = np.exp(-z ** 2 / (2 * 0.1 ** 2)) + 1e-3
    np.testing.assert_allclose(V[0], np.broadcast_to(rho, V[0].shape), rtol=1e-15)
    np.testing.assert_allclose(V[7], 0.01 * V[0], rtol=1e-15)
    vk = x[:, None, None] ** -0.5
    assert np.all(np.abs(V[2] / vk - 1) <= 0.01 + 1e-15)      # 1% perturbation
    np.testing.assert_allclose(V[5] ** 2 * 10.0 / 2, V[7], rtol=1e-14)   # beta = 10
    assert np.all(V[[1, 3, 4, 6, 8]] == 0)


def test_beta_per_annulus_log_uniform():
    p = cm.params_config(3, 64, 8, 8)
    V = cm.model_disc(p)
    beta = 2 * V[7] / V[5] ** 2
    b_i = beta[:, 3, 3]
    np.testing.assert_allclose(beta, b_i[:, None, None] * np.ones_like(beta), rtol=1e-13)
    assert b_i.min() >= 0.1 * (1 - 1e-12) and b_i.max() <= 100.0 * (1 + 1e-12)
    assert b_i.max() / b_i.min() > 10


def test_dump_restart_roundtrip(tmp_path):
    p, U, V, _, _ = disc_case(1, 16, 16, 10)
    x, y, z = cm.grid(p)
    d = tmp_path / "7"
    cm.binary_output(str(d), 1.25, x, y, z, V)
    assert (d / "GridTime.dat").read_text() == "%d\t%d\t%d\t%.3e" % (18, 18, 12, 1.25)
    assert np.array_equal(np.fromfile(d / "r", dtype=np.float64), x)
    assert np.array_equal(np.fromfile(d / "pr", dtype=np.float64).reshape(V.shape[1:]), V[7])
    V2 = cm.restart_read(str(d), V.shape[1:])
    np.testing.assert_array_equal(V2, V)
    with pytest.raises(cm.CylmhdError):
        cm.restart_read(str(tmp_path / "missing"), V.shape[1:])


def test_bad_parameters_are_errors_not_exits():
    p = cm.params_config(1, 16, 16, 10)
    q = p.copy(); q.jy = 15                      # odd phi count: no across-axis map
    h = C.c_void_p()
    assert cm.lib().cylmhd_create(C.byref(q), C.byref(h)) == -1
    q = p.copy(); q.rk = 2; q.kz = 6             # RK2 z absorbing layer needs kz >= 9
    assert cm.lib().cylmhd_create(C.byref(q), C.byref(h)) == -1
    assert cm.lib().cylmhd_params_config(C.byref(q), 9, 0, 0, 0) == -1
    msg = cm.lib().cylmhd_error_string(-1).decode()
    assert msg.startswith("invalid argument")


def test_no_device_is_an_error_not_a_fallback():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    p = cm.params_config(1, 16, 16, 10)
    with pytest.raises(cm.CylmhdError):
        cm.Context(p)


def restart_states(p):
    """Primitives as two dumps would hold them: dump 0 (the initial state, one
    cell below the pressure floor) and a later dump with one cell below each of
    the three floors of Check (check.h:51-173)."""
    _, V0 = cm.disc_state(p)
    V0 = V0.copy()
    V0[7, 3, 4, 5] = 0.1 * p.prcr * V0[0, 3, 4, 5]                  # p < prcr rho
    Vk = V0.copy()
    Vk[2] *= 1.01
    Vk[0, 5, 6, 4] = 1.0e-2 * p.rhocr * V0[0, 5, 6, 4]              # rho < rhocr rho_init
    Vk[7, 6, 2, 3] = 0.2 * p.prcr * Vk[0, 6, 2, 3]                  # p < prcr rho
    Vk[5, 8, 9, 6] = math.sqrt(2.0 * Vk[7, 8, 9, 6] * 30.0)         # 1/beta = 30 > 1/betacr
    return V0, Vk


def test_check_host_is_the_oracle_check_at_restart():
    """cylmhd_check_host after cylmhd_vtou_host = RestartRead's VtoU + Check
    (restartread.h:132-139), bit for bit against the oracle's: first with the
    floors not yet set (RestartRead(0) before SaveInitial, main.c:250), then
    with the floors of the saved initial state (RestartRead(nout))."""
    p = cm.params_config(1, 16, 16, 10)
    V0, Vk = restart_states(p)
    o = Oracle(p)
    c0 = o.restart_state(V0)
    Uo0, Vo0 = o.state()
    V = V0.copy()
    U = cm.vtou(p, V)
    assert cm.check_host(p, U, V) == c0
    # no density floor before SaveInitial; the planted cell and the axis
    # ghosts (i = 0: the across-axis copy of r rho vphi at the ghost's own
    # radius leaves p < 0 on this annulus grid) take the pressure floor
    assert c0[0] == 0 and c0[1] >= 1 and c0[2] == 0
    assert V[7, 3, 4, 5] == V[0, 3, 4, 5] * p.prcr
    np.testing.assert_array_equal(U, Uo0)
    np.testing.assert_array_equal(V, Vo0)
    o.set_initial(Uo0, Vo0)
    ck = o.restart_state(Vk)
    Uok, Vok = o.state()
    V = Vk.copy()
    U = cm.vtou(p, V)
    assert cm.check_host(p, U, V, rho_init=Uo0[0]) == ck
    assert ck[0] >= 1 and ck[1] >= 1 and ck[2] >= 1
    np.testing.assert_array_equal(U, Uok)
    np.testing.assert_array_equal(V, Vok)
    # the branch table of check.h: the density floor sets rho = rhofloor and
    # zeroes vr, vz; the pressure floor sets p = prcr rho; the beta floor
    # scales rho and p by betacr / beta
    c = (slice(None), 5, 6, 4)
    assert V[0][c[1:]] == Uo0[0][c[1:]] * p.rhocr and V[1][c[1:]] == 0.0 and V[3][c[1:]] == 0.0
    assert U[2][c[1:]] == cm.grid(p)[0][5] * V[0][c[1:]] * V[2][c[1:]]
    assert V[7, 6, 2, 3] == V[0, 6, 2, 3] * p.prcr
    beta = V[7, 8, 9, 6] / (0.5 * (V[4, 8, 9, 6] ** 2 + V[5, 8, 9, 6] ** 2 + V[6, 8, 9, 6] ** 2))
    assert abs(beta - p.betacr) <= 1e-14
