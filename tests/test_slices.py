"""Text output (CPU, no GPU): GDLOutputCylinder's files and the slice
restatement the GPU slice tests check against.

* cylmhd_gdl_output (host arrays) is byte-identical to the reference's own
  GDLOutputCylinder (output.h:280-384), run by the compiled harness, and to
  the golden text fixture that harness wrote.
* The slice restatement (oracle/oracle_slices.py) is pinned by what the
  mathematics and the IDL scripts fix: an index-encoded field decodes to the
  cell each reader expects (fltarr(ny, nx) / fltarr(nz, nx) / fltarr(2 NX, nz)
  + readf: density.pro:27-83, density-rz.pro:27-61, density-xz.pro:27-61),
  div B of fields with known divergence, the column integral of 1 equals the
  column height, means of linear fields, and the reference's own formulas
  where it has them (benergy output.h:353-355, Setdt's cooling rate
  setdt.h:56-58 through the oracle's Cooling).
"""
import os

import numpy as np
import pytest

import cylmhd as cm
import oracle_slices as osl
from cases import disc_params, ref_disc_case
from golden.make_golden import build_inputs
from oracle_py import Oracle, ref_harness, run_ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_dir(d):
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d))
            if os.path.isfile(os.path.join(d, f))}


def test_gdl_text_matches_reference(tmp_path):
    """GDLOutputCylinder of the reference (compiled harness, dump 7) vs
    cylmhd_gdl_output and the restatement, on the reference's own state."""
    name = "disc16_mm_hlld"
    if ref_harness(name) is None:
        pytest.skip("reference harness not built")
    p, Ui, Vi, Uc, Vc = ref_disc_case(1, 16, 16, 10, epssp0=0.0)
    o = Oracle(p)
    gx, gz = o.gravity_axi()
    refdir = tmp_path / "ref"
    refdir.mkdir()
    r, _ = run_ref(name, Vi, Vc, gx, gz, 2, 3, dumpdir=str(refdir), U_init=Ui, U_cur=Uc)
    ref = read_dir(refdir)
    assert sorted(ref) == sorted(osl.gdl_files(r["V"], *cm.grid(p), 7))
    x, y, z = cm.grid(p)
    cm.gdl_output(str(tmp_path / "mine"), 7, x, y, z, r["V"])
    mine = read_dir(tmp_path / "mine")
    assert sorted(mine) == sorted(ref)
    for f in ref:
        assert mine[f] == ref[f], f
    for f, b in osl.gdl_files(r["V"], x, y, z, 7).items():
        assert b == ref[f], f


def test_gdl_text_matches_golden(tmp_path):
    """The same without the reference build: the reference's text files
    stored with the golden state they were written from."""
    f = np.load(os.path.join(GOLDEN, "disc16_mm_hlld_rk3.npz"), allow_pickle=False)
    t = np.load(os.path.join(GOLDEN, "text", "disc16_mm_hlld_rk3_gdl.npz"), allow_pickle=False)
    import json
    meta = json.loads(str(f["meta"]))
    p = build_inputs(meta["kind"], meta["kwargs"])[0]
    x, y, z = cm.grid(p)
    cm.gdl_output(str(tmp_path), 7, x, y, z, f["V"])
    mine = read_dir(tmp_path)
    assert sorted(mine) == sorted(t.files)
    for name in t.files:
        assert mine[name] == t[name].tobytes(), name
    # the fixture holds real data: 16 x 16 rho values at k = 1, positive
    rho = np.array(t["2d-density-rphi-7.dat"].tobytes().split(), dtype=float)
    assert rho.size == 16 * 16 and (rho > 0).all()


def encoded(p):
    """V with V[m][i][j][k] = m 1e9 + i 1e6 + j 1e3 + k (exact in FP64)."""
    NX, NY, NZ = p.dims
    i, j, k = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    return np.stack([m * 1e9 + i * 1e6 + j * 1e3 + k for m in range(9)]).astype(np.float64)


def decode(v):
    v = np.asarray(v).astype(np.int64)
    return v // 10**9, (v // 10**6) % 1000, (v // 1000) % 1000, v % 1000


def check_layout(p, vals, kind, a=0, b=0):
    """The decoded cell of every value is the one the reader expects."""
    NX, NY, NZ = p.dims
    ni, nj, nk = NX - 2, NY - 2, NZ - 2
    x, y, z = cm.grid(p)
    m, i, j, k = decode(vals)
    assert (m == 0).all()
    if kind == osl.PLANE_RPHI:          # fltarr(ny, nx): arr[i][j] at (r[i], phi[j])
        I, J = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nj + 1), indexing="ij")
        assert (i.reshape(ni, nj) == I).all() and (j.reshape(ni, nj) == J).all() and (k == a).all()
    elif kind == osl.PLANE_ZR:          # fltarr(nz, nx): arr[i][k]
        I, K = np.meshgrid(np.arange(1, ni + 1), np.arange(1, nk + 1), indexing="ij")
        assert (i.reshape(ni, nk) == I).all() and (k.reshape(ni, nk) == K).all() and (j == a).all()
    elif kind == osl.PLANE_RZ:          # output.h:366-380: k outer, i inner
        K, I = np.meshgrid(np.arange(1, nk + 1), np.arange(1, ni + 1), indexing="ij")
        assert (i.reshape(nk, ni) == I).all() and (k.reshape(nk, ni) == K).all() and (j == a).all()
    elif kind == osl.PLANE_XZ:          # fltarr(2 NX, nz): arr[k][col] at (x[col], z[k])
        xs = np.concatenate([-x[::-1], x])
        I, J, K = (t.reshape(nk, 2 * NX) for t in (i, j, k))
        assert (K == np.arange(1, nk + 1)[:, None]).all()
        np.testing.assert_array_equal(np.abs(xs)[None, :].repeat(nk, 0), x[I])
        # the negative half is the opposite azimuth: phi differs by pi
        jo = 1 + (a - 1 + nj // 2) % nj
        assert (J[:, :NX] == jo).all() and (J[:, NX:] == a).all()
        assert abs(abs(y[jo] - y[a]) - np.pi) < 1e-12
    elif kind == osl.PLANE_PHI:
        assert (i == a).all() and (k == b).all() and (j == np.arange(1, nj + 1)).all()
    else:
        raise AssertionError(kind)


@pytest.mark.parametrize("kind,a,b", [(osl.PLANE_RPHI, 1, 0), (osl.PLANE_RPHI, 5, 0), (osl.PLANE_ZR, 1, 0),
                                      (osl.PLANE_ZR, 7, 0), (osl.PLANE_RZ, 0, 0), (osl.PLANE_XZ, 1, 0),
                                      (osl.PLANE_XZ, 6, 0), (osl.PLANE_PHI, 3, 4)])
def test_plane_layout_matches_idl_readers(kind, a, b):
    p = disc_params(1, 16, 16, 10)
    V = encoded(p)
    check_layout(p, osl.plane(V[0], kind, a, b), kind, a, b)


def test_divb_of_known_fields():
    """Br = r: div = (1/r) d(r^2)/dr = 2; Bz = z: div = 1; Bphi = const:
    div = 0 -- exact in the finite-volume form up to rounding, so a swapped
    metric factor, a wrong sign or a shifted face fails."""
    p = disc_params(1, 16, 16, 10)
    o = Oracle(p)
    geo = o.geometry()
    g = o.grid()
    NX, NY, NZ = p.dims
    V = np.zeros((9, NX, NY, NZ))
    V[0] = 1.0
    V[7] = 1.0
    inner = (slice(1, -1),) * 3
    V[4] = g["x"][:, None, None] * np.ones((1, NY, NZ))
    # the face-mean form needs xm = (x[i] + x[i+1]) / 2, as the grid builds it
    np.testing.assert_allclose(g["xm"][:-1], 0.5 * (g["x"][:-1] + g["x"][1:]), rtol=1e-14)
    d = osl.quantity(V, osl.Q_DIVB, geo)
    np.testing.assert_allclose(d[inner], 2.0, rtol=1e-10)
    V[4] = 0.0
    V[6] = g["z"][None, None, :] * np.ones((NX, NY, 1))
    d = osl.quantity(V, osl.Q_DIVB, geo)
    np.testing.assert_allclose(d[inner], 1.0, rtol=1e-10)
    V[6] = 0.0
    V[5] = 3.0
    assert np.abs(osl.quantity(V, osl.Q_DIVB, geo)[inner]).max() < 1e-12
    assert (osl.quantity(V, osl.Q_DIVB, geo)[0] == 0).all()     # zero outside the interior


def test_column_integral_and_means():
    p = disc_params(1, 16, 16, 10)
    o = Oracle(p)
    g = o.grid()
    dzm = o.dzm()
    NX, NY, NZ = p.dims
    one = np.ones((NX, NY, NZ))
    col = osl.plane(one, osl.PLANE_RPHI_ZINT, dzm=dzm)
    # sum of the cell heights of the interior = zm[NZ-2] - zm[0] (cylindricalgrid.h:93-94)
    zm = 0.5 * (g["z"][:-1] + g["z"][1:])
    np.testing.assert_allclose(col, zm[NZ - 2] - zm[0], rtol=1e-12)
    lin = np.arange(NZ, dtype=float)[None, None, :] * one
    np.testing.assert_allclose(osl.plane(lin, osl.PLANE_RPHI_ZMEAN), 0.5 * (1 + NZ - 2), rtol=1e-14)
    ph = np.cos(g["y"])[None, :, None] * one
    assert np.abs(osl.plane(ph, osl.PLANE_XZ_PHIMEAN)).max() < 1e-14
    np.testing.assert_array_equal(osl.plane(3.0 * one, osl.PLANE_XZ_PHIMEAN), 3.0)


def test_derived_quantities():
    p = disc_params(1, 16, 16, 10, cooling=cm.COOL_POWERLAW, cool_lambda0=0.1, cool_tmin=0.0, cool_index=0.5)
    o = Oracle(p)
    NX, NY, NZ = p.dims
    rng = np.random.default_rng(3)
    V = rng.uniform(0.5, 2.0, (9, NX, NY, NZ))
    b2 = V[4] ** 2 + V[5] ** 2 + V[6] ** 2
    np.testing.assert_allclose(osl.quantity(V, osl.Q_BENERGY), 0.5 * b2, rtol=1e-15)
    np.testing.assert_allclose(osl.quantity(V, osl.Q_BETA) * osl.quantity(V, osl.Q_BENERGY), V[7], rtol=1e-15)
    np.testing.assert_allclose(osl.quantity(V, osl.Q_TEMP) * V[0], V[7], rtol=1e-15)
    np.testing.assert_allclose(osl.quantity(V, osl.Q_VDOTB), (V[1:4] * V[4:7]).sum(0), rtol=1e-14)
    np.testing.assert_allclose(osl.quantity(V, osl.Q_RHOBZ), V[0] * V[6], rtol=1e-15)
    cool = lambda rho, pr: o.L.orc_cooling(p, rho, pr)
    lam = np.vectorize(cool)(V[0], V[7])
    assert (lam > 0).any()
    # Setdt's cooling rate: 2 rho^2 Lambda / (1.5 pr) = rho^2 Lambda / e_th (e_th = 1.5 pr)
    np.testing.assert_allclose(osl.quantity(V, osl.Q_ONEOTCOOL, cooling=cool), 2 * V[0] ** 2 * lam / (1.5 * V[7]),
                               rtol=1e-14)


def test_spiral_potential_matches_the_arm_force():
    """-(1/r) dPhi/dphi of the arm potential against the phi component of
    the oracle's spiral gravity (the reference's SpiralGravity, pinned by the
    compiled reference): central differences, error O(dphi^2)."""
    p = disc_params(2, 24, 64, 10)
    t = 0.5 * p.tausp                               # the arms at half amplitude
    x, y, z = cm.grid(p)
    NX, NY, NZ = p.dims
    sp = osl.spiral_tables(p, x, y, z, t)
    assert 0 < sp["eps"] < p.epssp0
    phi = osl.quantity(np.zeros((9, NX, NY, NZ)), osl.Q_SPIRALPOT, spiral=sp)
    o = Oracle(p)
    o.gravity(t)
    gy = o.get_gravity()[1]
    irin = int(p.rin / p.dx0)
    inner = (slice(irin + 1, NX - 1), slice(1, NY - 1), slice(1, NZ - 1))
    dphi = (phi[:, 2:, :] - phi[:, :-2, :]) / (y[2:] - y[:-2])[None, :, None]
    f = -dphi / x[:, None, None]
    want = gy[inner]
    got = f[irin + 1:NX - 1, :, 1:NZ - 1]
    assert np.abs(want).max() > 0
    err = np.abs(got - want).max() / np.abs(want).max()
    dm = p.armnum * (y[2] - y[1])
    assert err <= 0.5 * dm * dm, err                # (m dphi)^2 / 6 of a pure cosine, with margin


def test_arm_frame_velocities():
    p = disc_params(2, 16, 16, 10)
    NX, NY, NZ = p.dims
    x, y, z = cm.grid(p)
    sp = osl.spiral_tables(p, x, y, z, 0.0)
    rng = np.random.default_rng(5)
    V = rng.uniform(-1.0, 1.0, (9, NX, NY, NZ))
    va, vp = osl.quantity(V, osl.Q_VPARA, spiral=sp), osl.quantity(V, osl.Q_VPERP, spiral=sp)
    np.testing.assert_allclose(va * va + vp * vp, V[1] ** 2 + V[2] ** 2, rtol=1e-13)
    V[1] = 0.0                                      # pure rotation: vpara = vphi cos(pitch)
    np.testing.assert_allclose(osl.quantity(V, osl.Q_VPARA, spiral=sp), V[2] * np.cos(p.pitch), rtol=1e-13)


def test_idl_file_set_matches_scripts():
    """Every file name the restatement writes is one an IDL script opens."""
    ref = "/root/reference"
    if not os.path.isdir(ref):
        pytest.skip("reference scripts not present")
    import re
    opened = set()
    for f in os.listdir(ref):
        if f.endswith(".pro"):
            txt = open(os.path.join(ref, f), errors="replace").read()
            opened |= set(re.findall(r"'/([a-z0-9-]+)-'\+idlnumstring\+'\.dat'", txt))
    names = {n for n, *_ in osl.IDL_FILES}
    assert names <= opened, names - opened
    missing = opened - names
    assert not missing, missing
    # the azimuthal profiles of phiout*.pro / fourieramp.pro under ./azimuth-K/
    azim = set()
    for f in os.listdir(ref):
        if f.endswith(".pro"):
            txt = open(os.path.join(ref, f), errors="replace").read()
            azim |= set(re.findall(r"'\./azimuth-\d/([a-z]+)-'\+idlnumstring", txt))
    written = {"potential", "rho", "te", "vperp", "vpara", "velocitypara"}
    assert azim == written, azim ^ written
