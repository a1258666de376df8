"""GPU: slices cut on the device and the text writers built on them.

cylmhd_slice against the restatement (oracle/oracle_slices.py) applied to
the state the same context downloads: bit-identical for every quantity and
plane on every path (one FP64 expression per value; the cooling rate, whose
pow may differ from the host's libm in the last place, to 1e-14).  The
writers: cylmhd_gdl_output_ctx byte-identical to cylmhd_gdl_output on the
downloaded state (which is byte-identical to the reference's
GDLOutputCylinder, test_slices.py), and cylmhd_idl_output byte-identical
to the restatement's files.
"""
import os

import numpy as np
import pytest

import cylmhd as cm
import oracle_slices as osl
from cases import disc_case
from oracle_py import Oracle
from test_slices import check_layout, encoded, read_dir

pytestmark = pytest.mark.gpu

PATHS = [cm.PATH_PLAIN64, cm.PATH_FUSED64, cm.PATH_FUSED32]
PLANES = [(cm.PLANE_RPHI, 1, 0), (cm.PLANE_RPHI, 4, 0), (cm.PLANE_RPHI, 0, 0), (cm.PLANE_RPHI_ZMEAN, 0, 0),
          (cm.PLANE_RPHI_ZINT, 0, 0), (cm.PLANE_RZ, 0, 0), (cm.PLANE_RZ, 3, 0), (cm.PLANE_ZR, 2, 0),
          (cm.PLANE_XZ, 1, 0), (cm.PLANE_XZ, 5, 0), (cm.PLANE_XZ_PHIMEAN, 0, 0), (cm.PLANE_PHI, 6, 3)]


def stepped(path, nsteps=2):
    """A 20 x 16 x 12 disc with power-law cooling and 5 % spiral arms on,
    nsteps steps on the path."""
    p, Ui, Vi, Uc, Vc = disc_case(1, 20, 16, 12, cooling=cm.COOL_POWERLAW, cool_lambda0=0.05, cool_tmin=0.0,
                                  cool_index=0.5, epssp0=0.05)
    p.path = path
    g = cm.Context(p)
    g.set_initial(Ui, Vi)
    g.upload(Uc, Vc)
    t, dts = g.run(nsteps)
    g.t_gravity = t - dts[-1] if dts else 0.0     # time of the last SpiralGravity
    return p, g


def spiral(p, g):
    x, y, z = cm.grid(p)
    return osl.spiral_tables(p, x, y, z, g.t_gravity)


# quantities whose device and host values can differ in the last place
# (libm's cos / tan of the arm phase and pitch): compared to 1e-12 relative
NEAR = (cm.Q_SPIRALPOT, cm.Q_VPARA, cm.Q_VPERP)


@pytest.mark.parametrize("path", PATHS, ids=[cm.PATH_NAMES[q] for q in PATHS])
def test_slices_match_restatement(gpu, path):
    p, g = stepped(path)
    o = Oracle(p)
    geo, dzm = o.geometry(), o.dzm()
    cool = lambda rho, pr: o.L.orc_cooling(p, rho, pr)
    _, V = g.download()
    sp = spiral(p, g)
    for q in range(cm.NQ):
        Q = osl.quantity(V, q, geo, cool, sp)
        for kind, a, b in PLANES:
            got = g.slice(q, kind, a, b)
            want = osl.plane(Q, kind, a, b, dzm)
            print(f"{cm.PATH_NAMES[path]} q={q} plane={kind} a={a} b={b} n={got.size}", flush=True)
            if q == cm.Q_ONEOTCOOL:
                np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)
            elif q in NEAR:
                # relative to the quantity's magnitude (the phi-mean of the
                # arm potential is a cancellation: roundoff of that size)
                np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(np.abs(Q).max(), 1e-300))
            else:
                np.testing.assert_array_equal(got, want)
    oc = osl.plane(osl.quantity(V, cm.Q_ONEOTCOOL, geo, cool), cm.PLANE_RPHI, 1)
    assert (oc > 0).any()            # the cooling term was live
    g.close()


def test_slice_layout_on_device(gpu):
    """An index-encoded state uploaded to the plain path (which keeps V as
    given) decodes, plane by plane, to the cells the IDL readers expect."""
    p, Ui, Vi, Uc, Vc = disc_case(1, 16, 16, 10)
    g = cm.Context(p)
    g.set_initial(Ui, Vi)
    V = encoded(p)
    g.upload(Uc, V)
    for kind, a, b in [(cm.PLANE_RPHI, 1, 0), (cm.PLANE_RPHI, 7, 0), (cm.PLANE_ZR, 1, 0), (cm.PLANE_RZ, 0, 0),
                       (cm.PLANE_XZ, 1, 0), (cm.PLANE_XZ, 9, 0), (cm.PLANE_PHI, 2, 5)]:
        check_layout(p, g.slice(0, kind, a, b), kind, a, b)
    g.close()


def test_slice_rejects_bad_arguments(gpu):
    p, g = stepped(cm.PATH_PLAIN64, 0)
    NX, NY, NZ = p.dims
    for q, kind, a, b in [(cm.NQ, cm.PLANE_RPHI, 1, 0), (-1, cm.PLANE_RPHI, 1, 0), (0, 8, 0, 0),
                          (0, cm.PLANE_RPHI, NZ, 0), (0, cm.PLANE_RZ, -1, 0), (0, cm.PLANE_XZ, 0, 0),
                          (0, cm.PLANE_XZ, NY - 1, 0), (0, cm.PLANE_PHI, NX, 0), (0, cm.PLANE_PHI, 1, NZ)]:
        with pytest.raises(cm.CylmhdError):
            g.slice(q, kind, a, b)
    g.close()


@pytest.mark.parametrize("path", [cm.PATH_PLAIN64, cm.PATH_FUSED32], ids=["plain64", "fused32"])
def test_gdl_output_ctx_matches_host_writer(gpu, path, tmp_path):
    p, g = stepped(path)
    _, V = g.download()
    x, y, z = cm.grid(p)
    cm.gdl_output(str(tmp_path / "host"), 3, x, y, z, V)
    g.gdl_output(str(tmp_path / "dev"), 3)
    host, dev = read_dir(tmp_path / "host"), read_dir(tmp_path / "dev")
    assert sorted(host) == sorted(dev) and len(host) == 7
    for f in host:
        assert host[f] == dev[f], f
    g.close()


@pytest.mark.parametrize("path", [cm.PATH_FUSED64, cm.PATH_FUSED32], ids=["fused64", "fused32"])
def test_idl_output_matches_restatement(gpu, path, tmp_path):
    p, g = stepped(path)
    o = Oracle(p)
    cool = lambda rho, pr: o.L.orc_cooling(p, rho, pr)
    _, V = g.download()
    x, y, z = cm.grid(p)
    sp = spiral(p, g)
    assert sp["eps"] > 0
    for nout in (0, 12):
        g.idl_output(str(tmp_path), nout, kplane=2, k100=100, jplane=3, azim_i=(3, 0, 9, 0), azim_k=2)
        want = osl.idl_files(V, x, y, z, o.geometry(), o.dzm(), cool, nout, kplane=2, k100=100, jplane=3,
                             spiral=sp, azim_i=(3, 0, 9, 0), azim_k=2)
        for rel, b in want.items():
            got = open(os.path.join(tmp_path, rel), "rb").read()
            if any(s in rel for s in ("potential", "vpara", "vperp", "velocitypara")):
                gv, wv = np.array(got.split(), float), np.array(b.split(), float)
                np.testing.assert_allclose(gv, wv, rtol=0, atol=2e-6 * max(np.abs(wv).max(), 1e-300))
            elif "oneocoolingtime" in rel:
                np.testing.assert_allclose(np.array(got.split(), float), np.array(b.split(), float), rtol=1e-5)
            else:
                assert got == b, rel
        names = set(os.listdir(tmp_path / ("idl%d" % nout)))
        assert names == {os.path.basename(r) for r in want if r.startswith("idl%d/" % nout)}
    g.close()
