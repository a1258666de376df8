"""GPU: cylmhd_run, the reference's driver loop (main.c:224-398) on the
library, with its text outputs and the dump -> text converter.

The run writes, with every dump, GDLOutputCylinder's files (--gdl) and the
IDL text slices (--idl); --convert then rebuilds the same text files from
the binary dumps alone.  On the plain FP64 path the uploaded primitives are
kept as given, so the converted files must equal the ones written during
the run byte for byte; the dumps themselves must read back as the state the
run had (RestartRead -> the same GDL bytes).
"""
import os
import subprocess

import numpy as np
import pytest

import cylmhd as cm

from conftest import PKG
from test_slices import read_dir

pytestmark = pytest.mark.gpu

EXE = os.path.join(PKG, "host", "cylmhd_run")


def run(args, timeout=120):
    print("cylmhd_run", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def test_driver_text_output_and_converter(gpu, tmp_path):
    if not os.path.isfile(EXE):
        pytest.fail(f"{EXE} not built")
    out = tmp_path / "dumps"
    out.mkdir()
    common = ["--config", "1", "--grid", "16", "16", "10", "--path", "plain64", "--outdir", str(out)]
    r = run(common + ["--nstop", "3", "--tint", "1e-9", "--tstop", "1e9", "--quiet", "--gdl",
                      "--idl", str(tmp_path / "idl_run")])
    dumps = sorted(int(d) for d in os.listdir(out) if d.isdigit())
    assert dumps == list(range(len(dumps))) and len(dumps) >= 4, dumps
    gdl_run = read_dir(out)
    assert sum(f.startswith("2d-density-rphi-") for f in gdl_run) == len(dumps)
    idl_run = {d: read_dir(tmp_path / "idl_run" / d) for d in os.listdir(tmp_path / "idl_run")}
    assert sorted(d for d in idl_run if d.startswith("idl")) == sorted("idl%d" % n for n in dumps)
    assert sorted(d for d in idl_run if d.startswith("azimuth")) == ["azimuth-%d" % K for K in range(4)]
    assert all(len(v) == 25 for d, v in idl_run.items() if d.startswith("idl"))   # 4 coordinate lists + 21 slices
    assert all(len(v) == 1 + 6 * len(dumps) for d, v in idl_run.items() if d.startswith("azimuth"))
    run(common + ["--convert", "0", str(dumps[-1]), "--gdl", "--idl", str(tmp_path / "idl_conv")])
    assert read_dir(out) == gdl_run
    idl_conv = {d: read_dir(tmp_path / "idl_conv" / d) for d in os.listdir(tmp_path / "idl_conv")}
    # the same files, except the arm potential: the converter evaluates it at
    # nout * tint (the restart's time, main.c:269), the run at the dump time
    strip = lambda dd: {d: {f: b for f, b in v.items() if not f.startswith("potential")} for d, v in dd.items()}
    assert strip(idl_conv) == strip(idl_run)


def test_driver_converter_rejects_other_grid(gpu, tmp_path):
    out = tmp_path / "dumps"
    out.mkdir()
    run(["--config", "1", "--grid", "16", "16", "10", "--outdir", str(out), "--nstop", "0", "--quiet"])
    r = subprocess.run([EXE, "--config", "1", "--grid", "18", "16", "10", "--outdir", str(out), "--convert", "0",
                        "0", "--idl", str(tmp_path / "idl")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "grid" in r.stderr


def test_driver_restart_matches_oracle(gpu, tmp_path):
    """--restart NOUT (main.c:244-288): RestartRead(0) -- VtoU and Check with
    no floors set yet --, SaveInitial, RestartRead(NOUT) -- VtoU and Check
    with the saved floors --, t = NOUT*tint, then the loop.  The dumps carry
    a cell below each floor so that both Checks fire; the spiral amplitude
    ramps with t (epssp0 t / tausp, spiralgravity.h:41-42), so a wrong restart
    time changes the state.  On the plain FP64 path every dump written after
    the restart equals the oracle's state to 1e-12 relative per variable (the
    device's sin/cos of the arm phase may differ from glibc's in the last
    bit; with no spiral the two are bit-identical, test_gpu_parity), with
    its time."""
    from cases import rel_err
    from oracle_py import Oracle, parse_counts
    from test_host import restart_states

    out = tmp_path / "dumps"
    out.mkdir()
    p = cm.params_config(1, 16, 16, 10)
    p.epssp0, p.tausp = 0.05, 1.0
    V0, Vk = restart_states(p)
    x, y, z = cm.grid(p)
    nout, tint = 3, 1.0e-3
    cm.binary_output(str(out / "0"), 0.0, x, y, z, V0)
    cm.binary_output(str(out / str(nout)), nout * tint, x, y, z, Vk)
    r = run(["--config", "1", "--grid", "16", "16", "10", "--path", "plain64", "--outdir", str(out),
             "--restart", str(nout), "--nstop", "2", "--tint", str(tint), "--tstop", "1e9",
             "--param", "epssp0", "0.05", "--param", "tausp", "1.0"])
    o = Oracle(p)
    c0 = o.restart_state(cm.restart_read(str(out / "0"), p.dims))
    U0, V0c = o.state()
    o.set_initial(U0, V0c)
    ck = o.restart_state(cm.restart_read(str(out / str(nout)), p.dims))
    counts = parse_counts(r.stdout)
    assert counts[0].tolist() == c0 and counts[1].tolist() == ck, (counts[:2], c0, ck)
    assert ck[0] >= 1 and ck[1] >= 1 and ck[2] >= 1
    t = nout * tint                                   # main.c:269
    for n in range(3):
        dt, ch, _ = o.setdt()
        o.gravity(t)
        o.step(dt, ch)
        t += dt
        d = out / str(nout + 1 + n)
        Vd = cm.restart_read(str(d), p.dims)
        _, Vo = o.state()
        e = rel_err(Vd, Vo)
        print("dump", nout + 1 + n, "max rel err", e.max(), "t", t)
        assert np.isfinite(Vd).all() and e.max() <= 1e-12, e
        assert (d / "GridTime.dat").read_text().split("\t")[3] == "%.3e" % t
