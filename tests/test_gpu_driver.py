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

import pytest

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
