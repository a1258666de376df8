#!/usr/bin/env python3
"""Write tests/golden/*.npz: outputs of the reference's own routines.

Runs the compiled reference harnesses (oracle/_ref, built from /root/reference
by oracle/build_ref.py) on seeded inputs and stores what they produce:
U and V after n RK steps, dt and ch per step, and the Check floor counters
of the last step.  The inputs are not stored; test_golden.py regenerates
them from the same recipe and checks their SHA-256 first, so a changed input
generator is reported as such rather than as an oracle mismatch.

Usage (in a tree that has oracle/_ref built):  python3 tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "171026-1-3dhlld-cyli-gal_amd"), os.path.join(ROOT, "tests")]

import cylmhd as cm  # noqa: E402
from cases import rand_case, ref_disc_case  # noqa: E402
from oracle_py import Oracle, parse_counts, ref_harness, run_ref  # noqa: E402

CASES = {
    # name: (harness, builder, kwargs, nsteps, galactic)
    "disc16_mm_hlld_rk3": ("disc16_mm_hlld", "disc", dict(config=1, nr=16, nphi=16, nz=10, epssp0=0.0), 4, False),
    "disc16_mc_hll_rk2": ("disc16_mc_hll", "disc", dict(config=1, nr=16, nphi=16, nz=10, epssp0=0.0,
                                                        limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLL, rk=2), 4, False),
    "rand16_mc_hlld_rk3": ("rand16_mc_hlld", "rand", dict(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD), 2, True),
    "rand16_mm_hll_rk2": ("rand16_mm_hll", "rand", dict(limiter=cm.LIMITER_MINMOD, solver=cm.SOLVER_HLL, rk=2), 2, True),
    # the reference's other source functions (source.h:143-349), cooling band patch on
    "rand16_mc_hlld_axic_rk3": ("rand16_mc_hlld_axic", "rand", dict(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD,
                                                                  cooling_band=True, source=cm.SOURCE_AXI_COOLING), 2, True),
    "rand16_mc_hlld_axi_rk3": ("rand16_mc_hlld_axi", "rand", dict(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD,
                                                                cooling_band=True, source=cm.SOURCE_AXI), 2, True),
    "rand16_mc_hlld_ugz_rk3": ("rand16_mc_hlld_ugz", "rand", dict(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD,
                                                                cooling_band=True, source=cm.SOURCE_UNIFORM_GZ), 2, True),
    "rand16_mc_hlld_none_rk3": ("rand16_mc_hlld_none", "rand", dict(limiter=cm.LIMITER_MC, solver=cm.SOLVER_HLLD,
                                                                  cooling_band=True, source=cm.SOURCE_NONE), 2, True),
}


# cases whose run also writes GDLOutputCylinder's text files (output.h:280-384,
# dump number 7): stored as bytes in golden/text/<name>_gdl.npz
TEXT_CASES = ("disc16_mm_hlld_rk3",)


def build_inputs(kind, kw):
    if kind == "disc":
        kw = dict(kw)
        cfg = kw.pop("config")
        return ref_disc_case(cfg, kw.pop("nr"), kw.pop("nphi"), kw.pop("nz"), **kw)
    return rand_case(**kw)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    for name, (harness, kind, kw, nsteps, galactic) in CASES.items():
        if ref_harness(harness) is None:
            print(f"skip {name}: {harness} not built")
            continue
        p, Ui, Vi, Uc, Vc = build_inputs(kind, kw)
        o = Oracle(p)
        gx, gz = o.gravity_axi()
        with tempfile.TemporaryDirectory() as td:
            dumpdir = td if name in TEXT_CASES else None
            r, out = run_ref(harness, Vi, Vc, gx, gz, nsteps, p.rk, galactic=galactic, U_init=Ui, U_cur=Uc,
                             dumpdir=dumpdir)
            if dumpdir:
                texts = {f: np.frombuffer(open(os.path.join(td, f), "rb").read(), dtype=np.uint8)
                         for f in sorted(os.listdir(td)) if os.path.isfile(os.path.join(td, f))}
                os.makedirs(os.path.join(HERE, "text"), exist_ok=True)
                np.savez_compressed(os.path.join(HERE, "text", name + "_gdl.npz"), **texts)
                print(f"wrote text/{name}_gdl.npz: {sorted(texts)}")
        counts = parse_counts(out)
        nst = 3 if p.rk == 3 else 2
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            U=r["U"], V=r["V"], dt=r["dt"], ch=r["ch"], counts=counts[-nst:],
            meta=json.dumps(dict(harness=harness, kind=kind, kwargs=kw, nsteps=nsteps, galactic=galactic,
                                 input_sha256=digest(Ui, Vi, Uc, Vc))))
        print(f"wrote {name}: {len(r['dt'])} steps")
    return 0


if __name__ == "__main__":
    sys.exit(main())
