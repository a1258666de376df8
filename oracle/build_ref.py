#!/usr/bin/env python3
"""Compile the reference's own routines into small harness binaries.

TEST INFRASTRUCTURE ONLY.  Reads the reference sources where they lie
(/root/reference, or $CYLMHD_REFERENCE) and writes only into oracle/_ref/
(git-ignored).  Nothing is copied: the harness (oracle/ref_harness.c)
#includes the reference headers through -I.  The one file that differs per
build is ref_input.h, written here, which takes the place of the reference's
input.h: the reference fixes its grid sizes with compile-time #defines
(input.h:45-72), so each grid needs its own build.  Values are written as
hexadecimal floating literals so they equal the Python-side parameters
bit for bit.

Usage: python3 oracle/build_ref.py        (no-op with a message if the
reference tree is absent, e.g. on a GPU box that only carries the builds)
"""
import math
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
REF = os.environ.get("CYLMHD_REFERENCE", "/root/reference")

# grid families: name -> (nr, nphi, nz, kind)
#   kind "disc": benchmark-disc geometry (r in [0.1,1], z in [-0.5,0.5],
#                inertial frame); "rand": same geometry, reference rotating
#                frame (omegasp from input.h:94) for random-state tests.
GRIDS = {
    "disc16": (16, 16, 10, "disc"),
    "rand16": (16, 8, 10, "rand"),
}
# scheme name -> (limiter define or None, HLL instead of HLLD, source define or None)
SCHEMES = {"mm_hlld": (None, False, None), "mc_hlld": ("REF_LIMITER_MC", False, None),
           "mm_hll": (None, True, None), "mc_hll": ("REF_LIMITER_MC", True, None),
           "vl_hlld": ("REF_LIMITER_VL", False, None), "sb_hlld": ("REF_LIMITER_SB", False, None),
           "mc_hlld_axic": ("REF_LIMITER_MC", False, "REF_SOURCE_AXIC"),
           "mc_hlld_axi": ("REF_LIMITER_MC", False, "REF_SOURCE_AXI"),
           "mc_hlld_ugz": ("REF_LIMITER_MC", False, "REF_SOURCE_UGZ"),
           "mc_hlld_none": ("REF_LIMITER_MC", False, "REF_SOURCE_NONE")}


def hexf(v):
    return float(v).hex()


def grid_values(nr, nphi, nz, kind):
    dx0 = 0.9 / nr
    dz0 = 1.0 / nz
    vals = dict(ix=nr, jy=nphi, kz=nz, dx0=dx0, x0=0.1 + 0.5 * dx0,
                dz0=dz0, z0=-0.5 + 0.5 * dz0, lx=1.0e300, lz=1.0e300)
    if kind == "disc":
        vals["omegasp"] = 0.0
        vals["omegapattern"] = 0.0
    else:
        om = 10.0e+5 / (1.0 / 4.7058e+15) * (1.0 / 3.08568e+21)
        vals["omegasp"] = om
        vals["omegapattern"] = om
    return vals


def ref_input_h(v):
    return f"""/* generated by oracle/build_ref.py: stands in the place of input.h
 * (input.h:45-102) with this build's grid; all other reference code is used
 * as it is. */
#define ix {v['ix']}
#define jy {v['jy']}
#define kz {v['kz']}
#define lx {hexf(v['lx'])}
#define ly pi2
#define lz {hexf(v['lz'])}
#define x0 {hexf(v['x0'])}
#define y0 0.0
#define z0 {hexf(v['z0'])}
#define stretchratio 1.05
#define rin 0.1
#define margin 2
#define cfl 0.4
#define dtmin 1.0e-9
#define ixmax (ix+margin)
#define jymax (jy+margin)
#define kzmax (kz+margin)
#define ixmax1 (ixmax-1)
#define jymax1 (jymax-1)
#define kzmax1 (kzmax-1)
#define dx0 {hexf(v['dx0'])}
#define dy0 (ly/(double)jy)
#define dz0 {hexf(v['dz0'])}
#define irin ((int)(rin/dx0))
double dxmax=50.0*dx0;
double dzmax=50.0*dz0;
#define rhocr 1.0e-4
#define prcr 0.015
#define betacr 0.1
#define oneobetacr (1.0/betacr)
#define nstop 10000000
#define tint 0.05
#define tstop 40.0
double tout=tint;
double epssp0=0.0;
double tausp=200.0*1.0e+6*year/t0;
double scalez=0.03;
double scaler=0.7;
double scalephase=0.1;
int armnum=2;
double omegasp={hexf(v['omegasp'])};
double omegapattern={hexf(v['omegapattern'])};
double pitch=15.0*M_PI/180.0;
int restartsw=0;
int nout=0;
"""


def main():
    if not os.path.isfile(os.path.join(REF, "firststep.h")):
        print(f"build_ref: reference tree not found at {REF}; skipping (prebuilt "
              f"harnesses in {OUT} are used if present)")
        return 0
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(HERE, "ref_harness.c")
    jobs = []
    for gname, (nr, nphi, nz, kind) in GRIDS.items():
        for sname, (limdef, hll, srcdef) in SCHEMES.items():
            name = f"{gname}_{sname}"
            d = os.path.join(OUT, name)
            os.makedirs(d, exist_ok=True)
            with open(os.path.join(d, "ref_input.h"), "w") as f:
                f.write(ref_input_h(grid_values(nr, nphi, nz, kind)))
            cmd = ["gcc", "-O2", "-std=gnu99", "-fopenmp", "-ffp-contract=off", "-w",
                   "-I", d, "-I", REF, src, "-o", os.path.join(d, "ref_harness"), "-lm"]
            if limdef:
                cmd.insert(1, "-D" + limdef)
            if hll:
                cmd.insert(1, "-DREF_SOLVER_HLL")
            if srcdef:
                cmd.insert(1, "-D" + srcdef)
            jobs.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
    rc = 0
    for name, p in jobs:
        out, _ = p.communicate()
        if p.returncode != 0:
            print(f"build_ref: {name} FAILED\n{out}")
            rc = 1
    print(f"build_ref: built {len(jobs) - rc} harness variants into {OUT}")
    return rc


if __name__ == "__main__":
    sys.exit(main())
