#!/usr/bin/env python3
"""bench.py -- cell-updates/s of the RK3 step on one MI355X.

Workload (the metric's own configuration): the 512x512x256 disc -- FP32 state
with FP64 pressure recovery, conserved sums and CFL reduction, MC + HLLD,
TVD-RK3, spiral gravity m=2 at 5 %, beta log-uniform in [0.1, 100] per radial
annulus.  A step is the reference's main-loop body: Setdt (with its
device->host dt), SpiralGravity, TVDRK.  The same disc is also timed on the
two FP64 paths (fused, and the plain pass-by-pass drop-in) for the FP32 vs
FP64 comparison.

value     = interior cells x K / wall seconds of K steps (state resident in
            HBM; host<->device traffic is only the scalars of Setdt).
roofline  = the RK stage, two kernels (the r-march: limiter states, HLLD
            faces and each cell's flux divergence; then the per-cell update:
            sources, RK combination, UtoV, Check, store): algorithmic HBM
            bytes per stage (288 B per cell-update / 3 stages, FP32 state) /
            the pair's average duration from HIP events on the library's
            stream, each kernel's share under "kernels"; traffic (HBM bytes
            per stage) and valu_issue_frac (the time the VALU instructions
            take at one wave64 instruction per 4 cycles on each of the 1024
            SIMDs at 2.4 GHz, over the measured time) from the rocprofv3
            counter summaries under profiles/.
precision = relative L1 / Linf error per conserved variable of the FP32
            path against the plain FP64 path after 100 steps on the same
            dt sequence, with div B and floor hits (precision.py).
batch     = the batching form: 48 uneven discs advanced together, per RK
            stage one r-march launch over a tile work list of every disc and
            one update launch over every disc's cells, per-disc dt on the
            device, the step replayed from a hipGraph, launch shape
            autotuned; compared with the batch's scheme on one 512x512x256
            disc (single_disc_same_scheme).
cpu_baseline = the CPU restatement of the reference (the oracle: the
            reference's data layout and pass structure, OpenMP on all host
            cores) on the 256x256x128 disc of the FP64 parity runs (MC +
            HLLD, spiral 5 %), a bounded number of steps; same_grid times
            the GPU paths on that disc in the same run.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "171026-1-3dhlld-cyli-gal_amd")
sys.path[:0] = [PKG, os.path.join(ROOT, "tests")]

HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
METRIC = "cell-updates/s per RK3 step (FP32 vs FP64 path) + % HBM roofline, 1 MI355X"
BYTES_PER_CELL_UPDATE = {"fused32": 288.0, "fused64": 576.0, "plain64": 576.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_text_output(cm, ctx):
    """The text writers on the stepped state: one device-cut r-phi slice,
    the IDL slice set and the GDLOutputCylinder set written to a temporary
    directory, against bringing the whole state to the host."""
    import shutil
    import tempfile
    res = {}
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.slice(cm.Q_RHO, cm.PLANE_RPHI, 1)
    res["slice_rphi_ms"] = 1e3 * (time.perf_counter() - t0) / 3
    td = tempfile.mkdtemp(prefix="cylmhd_text_")
    try:
        t0 = time.perf_counter()
        ctx.idl_output(td, 1)
        res["idl_set_ms"] = 1e3 * (time.perf_counter() - t0)
        res["idl_set_bytes"] = sum(os.path.getsize(os.path.join(d, f)) for d, _, fs in os.walk(td) for f in fs)
        t0 = time.perf_counter()
        ctx.gdl_output(td, 1)
        res["gdl_set_ms"] = 1e3 * (time.perf_counter() - t0)
    finally:
        shutil.rmtree(td, ignore_errors=True)
    t0 = time.perf_counter()
    U, V = ctx.download()
    res["download_state_ms"] = 1e3 * (time.perf_counter() - t0)
    del U, V
    log(f"[bench] text output: slice {res['slice_rphi_ms']:.2f} ms, IDL set {res['idl_set_ms']:.0f} ms, "
        f"GDL set {res['gdl_set_ms']:.0f} ms, full download {res['download_state_ms']:.0f} ms")
    return res


def run_path(cm, np, path, grid, steps, warmup, config=3, text=False):
    p = cm.params_config(config, *grid)
    p.path = path
    t0 = time.perf_counter()
    U, V = cm.disc_state(p)
    log(f"[bench] {cm.PATH_NAMES[path]}: host init {time.perf_counter() - t0:.1f}s")
    ctx = cm.Context(p)
    ctx.set_initial(U, V)
    del U, V
    t = 0.0
    for _ in range(warmup):
        dt, ch, _ = ctx.setdt()
        ctx.gravity(t)
        ctx.step(dt, ch)
        t += dt
    ctx.sync()
    ctx.kernel_timing(0, reset=True)          # start timing the stage kernels
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt, ch, _ = ctx.setdt()
        ctx.gravity(t)
        ctx.step(dt, ch)
        t += dt
    ctx.sync()
    el = time.perf_counter() - t0
    march_ms, nmarch = ctx.kernel_timing(1)     # the two kernels of a split stage, each alone
    update_ms, _ = ctx.kernel_timing(2)
    kms, nlaunch = ctx.kernel_timing(0, reset=True)
    diag = ctx.diagnostics()
    counts = ctx.floor_counts()
    text_res = None
    if text:
        try:
            text_res = time_text_output(cm, ctx)
        except Exception as e:                     # reported, never fatal to the GPU number
            text_res = {"error": str(e)}
    ctx.close()
    ncell = p.ix * p.jy * p.kz
    name = cm.PATH_NAMES[path]
    res = dict(path=name, value=ncell * steps / el, ms_per_step=1e3 * el / steps, steps=steps,
               kernel_ms=kms, kernel_launches=nlaunch, interior_cells=ncell, t=t,
               march_ms=march_ms if nmarch else None, update_ms=update_ms if nmarch else None,
               mass=diag["mass"], energy=diag["energy"], divb_l1=diag["divb_l1"],
               pressure_floor_hits_last_step=int(counts[:, 1].sum()))
    if text_res is not None:
        res["text_output"] = text_res
    split = f" (r-march {march_ms:.3f} + update {update_ms:.3f})" if nmarch else ""
    log(f"[bench] {name}: {res['value']:.4g} cell-updates/s, {res['ms_per_step']:.2f} ms/step, "
        f"stage kernels {kms:.3f} ms{split} x {nlaunch}")
    return res


def run_batch(cm, steps, n=48, seed=20171026):
    """The batching form: n discs (dims from {32..128}, nz <= nr, beta
    [1, 100], spiral [0, 0.1]) advanced together; autotuned, then `steps`
    graph-replayed steps timed with HIP events."""
    ps = cm.params_batch(n, seed)
    cells = sum(p.ix * p.jy * p.kz for p in ps)
    t0 = time.perf_counter()
    b = cm.Batch(ps)
    for d, p in enumerate(ps):
        U, V = cm.disc_state(p)
        b.disc(d).set_initial(U, V)
    log(f"[bench] batch: {n} discs, {cells / 1e6:.2f} M cells, set-up {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    choice = b.tune(2)
    tune_s = time.perf_counter() - t0
    tiles = {}
    for d, p in enumerate(ps):
        sh, tj, tk, th = b.disc_tile(d)
        tiles.setdefault(f"{p.jy}x{p.kz}", f"{tj}x{tk}/{th}")
    b.run(2)                                   # warm-up (graph capture)
    ms = b.run(steps)
    st = [b.status(d) for d in range(n)]
    alive = sum(1 for x in st if x["alive"])
    done = sum(x["steps"] for x in st)
    b.close()
    updates = sum(p.ix * p.jy * p.kz * (x["steps"] - 2) for p, x in zip(ps, st))
    res = dict(discs=n, interior_cells=cells, steps=steps, ms=ms, value=updates / (ms * 1e-3),
               unit="cell-updates/s", alive=alive, steps_done=done,
               tuned={"common_tile": ["14x16/256", "6x32/256", "6x16/128", "10x16/192"][choice[0]],
                      "segment": choice[1] or "whole r", "grid": choice[2] or "resident blocks", "kernels_per_stage": choice[3],
                      "tile_by_class_nphi_x_nz": tiles, "tune_s": tune_s})
    log(f"[bench] batch: {res['value']:.4g} cell-updates/s over {steps} steps ({ms:.1f} ms), tuned {res['tuned']}")
    # the batch's scheme (minmod + HLLD, the batch discs' parameters) on one
    # 512x512x256 disc, the single-disc rate the batch is compared with
    import numpy as np
    one = run_path(cm, np, cm.PATH_FUSED32, (512, 512, 256), 6, 1, config=4)
    res["single_disc_same_scheme"] = dict(grid=[512, 512, 256], value=one["value"], ms_per_step=one["ms_per_step"])
    res["over_single_disc"] = res["value"] / one["value"]
    return res


def run_large(cm, np, steps):
    """The large disc (768x768x384, ~226 M cells) with every source active:
    power-law cooling Lambda ~ T^0.5 above T = 1e-2, spiral m=2 at 10 %,
    GLM alpha = 0.1, MC + HLLD.  FP32 and FP64 fused paths on the same grid
    (the plain pass-by-pass FP64 path does not fit one GPU's memory here),
    with the FP32 stage kernel's HBM roofline fraction and the FP32 error
    against FP64 after the same steps on the same dt sequence."""
    import precision
    grid = (768, 768, 384)
    r32 = run_path(cm, np, cm.PATH_FUSED32, grid, steps, 1, config=5)
    r64 = run_path(cm, np, cm.PATH_FUSED64, grid, steps, 1, config=5)
    ncell = r32["interior_cells"]
    bytes_launch = BYTES_PER_CELL_UPDATE["fused32"] * ncell / 3.0
    achieved = bytes_launch / (r32["kernel_ms"] * 1e-3)
    prec = precision.compare(cm.params_config(5, *grid), steps, cm.PATH_FUSED64, cm.PATH_FUSED32, log=log)
    keep = ("l1_max", "physical", "divb_l1", "mass_rel", "energy_rel", "floor_hits")
    return dict(grid=list(grid), interior_cells=ncell, steps=steps,
                fused32={k: r32[k] for k in ("value", "ms_per_step", "kernel_ms", "march_ms", "update_ms")},
                fused64={k: r64[k] for k in ("value", "ms_per_step", "kernel_ms", "march_ms", "update_ms")},
                fp32_over_fp64=r32["value"] / r64["value"],
                roofline=dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                              frac=achieved / HBM_PEAK, kernel="stage: k_march<float,MC,HLLD> + k_update<float>"),
                precision={k: prec[k] for k in keep if k in prec})


CPU_GRID = (256, 256, 128)


def cpu_baseline(cm, np, budget_s=15.0):
    """The CPU restatement of the reference on the 256x256x128 disc (MC +
    HLLD, spiral m = 2 at 5 %), FP64 in the reference's layout: one warm-up
    step, then as many whole steps as fit the time budget (1 to 20)."""
    from oracle_py import Oracle
    p = cm.params_config(2, *CPU_GRID)
    U, V = cm.disc_state(p)
    o = Oracle(p)
    o.set_initial(U, V)
    del U, V
    threads = o.L.orc_threads()
    t, el = o.run(1)                          # first step also warms the caches
    n = max(1, min(20, int(budget_s / max(el, 1e-3))))
    t, el = o.run(n, t)
    ncell = p.ix * p.jy * p.kz
    o.close()
    log(f"[bench] cpu oracle: {ncell * n / el:.4g} cell-updates/s on {threads} threads ({n} steps, {el:.1f} s)")
    return dict(value=ncell * n / el, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"256x256x128 disc (FP64, MC+HLLD, spiral 5%, RK3; the FP64 parity disc), {n} steps after "
                       f"1 warm-up step, oracle/liboracle.so (reference layout and pass structure) with OpenMP "
                       f"on {threads} threads")


def same_grid(cm, np, cpu):
    """The GPU paths on the CPU baseline's disc, in the same run."""
    res = {"grid": list(CPU_GRID), "cpu_oracle": cpu.get("value")}
    for path in (cm.PATH_PLAIN64, cm.PATH_FUSED64, cm.PATH_FUSED32):
        r = run_path(cm, np, path, CPU_GRID, 10, 1, config=2)
        res[cm.PATH_NAMES[path]] = r["value"]
        if cpu.get("value"):
            res[cm.PATH_NAMES[path] + "_over_cpu"] = r["value"] / cpu["value"]
    return res


def counters(grid):
    """Per-stage HBM traffic and VALU instructions of the FP32 stage kernels
    (the r-march and the per-cell update, summed over the two, and each
    alone under "kernels") from the rocprofv3 summaries under profiles/
    (None where absent)."""
    out = {"kernels": {}}
    for key, name in (("traffic", "traffic"), ("valu", "counters")):
        f = os.path.join(ROOT, "profiles", f"{name}_fused32_{grid[0]}x{grid[1]}x{grid[2]}.json")
        if os.path.isfile(f):
            with open(f) as fh:
                d = json.load(fh)
            field = "bytes_per_launch" if key == "traffic" else "sq_insts_valu_per_launch"
            out[key] = d.get(field)
            for kn, kd in d.get("kernels", {}).items():
                out["kernels"].setdefault(kn, {})[key] = kd.get(field)
    return out


def valu_floor_ms(valu):
    """VALU issue floor: one wave64 VALU instruction holds a SIMD for 4
    cycles (FP32 unpacked and FP64 alike); 256 CUs x 4 SIMDs at 2.4 GHz."""
    return valu * 4.0 / (1024 * 2.4e9) * 1e3 if valu else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, nargs=3, default=[512, 512, 256])
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison paths")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--precision-steps", type=int, default=100,
                    help="steps of the FP32-vs-FP64 error report (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=200,
                    help="steps of the 48-disc batching form (0: skip)")
    ap.add_argument("--large-steps", type=int, default=50,
                    help="steps of the 768x768x384 disc with every source active, FP32 and FP64 "
                         "fused paths (0: skip)")
    ap.add_argument("--no-text", action="store_true", help="skip timing the text writers")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (the calibration "
                         "kernel of the FETCH_SIZE / WRITE_SIZE counter runs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {a.gpus}); the path would shard as "
              f"phi-slabs with one-cell halo exchanges per stage", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import cylmhd as cm

    grid = tuple(a.grid)
    if a.copy_mib > 0:
        import torch
        n = a.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: 3 x {a.copy_mib} MiB device-to-device")
        del src, dst
    main_res = run_path(cm, np, cm.PATH_FUSED32, grid, a.steps, a.warmup, text=not a.no_text)
    fp64 = {}
    if not a.no_fp64:
        k64 = max(2, a.steps // 2)
        for path in (cm.PATH_FUSED64, cm.PATH_PLAIN64):
            fp64[cm.PATH_NAMES[path]] = run_path(cm, np, path, grid, k64, 1)
    ncell = main_res["interior_cells"]
    bytes_launch = BYTES_PER_CELL_UPDATE["fused32"] * ncell / 3.0
    achieved = bytes_launch / (main_res["kernel_ms"] * 1e-3) if main_res["kernel_ms"] > 0 else None
    cnt = counters(grid)
    traffic = cnt.get("traffic")
    valu = cnt.get("valu")
    vfloor = valu_floor_ms(valu)
    kms = main_res["kernel_ms"]
    # the stage runs as two kernels: the r-march (states, faces, divergence)
    # and the per-cell update; the roofline is that of the pair per stage
    kernels = {}
    for kn, ms in (("k_march", main_res["march_ms"]), ("k_update", main_res["update_ms"])):
        kc = cnt["kernels"].get(kn, {})
        kf = valu_floor_ms(kc.get("valu"))
        kernels[kn] = dict(ms=ms, traffic=kc.get("traffic"), valu_insts_per_launch=kc.get("valu"),
                           valu_issue_floor_ms=kf, valu_issue_frac=kf / ms if kf and ms else None,
                           traffic_GBps=kc["traffic"] / (ms * 1e-3) / 1e9 if kc.get("traffic") and ms else None)
    roofline = dict(bound="hbm", achieved=achieved / 1e9 if achieved else None, peak=HBM_PEAK / 1e9, unit="GB/s",
                    frac=achieved / HBM_PEAK if achieved else None, traffic=traffic,
                    kernel="stage: k_march<float,MC,HLLD> + k_update<float>",
                    algorithmic_bytes_per_launch=bytes_launch, kernel_ms=kms,
                    valu_insts_per_launch=valu, valu_issue_floor_ms=vfloor,
                    valu_issue_frac=vfloor / kms if vfloor and kms else None, kernels=kernels)
    out = {
        "metric": METRIC,
        "value": main_res["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"disc {grid[0]}x{grid[1]}x{grid[2]}: FP32 state + FP64 recovery, MC+HLLD, RK3, "
                               "spiral m=2 5%, beta log-uniform [0.1,100] per annulus",
                   "grid": list(grid), "interior_cells": ncell, "path": "fused32"},
        "roofline": roofline,
    }
    if "text_output" in main_res:
        out["text_output"] = main_res["text_output"]
    if fp64:
        out["fp64"] = {k: {kk: v[kk] for kk in ("value", "ms_per_step", "kernel_ms", "march_ms", "update_ms", "steps")}
                       for k, v in fp64.items()}
        out["fp32_over_fp64"] = {k: main_res["value"] / v["value"] for k, v in fp64.items()}
    if a.precision_steps > 0:
        import precision
        try:
            out["precision"] = precision.compare(cm.params_config(3, *grid), a.precision_steps,
                                                 cm.PATH_PLAIN64, cm.PATH_FUSED32, log=log)
        except Exception as e:                     # reported, never fatal to the GPU number
            out["precision"] = {"error": str(e)}
    if a.large_steps > 0:
        try:
            out["large"] = run_large(cm, np, a.large_steps)
        except Exception as e:                     # reported, never fatal to the GPU number
            out["large"] = {"error": str(e)}
    if a.batch_steps > 0:
        try:
            out["batch"] = run_batch(cm, a.batch_steps)
        except Exception as e:                     # reported, never fatal to the GPU number
            out["batch"] = {"error": str(e)}
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cm, np)
            out["same_grid"] = same_grid(cm, np, out["cpu_baseline"])
        except Exception as e:                     # reported, never fatal to the GPU number
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
