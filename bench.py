#!/usr/bin/env python3
"""Throughput of the OpenCFD-SC right-hand side + RK3 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload iso256]
                    [--precision fp64|mixed] [--cpu-baseline auto|reference|port|none]

A step is one full RK3 time step (three right-hand-side evaluations, stage
updates, boundary refresh, filtering when configured) of one block already
resident in HBM.  Prints one JSON line:

  value      Mcell-RK3-steps/s over the K timed steps
  mixed / fp64  the other precision form on the same workload, run from the
             same initial field for the same number of steps: its throughput,
             the ratio of the two, and the per-field relative L2 difference
             of the final primitives (mixed against FP64)
  roofline   the dominant kernel (largest share of the timed region, timed
             with HIP events on the library's stream) placed on its roofline:
             inviscid sweeps by algorithmic FP64 FLOPs (exact counts of the
             reference algorithm, ocfd/perfmodel.py) against 78.6 TFLOP/s,
             memory-bound kernels by compulsory HBM bytes against 8 TB/s
  cpu_baseline  the compiled reference solver under MPI on this host's cores
             on a bounded sample of the same workload (falls back to the
             single-threaded C oracle when the reference build is absent)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opencfd-sc_amd"))

from ocfd import cases, native, perfmodel  # noqa: E402
from ocfd.config import METRIC_FIELD, METRIC_UNIFORM, PREC_FP64, PREC_MIXED  # noqa: E402

METRIC = "Mcell-RK3-steps/s, FP64 vs mixed FP32/FP64 at stated L2 error; HBM GB/s"
HYB_THRESHOLDS = (0.01, 0.05)    # Hybrid_para(2:3): UD7L below, WENO7 between, WENO5 above
SPINUP_STEPS = 100               # untimed steps that turn the seeded field into developed turbulence


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(workload: str, n_override=None):
    """Returns (case, description dict, FLOP-model variant)."""
    if workload.startswith("iso"):
        n = n_override or int(workload[3:].split("-")[0] or 256)
        scheme = "WENO7" if workload.endswith("-weno7") else "HYBRID"
        case = cases.isotropic(n, scheme=scheme, split=2, char=1, hybrid_thresholds=HYB_THRESHOLDS,
                               metric_mode=METRIC_UNIFORM)
        desc = {
            "workload": "decaying compressible isotropic turbulence %d^3 periodic" % n,
            "scheme": ("characteristic hybrid UD7L/WENO7/WENO5 (sensor 1, thresholds %g/%g)" % HYB_THRESHOLDS
                       if scheme == "HYBRID" else "characteristic WENO7"),
            "splitting": "local Lax-Friedrichs", "viscous": "CD6", "time": "RK3",
            "grid": "uniform box, exact metrics", "Mt": 0.5, "Re": round(case.cfg.re, 3),
            "k0": 8, "seed": case.info["seed"], "cells": n ** 3,
        }
        variant = "HYBRID/LLF/char" if scheme == "HYBRID" else "WENO7/LLF/char"
        return case, desc, variant
    if workload.startswith("tgv"):
        n = n_override or int(workload[3:] or 64)
        case = cases.tgv(n, metric_mode=METRIC_FIELD)
        desc = {"workload": "Taylor-Green vortex %d^3 periodic" % n, "scheme": "WENO5 Steger-Warming",
                "viscous": "CD6", "time": "RK3 CFL 0.5", "grid": "reference box generator",
                "Ma": 0.1, "Re": 1600, "cells": n ** 3}
        return case, desc, "WENO5/SW/nonchar"
    raise SystemExit("unknown workload %s" % workload)


def run_block(case, precision, spinup, warmup, steps, prof_steps, autotune=True):
    """Set up one block, develop the field, warm up, time `steps` whole steps
    (replayed from the captured graph), then run `prof_steps` more steps with
    every kernel bracketed by HIP events on the library's stream (eager
    launches) for the per-kernel times.
    Returns (block, elapsed seconds of the timed steps, per-kernel profile)."""
    b = native.setup_block(native.Block(case.cfg.replace(precision=precision)), case)
    if spinup:
        t1 = time.time()
        b.step(spinup)
        b.synchronize()
        log("spin-up (%s): %d steps in %.1f s" % ("mixed" if precision == PREC_MIXED else "fp64", spinup,
                                                 time.time() - t1))
    if autotune:
        b.tiles = b.autotune()
        log("autotuned sweep tiles (x, y, z): %s" % b.tiles)
    b.step(warmup)
    b.synchronize()
    t = time.perf_counter()
    b.step(steps)
    b.synchronize()
    elapsed = time.perf_counter() - t
    prof = {}
    if prof_steps:
        b.profile_reset()
        b.profile_enable(True)
        b.step(prof_steps)
        b.synchronize()
        prof = b.profile_read()
        b.profile_enable(False)
    return b, elapsed, prof


def compare_forms(bm, b64):
    """Relative L2 of the mixed-precision primitives against FP64, and of the
    kinetic energy / enstrophy (ana_box)."""
    import numpy as np
    pm, p64 = bm.download_primitives(), b64.download_primitives()
    L = 4
    out = {}
    for q, nm in enumerate("duvwT"):
        a = pm[q][L:-L, L:-L, L:-L]
        r = p64[q][L:-L, L:-L, L:-L]
        out[nm] = float("%.3e" % (np.linalg.norm(a - r) / np.linalg.norm(r)))
    am, a64 = bm.ana_box(), b64.ana_box()
    return out, float("%.3e" % (abs(am[12] - a64[12]) / abs(a64[12]))), \
        float("%.3e" % (abs(am[13] - a64[13]) / abs(a64[13])))


# ---------------------------------------------------------------------------
# batched ensemble of independent periodic boxes
# ---------------------------------------------------------------------------
BATCH_SIZES = (32, 48, 64, 96, 128)


def batch_cases(nbox: int, precision: int, seed: int = 5):
    """nbox boxes: n^3 with n uniform over BATCH_SIZES, a Taylor-Green or a
    seeded isotropic initial field (seed 20261202 + box), Ma uniform in
    [0.1, 0.5], Re uniform in [400, 3200], WENO5 or OMP6 with Steger-Warming
    splitting (component-wise), CD6, CFL 0.5, uniform-box metrics."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nbox):
        n = int(rng.choice(BATCH_SIZES))
        kind = "tgv" if rng.random() < 0.5 else "iso"
        ma = float(rng.uniform(0.1, 0.5))
        re = float(rng.uniform(400.0, 3200.0))
        sch = "WENO5" if rng.random() < 0.5 else "OMP6"
        if kind == "tgv":
            c = cases.tgv(n, ma=ma, re=re, scheme=sch, split=1, metric_mode=METRIC_UNIFORM)
        else:
            c = cases.isotropic(n, seed=20261202 + i, mt=ma, re=re, scheme=sch, split=1, char=0,
                                metric_mode=METRIC_UNIFORM)
        c.cfg = c.cfg.replace(precision=precision)
        out.append(c)
    return out


def run_batch(args):
    nbox = int(args.workload[5:] or 96)
    prec = PREC_MIXED if args.precision == "mixed" else PREC_FP64
    t0 = time.time()
    cs = batch_cases(nbox, prec)
    cells = sum(c.cfg.nx * c.cfg.ny * c.cfg.nz for c in cs)
    log("%d boxes (%.1f Mcells) built in %.1f s" % (nbox, cells / 1e6, time.time() - t0))
    with native.Batch([c.cfg for c in cs]) as batch:
        for i, c in enumerate(cs):
            native.setup_block(batch.box(i), c, use_uniform=True)
        tiles = batch.autotune() if args.autotune else None
        batch.step(args.warmup)
        t = time.perf_counter()
        batch.step(args.steps)
        el_b = time.perf_counter() - t
        # the same boxes advanced one at a time, each through its own handle
        # (one captured graph per box)
        for bx in batch.boxes:
            bx.step(args.warmup)
        t = time.perf_counter()
        for bx in batch.boxes:
            bx.step(args.steps)
        el_s = time.perf_counter() - t
    value = cells * args.steps / el_b / 1e6
    single = cells * args.steps / el_s / 1e6
    log("batched: %.2f ms/step, %.1f Mcell-RK3-steps/s; one at a time: %.1f (x%.2f)" % (
        el_b / args.steps * 1e3, value, single, el_s / el_b))
    sizes = [c.cfg.nx for c in cs]
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el_b / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded initial fields generated on the host)",
        "config": {"workload": "batched ensemble of %d independent periodic boxes" % nbox,
                   "box_sizes": {str(n): sizes.count(n) for n in BATCH_SIZES}, "cells": cells,
                   "initial": "Taylor-Green or seeded isotropic per box; Ma U[0.1,0.5], Re U[400,3200]",
                   "scheme": "WENO5 or OMP6 per box, Steger-Warming, CD6, RK3 CFL 0.5",
                   "precision": "FP64" if prec == PREC_FP64 else "mixed",
                   "execution": "one launch per kernel phase over all boxes, graph-captured steps",
                   "autotuned": tiles},
        "one_at_a_time": {"value": round(single, 3), "ms_per_step": round(el_s / args.steps * 1e3, 4),
                          "batched_speedup": round(el_s / el_b, 4)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


# kernel symbols behind the library's profile names
KERNEL_SYMBOL = {"stage_update": "k_visc_div_update", "visc_flux": "k_visc_flux", "halo_fill": "k_halo",
                 "inviscid_x": "k_inviscid<0", "inviscid_y": "k_inviscid<1", "inviscid_z": "k_inviscid<2",
                 "inviscid_char_x": "k_inviscid<0", "inviscid_char_y": "k_inviscid<1",
                 "inviscid_char_z": "k_inviscid<2", "weno_faces": "k_weno_faces", "sensor": "k_sensor1",
                 "hyb_list": "k_hyb_list"}


def profile_traffic(workload: str, precision: str, name: str):
    """HBM bytes per launch of kernel `name` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed for this workload and precision
    (profiles/<workload>_<precision>_kernels.md, counters scaled by the copy
    kernel), or None when there is no such summary."""
    path = os.path.join(ROOT, "profiles", "%s_%s_kernels.md" % (workload, precision))
    sym = KERNEL_SYMBOL.get(name)
    if sym is None or not os.path.exists(path):
        return None
    best = None
    for line in open(path):
        if line.startswith("Occupancy and issue"):
            break  # the per-kernel HBM table ends where the occupancy table starts
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) < 7 or not cols[0].startswith("`" + sym):
            continue
        try:
            total, rd, wr = float(cols[3]), float(cols[5]), float(cols[6])
        except ValueError:
            continue
        if best is None or total > best[0]:
            best = (total, (rd + wr) * 1e6)
    return best[1] if best else None


def hybrid_mix(b):
    b.exchange_primitives()
    c = b.comput_rhybrid(counts=True)
    return [c[0] / c[3], c[1] / c[3], c[2] / c[3]]


# ---------------------------------------------------------------------------
def cpu_baseline(workload: str, mode: str):
    """Time the CPU path on a bounded sample; returns the cpu_baseline dict."""
    if mode == "none":
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    n = 128 if workload.startswith("iso") else 64
    steps = 4
    case, _, _ = build_case(workload, n_override=n)
    case.cfg.metric_mode = METRIC_FIELD
    case.cfg.grid_type = cases.GRID_BOX
    if mode in ("auto", "reference") and pyoracle.reference_available():
        ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        p = 16 if ncpu >= 16 else (8 if ncpu >= 8 else (4 if ncpu >= 4 else 1))
        dec = {16: (4, 2, 2), 8: (2, 2, 2), 4: (2, 2, 1), 1: (1, 1, 1)}[p]
        with tempfile.TemporaryDirectory(prefix="ocfd_cpu_") as d:
            try:
                wall, _, _ = pyoracle.run_reference(case, d, steps, nranks=p, decomposition=dec, timeout=900)
            except Exception as e:  # reference could not run here: fall back to the oracle
                log("reference CPU run failed (%s); timing the oracle instead" % str(e)[:200])
                wall = None
        if wall:
            return {"value": round(n ** 3 * steps / wall / 1e6, 4), "unit": "Mcell-RK3-steps/s", "cores": p,
                    "kind": "reference",
                    "sample": "reference Fortran (amdflang -O3, MPICH) on %d ranks %s, %d^3 block of the same "
                              "case/scheme, %d RK3 steps, its own 'total wtime cost' timer" % (p, dec, n, steps)}
    if mode == "reference":
        return None
    n = 64
    steps = 2
    case, _, _ = build_case(workload, n_override=n)
    case.cfg.metric_mode = METRIC_FIELD
    ob = pyoracle.setup_oracle(case, uniform=False)
    t = time.perf_counter()
    ob.step(steps)
    wall = time.perf_counter() - t
    return {"value": round(n ** 3 * steps / wall / 1e6, 4), "unit": "Mcell-RK3-steps/s", "cores": 1,
            "kind": "port", "sample": "C restatement (gcc -O2, 1 thread), %d^3 block, %d steps" % (n, steps)}


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--spinup", type=int, default=-1,
                    help="untimed steps before warm-up that develop the initial field (default %d for the "
                         "isotropic workloads, 0 otherwise)" % SPINUP_STEPS)
    ap.add_argument("--workload", default="iso256")
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick the sweep tile shapes with the library's autotuner after the spin-up (untimed)")
    ap.add_argument("--prof-steps", type=int, default=3,
                    help="steps after the timed region with per-kernel HIP events (roofline, kernel table)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--other-form", type=int, default=1,
                    help="also run the other precision form on the same workload and report it (default 1)")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after the timed region, launch the plain fn=f copy kernel (5 planes of FP64 "
                         "read and written) this many times: the reference for scaling HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one MI355X: every box in this setup has one GPU, so --gpus %d is not "
            "supported (the multi-GPU halo-exchange path is design-only)" % args.gpus)
        sys.exit(2)

    if args.workload.startswith("batch"):
        return run_batch(args)
    t0 = time.time()
    case, desc, variant = build_case(args.workload)
    cfg = case.cfg
    n = cfg.nx
    cells = cfg.nx * cfg.ny * cfg.nz
    log("case %s built in %.1f s (dt=%.4g)" % (case.name, time.time() - t0, cfg.dt))
    prec = PREC_MIXED if args.precision == "mixed" else PREC_FP64
    spinup = args.spinup if args.spinup >= 0 else (SPINUP_STEPS if args.workload.startswith("iso") else 0)
    if spinup:
        desc["spinup_steps"] = spinup
    b, elapsed, prof = run_block(case, prec, spinup, args.warmup, args.steps, args.prof_steps, args.autotune)
    mix = hybrid_mix(b) if cfg.scheme_invis == "HYBRID" else None
    value = cells * args.steps / elapsed / 1e6
    log("timed %d steps: %.3f ms/step, %.1f Mcell-RK3-steps/s" % (args.steps, elapsed / args.steps * 1e3, value))
    for k, (ms, cnt) in sorted(prof.items(), key=lambda x: -x[1][0]):
        log("  %-18s %9.3f ms  %5d launches  %8.4f ms/launch" % (k, ms, cnt, ms / cnt))
    if not prof:
        raise SystemExit("--prof-steps must be >= 1: the roofline needs per-kernel times")

    # dominant kernel on its roofline
    name, (ms, cnt) = max(prof.items(), key=lambda x: x[1][0])
    avg_s = ms / cnt / 1e3
    uniform = cfg.metric_mode == METRIC_UNIFORM
    if name.startswith("inviscid"):
        flops = perfmodel.inviscid_axis_flops(variant, n, mix)
        ach = flops / avg_s / 1e12
        peak = perfmodel.PEAK_FP64_TFLOPS if prec == PREC_FP64 else perfmodel.PEAK_FP32_TFLOPS
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(ach / peak, 4), "traffic": None,
                "kernel": name, "algorithmic_per_launch": flops,
                "note": ("FP64 vector rate (equal to the MI355X dense FP64 matrix rate)" if prec == PREC_FP64 else
                         "FP32 vector rate (157.3 TFLOP/s, no packed FP32)")
                        + "; FLOPs are exact counts of the reference algorithm"}
    else:
        byts = perfmodel.algorithmic_bytes(name, cells, uniform, viscous=cfg.if_viscous == 1,
                                           rhs_bytes=4 if prec == PREC_MIXED else 8)
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": perfmodel.PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / perfmodel.PEAK_HBM_GBS, 4), "traffic": None, "kernel": name,
                "algorithmic_per_launch": byts}
    roof["avg_launch_ms"] = round(ms / cnt, 5)
    roof["traffic"] = profile_traffic(args.workload, args.precision, name)

    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded initial field generated on the host)", "config": desc,
        "roofline": roof,
    }
    if mix:
        res["config"]["face_mix_linear_weno7_weno5"] = [round(x, 4) for x in mix]
    if args.autotune:
        res["config"]["sweep_tiles_xyz"] = ["%dx%d" % t for t in
                                            [((64, 32, 128)[s], 256 // (64, 32, 128)[s]) if a == 0 else
                                             ((32, 16, 64)[s], 256 // (32, 16, 64)[s])
                                             for a, s in enumerate(b.tiles)]]
    res["config"]["precision"] = ("FP64" if prec == PREC_FP64 else
                                  "mixed: FP32 splitting, reconstruction, viscous derivatives, du and stored "
                                  "primitives; FP64 conserved variables, RK3, metrics and the characteristic "
                                  "basis and projections of the sensor-flagged faces")
    res["kernels_ms_per_step"] = {k: round(v[0] / args.prof_steps, 4) for k, v in prof.items()}
    if args.calibrate_copy:
        # k_copy moves exactly 5*cells*8 bytes in and 5*cells*8 bytes out per launch
        for _ in range(args.calibrate_copy):
            b.save_fn()
        b.synchronize()
        res["copy_kernel_bytes_each_way"] = 5 * cells * 8
    if args.other_form:
        other = PREC_FP64 if prec == PREC_MIXED else PREC_MIXED
        bo, el_o, _ = run_block(case, other, spinup, args.warmup, args.steps, args.prof_steps, args.autotune)
        bm, b64 = (b, bo) if prec == PREC_MIXED else (bo, b)
        el_m, el_64 = (elapsed, el_o) if prec == PREC_MIXED else (el_o, elapsed)
        errs, ke, en = compare_forms(bm, b64)
        key = "fp64" if prec == PREC_MIXED else "mixed"
        res[key] = {"value": round(cells * args.steps / el_o / 1e6, 3),
                    "ms_per_step": round(el_o / args.steps * 1e3, 4),
                    "mixed_speedup_over_fp64": round(el_64 / el_m, 4),
                    "steps_compared": spinup + args.warmup + args.steps + args.prof_steps,
                    "rel_l2_mixed_vs_fp64": errs, "kinetic_energy_rel": ke, "enstrophy_rel": en}
        log("%s: %.1f Mcell-RK3-steps/s, mixed/fp64 speedup %.3f, relative L2 %s" % (
            key, res[key]["value"], res[key]["mixed_speedup_over_fp64"], errs))
        bo.close()
    b.close()
    res["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_baseline)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
