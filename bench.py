#!/usr/bin/env python3
"""Throughput of the OpenCFD-SC right-hand side + RK3 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload iso256]
                    [--precision fp64] [--cpu-baseline auto|reference|port|none]

A step is one full RK3 time step (three right-hand-side evaluations, stage
updates, boundary refresh, filtering when configured) of one block already
resident in HBM.  Prints one JSON line:

  value      Mcell-RK3-steps/s over the K timed steps
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
from ocfd.config import METRIC_FIELD, METRIC_UNIFORM  # noqa: E402

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
    ap.add_argument("--precision", default="fp64", choices=["fp64"])
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after the timed region, launch the plain fn=f copy kernel (5 planes of FP64 "
                         "read and written) this many times: the reference for scaling HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one MI355X: every box in this setup has one GPU, so --gpus %d is not "
            "supported (the multi-GPU halo-exchange path is design-only)" % args.gpus)
        sys.exit(2)

    t0 = time.time()
    case, desc, variant = build_case(args.workload)
    cfg = case.cfg
    n = cfg.nx
    cells = cfg.nx * cfg.ny * cfg.nz
    log("case %s built in %.1f s (dt=%.4g)" % (case.name, time.time() - t0, cfg.dt))
    b = native.setup_block(native.Block(cfg), case)
    spinup = args.spinup if args.spinup >= 0 else (SPINUP_STEPS if args.workload.startswith("iso") else 0)
    if spinup:
        t1 = time.time()
        b.step(spinup)
        b.synchronize()
        log("spin-up: %d steps in %.1f s" % (spinup, time.time() - t1))
        desc["spinup_steps"] = spinup
    b.step(args.warmup)
    b.synchronize()
    mix0 = hybrid_mix(b) if cfg.scheme_invis == "HYBRID" else None
    b.profile_reset()
    b.profile_enable(True)
    b.synchronize()
    t = time.perf_counter()
    b.step(args.steps)
    b.synchronize()
    elapsed = time.perf_counter() - t
    prof = b.profile_read()
    b.profile_enable(False)
    mix1 = hybrid_mix(b) if cfg.scheme_invis == "HYBRID" else None
    mix = [(a + c) / 2 for a, c in zip(mix0, mix1)] if mix0 else None
    value = cells * args.steps / elapsed / 1e6
    log("timed %d steps: %.3f ms/step, %.1f Mcell-RK3-steps/s" % (args.steps, elapsed / args.steps * 1e3, value))
    for k, (ms, cnt) in sorted(prof.items(), key=lambda x: -x[1][0]):
        log("  %-18s %9.3f ms  %5d launches  %8.4f ms/launch" % (k, ms, cnt, ms / cnt))

    # dominant kernel on its roofline
    name, (ms, cnt) = max(prof.items(), key=lambda x: x[1][0])
    avg_s = ms / cnt / 1e3
    uniform = cfg.metric_mode == METRIC_UNIFORM
    if name.startswith("inviscid"):
        flops = perfmodel.inviscid_axis_flops(variant, n, mix)
        ach = flops / avg_s / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": perfmodel.PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / perfmodel.PEAK_FP64_TFLOPS, 4), "traffic": None,
                "kernel": name, "algorithmic_per_launch": flops,
                "note": "FP64 vector rate (equal to the MI355X dense FP64 matrix rate); FLOPs are exact "
                        "counts of the reference algorithm"}
    else:
        byts = perfmodel.algorithmic_bytes(name, cells, uniform, viscous=cfg.if_viscous == 1)
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": perfmodel.PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / perfmodel.PEAK_HBM_GBS, 4), "traffic": None, "kernel": name,
                "algorithmic_per_launch": byts}
    roof["avg_launch_ms"] = round(ms / cnt, 5)

    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded initial field generated on the host)", "config": desc,
        "roofline": roof,
    }
    if mix:
        res["config"]["face_mix_linear_weno7_weno5"] = [round(x, 4) for x in mix]
    res["kernels_ms_per_step"] = {k: round(v[0] / args.steps, 4) for k, v in prof.items()}
    if args.calibrate_copy:
        # k_copy moves exactly 5*cells*8 bytes in and 5*cells*8 bytes out per launch
        for _ in range(args.calibrate_copy):
            b.save_fn()
        b.synchronize()
        res["copy_kernel_bytes_each_way"] = 5 * cells * 8
    b.close()
    res["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_baseline)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
