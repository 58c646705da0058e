#!/usr/bin/env python3
"""Throughput of the OpenCFD-SC right-hand side + RK3 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload iso512|iso256|tgv64|batch96]
                    [--precision fp64|mixed] [--cpu-baseline auto|reference|port|none] [--dropin 0|1]

A step is one full RK3 time step (three right-hand-side evaluations, stage
updates, boundary refresh, filtering when configured) of one block already
resident in HBM.  The default workload is decaying compressible isotropic
turbulence on a 512^3 periodic box (characteristic hybrid UD7L/WENO7/WENO5
with local Lax-Friedrichs splitting, CD6 viscous terms), developed from the
seeded initial field to the physical time that 200 steps of the 256^3 run
reach (t = 200 dt(256^3), 400 steps at 512^3), so that the shock sensor
flags WENO faces.  Both precision forms start from the same initial field and
run the same steps.  Prints one JSON line:

  value      FP64 Mcell-RK3-steps/s over the K timed steps
  mixed      the mixed-precision form on the same workload: its throughput,
             the ratio to FP64, the per-field relative L2 difference of the
             final primitives and the largest relative difference of the
             kinetic energy and enstrophy histories (ana_box every 50 steps)
  roofline   the dominant kernel of the FP64 step (largest share, timed with
             HIP events on the library's stream): compulsory bytes per launch
             (120 B per cell and stage for the stage update: the conserved
             variables read, the stage-start values read and the result
             written, 3 x 5 FP64 values) against 8 TB/s, or exact algorithmic
             FLOPs of the reference algorithm (ocfd/perfmodel.py) against
             the FP64 vector rate; `design_bytes_frac` prices the same launch
             at every array this design reads and writes once
  step_roofline  the whole step: 360 B per cell-step against 8 TB/s and the
             exact FLOPs per cell-step at the measured face mix against
             78.6 TFLOP/s
  step_hbm_measured  the whole step's HBM bytes as rocprofv3 counted them in
             the committed summary of this workload, over this run's time
             per step, against 8 TB/s
  dropin     the reference's own driver (NS_solver) with this library as its
             hot path, on 256^3 isotropic turbulence with field metrics,
             timed by the reference's show_flow_msg timer, beside
             ocfd_step on the same case (--dropin 0 leaves it out; it runs
             mpiexec and must stay out of profiler runs)
  cpu_baseline  the compiled reference solver under MPI on this host's cores
             (median of 3) and on one rank, on a field developed on the GPU
             to the same physical time (falls back to the single-threaded C
             oracle when the reference build is absent)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
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
SPINUP_STEPS_256 = 200           # the 256^3 run length; the isotropic workloads develop to its end time
HISTORY_EVERY = 50               # ana_box sampling interval of the spin-up (steps)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def trace_marker(b, on: bool):
    """One launch of the CFL reduction kernel (k_cfl_rate), which the bench
    launches nowhere else: in a rocprofv3 trace or counter run, the dispatches
    after it are the timed and profiled steps of the developed field."""
    if on:
        b.cfl_dt(0.5)
        b.synchronize()


def build_case(workload: str, n_override=None):
    """Returns (case, description dict, FLOP-model variant)."""
    if workload.startswith("iso"):
        n = n_override or int(workload[3:].split("-")[0] or 512)
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


DROPIN_N = 256
DROPIN_STEPS = 20


def dropin_leg(n=DROPIN_N, steps=DROPIN_STEPS):
    """The reference's own driver with the hot path served by this library
    (opencfd-sc_amd/fortran: NS_solver, one MPI rank, state on the device)
    against ocfd_step on the same n^3 isotropic case: hybrid characteristic
    LLF in FP64 with the field metrics the reference's init_mesh computes
    (the drop-in's metric form).  The drop-in runs 2 x `steps` steps and is
    timed by the reference's own show_flow_msg timer over the second block
    of `steps`, once with stage fusion (its default) and once with one kernel
    sequence per routine."""
    import shutil
    from ocfd import dropin
    if not dropin.available():
        return {"skipped": "drop-in executable or mpiexec missing"}
    case = cases.isotropic(n, scheme="HYBRID", split=2, char=1, hybrid_thresholds=HYB_THRESHOLDS,
                           metric_mode=METRIC_FIELD)
    cells = case.cfg.nx * case.cfg.ny * case.cfg.nz
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.step(steps)
        b.synchronize()
        t = time.perf_counter()
        b.step(steps)
        b.synchronize()
        lib_s = time.perf_counter() - t
    out = {"workload": "decaying compressible isotropic turbulence %d^3, hybrid characteristic LLF, FP64, "
                       "field metrics, initial field" % n,
           "steps_timed": steps, "ocfd_step": round(cells * steps / lib_s / 1e6, 3)}
    log("drop-in leg: ocfd_step %.1f Mcell-RK3-steps/s on %d^3 field metrics" % (out["ocfd_step"], n))
    for key, env in (("dropin", {}), ("dropin_per_routine_kernels", {"OCFD_HIP_STAGE_FUSION": "0"})):
        wd = tempfile.mkdtemp(prefix="ocfd_dropin_")
        try:
            t = time.time()
            _, iv, total = dropin.run(case, wd, 2 * steps, show_every=steps, env=env)
            out[key] = round(cells * steps / iv[-1] / 1e6, 3)
            log("drop-in leg: %s %.1f Mcell-RK3-steps/s (%.3f s per %d steps; run %.1f s)" % (
                key, out[key], iv[-1], steps, time.time() - t))
        finally:
            shutil.rmtree(wd, ignore_errors=True)
    out["dropin_over_ocfd_step"] = round(out["dropin"] / out["ocfd_step"], 4)
    return out


def spinup_steps_for(workload: str, case) -> int:
    """Steps that take the isotropic field to the end time of the 256^3 run
    (200 steps; dt scales with the grid spacing, so 400 at 512^3)."""
    if not workload.startswith("iso"):
        return 0
    return int(math.ceil(SPINUP_STEPS_256 * case.cfg.nx / 256))


def run_block(case, precision, spinup, warmup, steps, prof_steps, autotune, marker):
    """Set up one block, develop the field (ana_box every HISTORY_EVERY
    steps), warm up, time `steps` whole steps (replayed from the captured
    graph), then run `prof_steps` more steps with every kernel bracketed by
    HIP events on the library's stream (eager launches) for the per-kernel
    times.  Returns (block, elapsed seconds of the timed steps, per-kernel
    profile, KE/enstrophy history)."""
    name = "mixed" if precision == PREC_MIXED else "fp64"
    b = native.setup_block(native.Block(case.cfg.replace(precision=precision)), case)
    hist = []
    if spinup:
        t1 = time.time()
        done = 0
        while done < spinup:
            k = min(HISTORY_EVERY, spinup - done)
            b.step(k)
            done += k
            a = b.ana_box()
            hist.append((done, float(a[12]), float(a[13])))
        b.synchronize()
        log("spin-up (%s): %d steps in %.1f s, t = %.4f" % (name, spinup, time.time() - t1, b.get_time()[1]))
    if autotune:
        b.tiles = b.autotune()
        log("autotuned sweep tiles (x, y, z): %s" % b.tiles)
    b.step(warmup)
    b.synchronize()
    trace_marker(b, marker)
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
    return b, elapsed, prof, hist


def compare_forms(bm, b64, hm, h64):
    """Relative L2 of the mixed-precision primitives against FP64, and the
    largest relative difference of the kinetic energy and enstrophy over the
    sampled history and at the end (ana_box)."""
    import numpy as np
    pm, p64 = bm.download_primitives(), b64.download_primitives()
    L = 4
    out = {}
    for q, nm in enumerate("duvwT"):
        a = pm[q][L:-L, L:-L, L:-L]
        r = p64[q][L:-L, L:-L, L:-L]
        out[nm] = float("%.3e" % (np.linalg.norm(a - r) / np.linalg.norm(r)))
    del pm, p64
    am, a64 = bm.ana_box(), b64.ana_box()
    ke = [abs(am[12] - a64[12]) / abs(a64[12])]
    en = [abs(am[13] - a64[13]) / abs(a64[13])]
    for (s1, k1, e1), (s2, k2, e2) in zip(hm, h64):
        assert s1 == s2
        ke.append(abs(k1 - k2) / abs(k2))
        en.append(abs(e1 - e2) / abs(e2))
    return out, float("%.3e" % max(ke)), float("%.3e" % max(en))


def hybrid_mix(b):
    b.exchange_primitives()
    c = b.comput_rhybrid(counts=True)
    return [c[0] / c[3], c[1] / c[3], c[2] / c[3]]


# ---------------------------------------------------------------------------
# roofline
# ---------------------------------------------------------------------------
PROFILE_FILE = {("iso512", "fp64"): "stage_update_iso512_fp64.md",
                ("iso512", "mixed"): "stage_update_iso512_mixed.md"}
KERNEL_SYMBOL = {"stage_update": "k_visc_div_update", "stage_update_fused": "k_visc_fused",
                 "visc_flux": "k_visc_flux", "halo_fill": "k_halo",
                 "inviscid_x": "k_inviscid<0", "inviscid_y": "k_inviscid<1", "inviscid_z": "k_inviscid<2",
                 "inviscid_char_x": "k_inviscid<0", "inviscid_char_y": "k_inviscid<1",
                 "inviscid_char_z": "k_inviscid<2", "weno_faces": "k_weno_faces", "sensor": "k_sensor1",
                 "hyb_list": "k_hyb_list"}


def profile_traffic(workload: str, precision: str, name: str):
    """HBM bytes per launch of kernel `name` as rocprofv3 FETCH_SIZE +
    WRITE_SIZE measured them (scaled by the copy kernel) in the committed
    summary of this workload and precision; (bytes, file) or (None, None)."""
    fn = PROFILE_FILE.get((workload, precision))
    sym = KERNEL_SYMBOL.get(name)
    if fn is None or sym is None:
        return None, None
    path = os.path.join(ROOT, "profiles", fn)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        # rows of the HBM table: | `symbol` | launches | avg us | read MB | write MB | ...
        if len(cols) >= 5 and cols[0].startswith("`" + sym):
            try:
                return (float(cols[3]) + float(cols[4])) * 1e6, "profiles/" + fn
            except ValueError:
                continue
    return None, None


def step_traffic(workload: str, precision: str):
    """HBM bytes per time step of every kernel of the step, from the committed
    summary of this workload and precision (rocprofv3 FETCH_SIZE + WRITE_SIZE
    per launch times launches, over the profiled steps, i.e. stage-update
    launches / 3; the diagnostics kernel k_hyb_counts, the runtime's copies
    and the calibration copy kernel left out); (bytes, file) or (None, None)."""
    fn = PROFILE_FILE.get((workload, precision))
    if fn is None or not os.path.exists(os.path.join(ROOT, "profiles", fn)):
        return None, None
    total, updates = 0.0, 0
    for line in open(os.path.join(ROOT, "profiles", fn)):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) < 5 or not cols[0].startswith("`"):
            continue
        name = cols[0].strip("`")
        try:
            n, rd, wr = int(cols[1]), float(cols[3]), float(cols[4])
        except ValueError:
            continue
        if name.startswith("__amd") or name.startswith("k_hyb_counts") or name.startswith("k_copy"):
            continue
        if name.startswith("k_visc_div_update") or name.startswith("k_visc_fused"):
            updates += n
        total += n * (rd + wr) * 1e6
    if updates == 0:
        return None, None
    return total / (updates / 3.0), "profiles/" + fn


def step_hbm(workload, prec, step_s):
    """The whole step's measured HBM traffic (committed profile of the same
    workload) over this run's time per step, against 8 TB/s."""
    b, src = step_traffic(workload, "mixed" if prec == PREC_MIXED else "fp64")
    if b is None:
        return None
    gbs = b / step_s / 1e9
    return {"bytes_per_step": round(b), "achieved_GBs": round(gbs, 1), "peak_GBs": perfmodel.PEAK_HBM_GBS,
            "frac": round(gbs / perfmodel.PEAK_HBM_GBS, 4),
            "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of every kernel of the step in the committed summary %s "
                      "(not collected in this run), over this run's time per step" % src}


def kernel_roofline(name, ms, cnt, cfg, variant, mix, prec, workload):
    """The kernel `name` on its roofline: compulsory bytes (stage update:
    120 B per cell, 3 x 5 FP64 values) or algorithmic FLOPs per launch over
    its average launch time."""
    n = cfg.nx
    cells = cfg.nx * cfg.ny * cfg.nz
    avg_s = ms / cnt / 1e3
    uniform = cfg.metric_mode == METRIC_UNIFORM
    rb = 4 if prec == PREC_MIXED else 8
    peak_f = perfmodel.PEAK_FP64_TFLOPS if prec == PREC_FP64 else perfmodel.PEAK_FP32_TFLOPS
    if name.startswith("inviscid") or name == "weno_faces":
        if name == "weno_faces":
            flops = perfmodel.weno_faces_axis_flops(variant, n, mix)
        else:
            flops = perfmodel.inviscid_axis_flops(variant, n, mix)
            if name == "inviscid_char_x" or name == "inviscid_char_y" or name == "inviscid_char_z":
                # the sweeps compute the linear faces; the listed WENO faces are k_weno_faces'
                flops -= perfmodel.weno_faces_axis_flops(variant, n, mix)
        ach = flops / avg_s / 1e12
        roof = {"bound": "valu", "achieved": round(ach, 3), "peak": peak_f, "unit": "TFLOP/s",
                "frac": round(ach / peak_f, 4), "kernel": name, "algorithmic_per_launch": flops,
                "note": "no MFMA: the stencils have no dense contraction; %s vector rate; exact operation "
                        "counts of the reference algorithm" % ("FP64" if prec == PREC_FP64 else "FP32")}
    else:
        if name in ("stage_update", "stage_update_fused"):
            byts = float(perfmodel.COMPULSORY_STAGE_BYTES * cells)
            basis = "120 B per cell and stage: 5 conserved variables read, 5 stage-start values read, 5 written (FP64)"
        else:
            byts = perfmodel.algorithmic_bytes(name, cells, uniform, viscous=cfg.if_viscous == 1, rhs_bytes=rb)
            basis = "every array of this kernel read or written once"
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": perfmodel.PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / perfmodel.PEAK_HBM_GBS, 4), "kernel": name,
                "algorithmic_per_launch": byts, "basis": basis}
    roof["avg_launch_ms"] = round(ms / cnt, 5)
    design = perfmodel.algorithmic_bytes(name, cells, uniform, viscous=cfg.if_viscous == 1, rhs_bytes=rb)
    if design:
        roof["design_bytes_per_launch"] = design
        roof["design_bytes_frac"] = round(design / avg_s / 1e9 / perfmodel.PEAK_HBM_GBS, 4)
    traffic, src = profile_traffic(workload, "mixed" if prec == PREC_MIXED else "fp64", name)
    roof["traffic"] = traffic
    if src:
        roof["traffic_source"] = ("rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel in the same workload, "
                                  "from the committed summary %s (not collected in this run)" % src)
    return roof


def step_roofline(rate_mcells, cfg, variant, mix, prec):
    """The whole step against both roofs: 360 compulsory bytes per cell-step
    and the exact algorithmic FLOPs per cell-step at the face mix."""
    n = cfg.nx
    f_cell = perfmodel.step_flops_per_cell(variant, n, mix, viscous=cfg.if_viscous == 1)
    rate = rate_mcells * 1e6
    hbm = perfmodel.COMPULSORY_STEP_BYTES * rate / 1e9
    if prec == PREC_FP64:
        tf = f_cell * rate / 1e12
        flop_frac = tf / perfmodel.PEAK_FP64_TFLOPS
        peak_note = "FP64 vector 78.6 TFLOP/s"
    else:
        # every operation priced at the FP32 rate: the shortest time the
        # mixed form's arithmetic could take, so the lowest fraction
        tf = f_cell * rate / 1e12
        flop_frac = tf / perfmodel.PEAK_FP32_TFLOPS
        peak_note = "all operations priced at the FP32 vector rate 157.3 TFLOP/s"
    return {"bytes_per_cell_step": perfmodel.COMPULSORY_STEP_BYTES, "hbm_achieved_GBs": round(hbm, 1),
            "hbm_frac": round(hbm / perfmodel.PEAK_HBM_GBS, 4),
            "flops_per_cell_step": round(f_cell, 1), "flop_achieved_TFs": round(tf, 3),
            "flop_frac": round(flop_frac, 4), "flop_peak": peak_note,
            "frac": round(max(flop_frac, hbm / perfmodel.PEAK_HBM_GBS), 4)}


# ---------------------------------------------------------------------------
# batched ensemble of independent periodic boxes
# ---------------------------------------------------------------------------
def run_batch(args):
    nbox = int(args.workload[5:] or 96)
    prec = PREC_MIXED if args.precision == "mixed" else PREC_FP64
    t0 = time.time()
    cs = cases.batch_ensemble(nbox, prec)
    cells = sum(c.cfg.nx * c.cfg.ny * c.cfg.nz for c in cs)
    log("%d boxes (%.1f Mcells) built in %.1f s" % (nbox, cells / 1e6, time.time() - t0))
    with native.Batch([c.cfg for c in cs]) as batch:
        for i, c in enumerate(cs):
            native.setup_block(batch.box(i), c, use_uniform=True)
        tiles = batch.autotune() if args.autotune else None
        # one launch per kernel phase (the persistent work-queue launch,
        # mode 1, is switched off in the library)
        batch.set_launch_mode(0)
        batch.step(args.warmup)
        batch.synchronize()
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
                   "box_sizes": {str(n): sizes.count(n) for n in cases.BATCH_SIZES}, "cells": cells,
                   "initial": "Taylor-Green or seeded isotropic per box; Ma U[0.1,0.5], Re U[400,3200]",
                   "scheme": "WENO5 or OMP6 per box, Steger-Warming, CD6, RK3 CFL 0.5",
                   "precision": "FP64" if prec == PREC_FP64 else "mixed",
                   "execution": "one launch per kernel phase over all boxes, graph-captured steps; in mixed "
                                "precision the boxes with whole 32x8 tiles take the fused viscous stage update "
                                "(the persistent work-queue launch is built but switched off: its kernel did not "
                                "finish on the GPU)",
                   "autotuned": tiles},
        "one_at_a_time": {"value": round(single, 3), "ms_per_step": round(el_s / args.steps * 1e3, 4),
                          "batched_speedup": round(el_s / el_b, 4)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def decomposition(p: int):
    """npx0 x npy0 x npz0 = p, as even as powers of two allow."""
    d = [1, 1, 1]
    q, a = p, 0
    while q > 1 and q % 2 == 0:
        d[a % 3] *= 2
        q //= 2
        a += 1
    d[0] *= q
    return tuple(d)


def developed_case(workload: str, n: int, spinup_steps_256: int):
    """The workload's case on an n^3 box, developed on the GPU (FP64) to the
    bench's physical time; its primitives replace the initial field."""
    case, _, _ = build_case(workload, n_override=n)
    steps = int(math.ceil(spinup_steps_256 * n / 256)) if workload.startswith("iso") else 0
    if steps:
        with native.setup_block(native.Block(case.cfg), case) as b:
            b.step(steps)
            pr = b.download_primitives()
            tt = b.get_time()[1]
        L = 4
        case.prims = [p[L:-L, L:-L, L:-L].copy() for p in pr]
    else:
        tt = 0.0
    case.cfg.metric_mode = METRIC_FIELD
    case.cfg.grid_type = cases.GRID_BOX
    return case, steps, tt


def cpu_baseline(workload: str, mode: str):
    """Time the CPU path on a bounded sample; returns the cpu_baseline dict."""
    if mode == "none":
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    # one GPU's share of the host is 16 cores (OMP_NUM_THREADS on the GPU hosts)
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    p = max(1, min(ncpu, share, 16))
    model = cpu_model()
    if mode in ("auto", "reference") and pyoracle.reference_available():
        n, steps, n1, steps1 = (256, 2, 128, 1) if workload.startswith("iso") else (64, 4, 64, 1)
        try:
            case, sp, tt = developed_case(workload, n, SPINUP_STEPS_256)
            dec = decomposition(p)
            walls = []
            with tempfile.TemporaryDirectory(prefix="ocfd_cpu_") as d:
                for rep in range(3):
                    wall, _, _ = pyoracle.run_reference(case, os.path.join(d, "r%d" % rep), steps, nranks=p,
                                                        decomposition=dec, timeout=900)
                    walls.append(wall)
                    log("reference on %d ranks %s, %d^3, %d steps: %.2f s" % (p, dec, n, steps, wall))
                case1, _, _ = developed_case(workload, n1, SPINUP_STEPS_256)
                wall1, _, _ = pyoracle.run_reference(case1, os.path.join(d, "one"), steps1, nranks=1,
                                                     timeout=900)
                log("reference on 1 rank, %d^3, %d step: %.2f s" % (n1, steps1, wall1))
            wall = statistics.median(walls)
            return {"value": round(n ** 3 * steps / wall / 1e6, 4), "unit": "Mcell-RK3-steps/s", "cores": p,
                    "kind": "reference", "cpu_model": model, "host_cpus_visible": ncpu,
                    "runs_s": [round(w, 3) for w in walls],
                    "one_rank": {"value": round(n1 ** 3 * steps1 / wall1 / 1e6, 4), "cores": 1,
                                 "sample": "%d^3, %d step" % (n1, steps1)},
                    "sample": "reference Fortran (amdflang -O3, MPICH) on %d ranks %s (this GPU's share of the "
                              "host), median of 3 runs of %d RK3 steps on a %d^3 box of the same case and scheme, "
                              "developed on the GPU to t = %.4f like the GPU run; timed by its own 'total wtime "
                              "cost'" % (p, dec, steps, n, tt)}
        except Exception as e:  # reference could not run here: fall back to the oracle
            log("reference CPU run failed (%s); timing the oracle instead" % str(e)[:300])
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
            "kind": "port", "cpu_model": model,
            "sample": "C restatement (gcc -O2, 1 thread), %d^3 block, %d steps" % (n, steps)}


# ---------------------------------------------------------------------------
def tiles_desc(b):
    """Sweep launch shapes per axis: tiles (points x lines) or y/z marches
    (one lane per line, segments of 32 or 64 cells)."""
    out = []
    for a, s in enumerate(b.tiles):
        if s >= 3:
            out.append("march%d" % (32 if s == 3 else 64))
        else:
            tl = (64, 32, 128)[s] if a == 0 else (32, 16, 64)[s]
            out.append("%dx%d" % (tl, 256 // tl))
    return out


def dominant(prof):
    name, (ms, cnt) = max(prof.items(), key=lambda x: x[1][0])
    return name, ms, cnt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--spinup", type=int, default=-1,
                    help="untimed steps before warm-up that develop the initial field (default: the isotropic "
                         "workloads run to t = %d dt(256^3), others 0)" % SPINUP_STEPS_256)
    ap.add_argument("--workload", default="iso512")
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick the sweep tile shapes with the library's autotuner after the spin-up (untimed)")
    ap.add_argument("--prof-steps", type=int, default=2,
                    help="steps after the timed region with per-kernel HIP events (roofline, kernel table)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="the form whose throughput is `value` (the other runs beside it)")
    ap.add_argument("--other-form", type=int, default=1,
                    help="also run the other precision form on the same workload and report it (default 1)")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "none"])
    ap.add_argument("--dropin", type=int, default=1,
                    help="also time the reference driver with this library as its hot path "
                         "(%d^3, %d steps; default 1)" % (DROPIN_N, DROPIN_STEPS))
    ap.add_argument("--trace-marker", type=int, default=0,
                    help="launch the CFL reduction kernel once right before the timed steps, so that a "
                         "rocprofv3 trace or counter run can keep only the dispatches after it")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after the timed region, launch the plain fn=f copy kernel (5 planes of FP64 "
                         "read and written) this many times: the reference for scaling HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one MI355X: --gpus %d is not supported (one GPU runs the whole block; the "
            "multi-GPU halo-exchange path is design-only)" % args.gpus)
        sys.exit(2)

    if args.workload.startswith("batch"):
        return run_batch(args)
    t0 = time.time()
    case, desc, variant = build_case(args.workload)
    cfg = case.cfg
    cells = cfg.nx * cfg.ny * cfg.nz
    log("case %s built in %.1f s (dt=%.4g)" % (case.name, time.time() - t0, cfg.dt))
    prec = PREC_MIXED if args.precision == "mixed" else PREC_FP64
    spinup = args.spinup if args.spinup >= 0 else spinup_steps_for(args.workload, case)
    if spinup:
        desc["spinup_steps"] = spinup
        desc["spinup_time"] = round(spinup * cfg.dt, 5)
    forms = [prec] + ([PREC_MIXED if prec == PREC_FP64 else PREC_FP64] if args.other_form else [])
    runs = {}
    for pr in forms:
        b, elapsed, prof, hist = run_block(case, pr, spinup, args.warmup, args.steps, args.prof_steps,
                                           args.autotune, bool(args.trace_marker))
        if not prof:
            raise SystemExit("--prof-steps must be >= 1: the roofline needs per-kernel times")
        mix = hybrid_mix(b) if cfg.scheme_invis == "HYBRID" else None
        rate = cells * args.steps / elapsed / 1e6
        nm = "mixed" if pr == PREC_MIXED else "fp64"
        log("%s: timed %d steps: %.3f ms/step, %.1f Mcell-RK3-steps/s, face mix %s" % (
            nm, args.steps, elapsed / args.steps * 1e3, rate, mix))
        for k, (ms, cnt) in sorted(prof.items(), key=lambda x: -x[1][0]):
            log("  %-18s %9.3f ms  %5d launches  %8.4f ms/launch" % (k, ms, cnt, ms / cnt))
        if args.calibrate_copy and pr == prec:
            # k_copy moves exactly 5*cells*8 bytes in and 5*cells*8 bytes out per launch
            for _ in range(args.calibrate_copy):
                b.save_fn()
            b.synchronize()
        runs[pr] = dict(b=b, elapsed=elapsed, prof=prof, hist=hist, mix=mix, rate=rate)

    def form_report(pr):
        r = runs[pr]
        name, ms, cnt = dominant(r["prof"])
        out = {"roofline": kernel_roofline(name, ms, cnt, cfg, variant, r["mix"], pr, args.workload),
               "step_roofline": step_roofline(r["rate"], cfg, variant, r["mix"], pr),
               "step_hbm_measured": step_hbm(args.workload, pr, r["elapsed"] / args.steps),
               "kernels_ms_per_step": {k: round(v[0] / args.prof_steps, 4) for k, v in r["prof"].items()}}
        if r["mix"]:
            out["face_mix_linear_weno7_weno5"] = [round(x, 5) for x in r["mix"]]
        if args.autotune:
            out["sweep_tiles_xyz"] = tiles_desc(r["b"])
        return out

    main_r = runs[prec]
    rep = form_report(prec)
    res = {
        "metric": METRIC, "value": round(main_r["rate"], 3), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(main_r["elapsed"] / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded initial field generated on the host)", "config": desc,
        "roofline": rep.pop("roofline"), "step_roofline": rep.pop("step_roofline"),
        "step_hbm_measured": rep.pop("step_hbm_measured"),
    }
    res["config"]["precision"] = ("FP64" if prec == PREC_FP64 else
                                  "mixed: FP32 splitting, reconstruction, viscous derivatives, du and stored "
                                  "primitives; FP64 conserved variables, RK3, metrics and the characteristic "
                                  "basis and projections of the sensor-flagged faces")
    for k in ("face_mix_linear_weno7_weno5", "sweep_tiles_xyz"):
        if k in rep:
            res["config"][k] = rep.pop(k)
    res["kernels_ms_per_step"] = rep.pop("kernels_ms_per_step")
    if args.calibrate_copy:
        res["copy_kernel_bytes_each_way"] = 5 * cells * 8
    if args.other_form:
        other = forms[1]
        o = runs[other]
        bm, b64 = (runs[PREC_MIXED]["b"], runs[PREC_FP64]["b"])
        el_m, el_64 = runs[PREC_MIXED]["elapsed"], runs[PREC_FP64]["elapsed"]
        errs, ke, en = compare_forms(bm, b64, runs[PREC_MIXED]["hist"], runs[PREC_FP64]["hist"])
        key = "fp64" if prec == PREC_MIXED else "mixed"
        res[key] = {"value": round(o["rate"], 3), "ms_per_step": round(o["elapsed"] / args.steps * 1e3, 4),
                    "mixed_speedup_over_fp64": round(el_64 / el_m, 4),
                    "steps_compared": spinup + args.warmup + args.steps + args.prof_steps,
                    "rel_l2_mixed_vs_fp64": errs,
                    "kinetic_energy_rel_max": ke, "enstrophy_rel_max": en,
                    "history_every_steps": HISTORY_EVERY}
        res[key].update(form_report(other))
        log("%s: %.1f Mcell-RK3-steps/s, mixed/fp64 speedup %.3f, relative L2 %s, KE %.2e, enstrophy %.2e" % (
            key, res[key]["value"], res[key]["mixed_speedup_over_fp64"], errs, ke, en))
    for r in runs.values():
        r["b"].close()
    runs.clear()
    if args.dropin:
        res["dropin"] = dropin_leg()
    res["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_baseline)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
