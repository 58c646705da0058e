"""The benchmark workloads against the compiled reference.

The reference is the unmodified Fortran solver built from its sources
(oracle/_ref/opencfd_ref; opencfd_ref_wall adds the isothermal-wall
ocfd_bc_User of oracle/ref_user_hooks.f90), run with mpiexec on this host's
cores inside the test on the same input files.  Each case runs once.

  C1  Taylor-Green vortex 64^3, WENO5 + Steger-Warming, CD6, 100 steps
  C2  isotropic turbulence 64^3, characteristic hybrid UD7L/WENO7/WENO5 +
      local Lax-Friedrichs, sensor thresholds that list WENO faces, 100
      steps; the periodic uniform-metric kernels of the benchmark, with the
      reference reading the same uniform metrics from its grid file
  C3  Mach-6 flow over a cylinder, 512 x 256 x 8 O-grid, WENO5
      characteristic + Steger-Warming, isothermal ghost-cell wall through
      BC_User_Def, Fo9p filter every step, 100 steps
  C4  the isotropic generator at 256^3, 10 steps (field metrics: bitwise;
      the uniform-metric kernels within 1e-12), and the mixed-precision form
      against FP64 at 512^3 over 200 steps with WENO faces listed
  C5  the 96-box batch in mixed precision against the same boxes stepped
      alone (bitwise), and sampled boxes against the FP64 oracle

Bars: FP64 relative L2 per field <= 1e-12 and kinetic-energy /
enstrophy histories (ana_box every 10 steps, statistics.dat) <= 1e-10;
mixed precision <= 1e-5 against FP64.
"""
import os

import numpy as np
import pytest

import pyoracle
from ocfd import cases, native
from ocfd.config import METRIC_FIELD, METRIC_UNIFORM, PREC_FP64, PREC_MIXED
from util import interior, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = "duvwT"


def _ranks():
    """MPI ranks for the reference: the host cores this process may use, at
    most 16 (one GPU's share of the host)."""
    n = len(os.sched_getaffinity(0))
    p = 16 if n >= 16 else (8 if n >= 8 else (4 if n >= 4 else 1))
    return p, {16: (4, 2, 2), 8: (2, 2, 2), 4: (2, 2, 1), 1: (1, 1, 1)}[p]


def _need_reference(exe=None):
    if not pyoracle.reference_available() or (exe and not os.path.exists(exe)):
        pytest.skip("reference executable not built (make -C oracle ref ref_wall)")


def _gpu_history(b, nsteps, every):
    """Step b nsteps steps; ana_box after every `every` steps -> rows."""
    rows = []
    for _ in range(nsteps // every):
        b.step(every)
        rows.append(b.ana_box())
    return np.array(rows)


def _check_fields(got_halo, want, bar, what):
    errs = {}
    for q, nm in enumerate(FIELDS):
        errs[nm] = rel_l2(interior(got_halo[q]), want[q])
    print("%s: relative L2 per field %s" % (what, {k: "%.2e" % v for k, v in errs.items()}), flush=True)
    for nm, e in errs.items():
        assert e <= bar, "%s: field %s relative L2 %.3e > %.0e" % (what, nm, e, bar)


def _check_history(got, want, what):
    """KE (column 12) and enstrophy (13) of ana_box against statistics.dat;
    the file holds 14 significant digits."""
    assert got.shape == want.shape, (got.shape, want.shape)
    assert np.allclose(got[:, 0], want[:, 0], rtol=1e-12, atol=0), "%s: tt columns differ" % what
    for col, nm in ((12, "kinetic energy"), (13, "enstrophy")):
        rel = np.abs(got[:, col] - want[:, col]) / np.abs(want[:, col])
        print("%s: %s history max relative difference %.2e" % (what, nm, rel.max()), flush=True)
        assert rel.max() <= 1e-10, "%s: %s history differs by %.3e" % (what, nm, rel.max())


def _reference(case, tmp, nsteps, ana_every=0, exe=None, dec=None):
    p, d = _ranks()
    if dec is not None:
        p, d = dec
    wall, prims, out = pyoracle.run_reference(case, str(tmp), nsteps, nranks=p, decomposition=d,
                                              ana_box_every=ana_every, exe=exe, timeout=900)
    assert prims is not None, out[-2000:]
    print("reference: %d ranks %s, %.1f s" % (p, d, wall), flush=True)
    stats = pyoracle.read_statistics(os.path.join(str(tmp), "statistics.dat")) if ana_every else None
    return prims, stats


def test_c1_tgv64_against_reference(tmp_path):
    _need_reference()
    case = cases.tgv(64)
    print("C1: TGV 64^3 WENO5-SW, 100 steps, dt=%.6g" % case.cfg.dt, flush=True)
    want, stats = _reference(case, tmp_path / "ref", 100, ana_every=10)
    with native.setup_block(native.Block(case.cfg), case) as b:
        hist = _gpu_history(b, 100, 10)
        _check_fields(b.download_primitives(), want, 1e-12, "C1")
    _check_history(hist, stats, "C1")


def test_c2_isotropic64_hybrid_against_reference(tmp_path):
    _need_reference()
    case = cases.isotropic(64, scheme="HYBRID", split=2, char=1, hybrid_thresholds=(0.01, 0.05),
                           metric_mode=METRIC_UNIFORM)
    print("C2: isotropic 64^3 hybrid char LLF, 100 steps, dt=%.6g" % case.cfg.dt, flush=True)
    want, stats = _reference(case, tmp_path / "ref", 100, ana_every=10)
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.autotune(reps=1)  # the benchmark's kernels: autotuned periodic sweeps
        hist = _gpu_history(b, 100, 10)
        _check_fields(b.download_primitives(), want, 1e-12, "C2")
        b.exchange_primitives()
        c = b.comput_rhybrid(counts=True)
        print("C2: WENO7 / WENO5 faces listed at step 100: %d / %d of %d" % (c[1], c[2], c[3]), flush=True)
        assert c[1] + c[2] > 0
    _check_history(hist, stats, "C2")


def test_c3_blunt_cylinder_against_reference(tmp_path):
    _need_reference(pyoracle.REF_WALL_EXE)
    case = cases.blunt_ogrid()
    print("C3: blunt cylinder %s, WENO5 char SW + Fo9p + wall, 100 steps, dt=%.6g, %s" % (
        case.cfg.shape, case.cfg.dt, case.info), flush=True)
    # the reference's result on this O-grid (i periodic with the coordinates
    # wrapping, zero span) changes when its i direction is split over ranks;
    # split in j only, it equals the one-rank run bit for bit
    p, _ = _ranks()
    dec = (p, (1, p, 1))
    want, _ = _reference(case, tmp_path / "ref", 100, exe=pyoracle.REF_WALL_EXE, dec=dec)
    with native.setup_block(native.Block(case.cfg), case) as b:
        n = b.step(100)
        print("C3: %d T <= 0 repairs in 100 steps" % n, flush=True)
        _check_fields(b.download_primitives(), want, 1e-12, "C3")


def test_c4_isotropic256_window_against_reference(tmp_path):
    _need_reference()
    case = cases.isotropic(256, scheme="HYBRID", split=2, char=1, hybrid_thresholds=(0.01, 0.05),
                           metric_mode=METRIC_FIELD)
    case.cfg.grid_type = cases.GRID_BOX
    print("C4: isotropic 256^3 hybrid, 10 steps against the reference", flush=True)
    want, _ = _reference(case, tmp_path / "ref", 10)
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.step(10)
        _check_fields(b.download_primitives(), want, 0.0, "C4 field metrics")
    case.uniform_metrics = cases.uniform_box_metrics(256, 256, 256)
    case.cfg.metric_mode = METRIC_UNIFORM
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.autotune(reps=1)
        b.step(10)
        _check_fields(b.download_primitives(), want, 1e-12, "C4 uniform-metric kernels")


def test_c4_mixed512_against_fp64():
    case = cases.isotropic(512, scheme="HYBRID", split=2, char=1, hybrid_thresholds=(0.01, 0.05),
                           metric_mode=METRIC_UNIFORM)
    print("C4: 512^3 mixed against FP64, 200 steps, history every 20", flush=True)
    hists = {}
    with native.setup_block(native.Block(case.cfg.replace(precision=PREC_FP64)), case) as b64, \
            native.setup_block(native.Block(case.cfg.replace(precision=PREC_MIXED)), case) as bm:
        for nm, b in (("fp64", b64), ("mixed", bm)):
            b.autotune(reps=1)
            hists[nm] = _gpu_history(b, 200, 20)
        b64.exchange_primitives()
        c = b64.comput_rhybrid(counts=True)
        print("C4: WENO7/WENO5 faces listed at step 200: %d / %d of %d" % (c[1], c[2], c[3]), flush=True)
        assert c[1] + c[2] > 0
        p64 = [interior(a) for a in b64.download_primitives()]
        _check_fields(bm.download_primitives(), p64, 1e-5, "C4 mixed vs FP64")
    for col, nm in ((12, "kinetic energy"), (13, "enstrophy")):
        rel = np.abs(hists["mixed"][:, col] - hists["fp64"][:, col]) / np.abs(hists["fp64"][:, col])
        print("C4: %s history max relative difference %.2e" % (nm, rel.max()), flush=True)
        assert rel.max() <= 1e-5


def test_c5_batch96_mixed():
    cs = cases.batch_ensemble(96, PREC_MIXED)
    cells = sum(c.cfg.nx * c.cfg.ny * c.cfg.nz for c in cs)
    print("C5: 96 boxes, %.1f M cells, mixed precision, 3 steps" % (cells / 1e6), flush=True)
    with native.Batch([c.cfg for c in cs]) as batch:
        for i, c in enumerate(cs):
            native.setup_block(batch.box(i), c, use_uniform=True)
        batch.autotune(reps=1)
        assert batch.step(3) == 0
        for i, c in enumerate(cs):
            with native.setup_block(native.Block(c.cfg), c, use_uniform=True) as alone:
                alone.step(3)
                assert np.array_equal(batch.box(i).download_f(), alone.download_f()), "box %d (%s)" % (i, c.name)
        # sampled boxes against the FP64 oracle: the smallest of each scheme
        picked = {}
        for i, c in enumerate(cs):
            key = c.cfg.scheme_invis
            if key not in picked or c.cfg.nx < cs[picked[key]].cfg.nx:
                picked[key] = i
        for key, i in sorted(picked.items()):
            c = cs[i]
            c.cfg = c.cfg.replace(precision=PREC_FP64)  # the oracle's arithmetic
            ob = pyoracle.setup_oracle(c)
            ob.step(3)
            want = ob.download_f()
            err = rel_l2(batch.box(i).download_f(), want)
            print("C5: box %d (%s) mixed against the FP64 oracle: relative L2 %.2e" % (i, c.name, err), flush=True)
            assert err <= 1e-5
