"""The viscous terms fused into the stage update (k_visc_fused: mixed
precision, periodic uniform-metric blocks with CD6 and whole 32 x 8 tiles)
against the two-kernel path (k_visc_flux + k_visc_div_update): the fused
march computes every value with the same expressions in the same order, so
whole steps are bitwise equal -- the conserved variables and the halo'd
primitives, with graph replay and eager launches."""
import numpy as np
import pytest

from ocfd import cases, native
from ocfd.config import PREC_MIXED

pytestmark = pytest.mark.gpu

CASES = [
    dict(seed=91, scheme="HYBRID", split=2, char=1, jump=0.5, hybrid=(0.002, 0.02)),
    dict(seed=92, scheme="WENO5", split=1, char=0),
    dict(seed=93, scheme="OMP6", split=1, char=0, noise=0.03),
]


def _case(kw, shape):
    case = cases.random_case(shape=shape, **kw)
    case.uniform_metrics = cases.uniform_box_metrics(case.cfg.nx, case.cfg.ny, case.cfg.nz)
    case.cfg = case.cfg.replace(metric_mode=1, precision=PREC_MIXED)
    return case


def _run(case, nsteps, monkeypatch, fused, graph=True):
    if fused:
        monkeypatch.delenv("OCFD_NO_FUSED", raising=False)
    else:
        monkeypatch.setenv("OCFD_NO_FUSED", "1")
    if graph:
        monkeypatch.delenv("OCFD_NO_GRAPH", raising=False)
    else:
        monkeypatch.setenv("OCFD_NO_GRAPH", "1")
    with native.setup_block(native.Block(case.cfg), case, use_uniform=True) as b:
        b.step(nsteps)
        return b.download_f(), b.download_primitives()


def _same(a, b, what):
    nbad = int(np.count_nonzero(np.asarray(a) != np.asarray(b)))
    assert nbad == 0, "%s: %d values differ" % (what, nbad)


def _kernels(case, monkeypatch, fused):
    """Names of the kernels one profiled (eager) step launches."""
    if fused:
        monkeypatch.delenv("OCFD_NO_FUSED", raising=False)
    else:
        monkeypatch.setenv("OCFD_NO_FUSED", "1")
    with native.setup_block(native.Block(case.cfg), case, use_uniform=True) as b:
        b.profile_enable(True)
        b.step(1)
        return sorted(b.profile_read())


# (nz, ny, nx): nx = 64 and 32, multiples of the fused tile's 32 x 8 columns
SHAPES = [(24, 32, 64), (40, 16, 32)]


@pytest.mark.parametrize("kw", CASES, ids=[c["scheme"] + str(c["seed"]) for c in CASES])
@pytest.mark.parametrize("shape", SHAPES, ids=["64x32x24", "32x16x40"])
def test_fused_steps_equal_two_kernel_steps(kw, shape, monkeypatch):
    case = _case(kw, shape)
    print("fused viscous stage update: %s nx, ny, nz = %s" % (kw["scheme"], shape[::-1]), flush=True)
    names = _kernels(case, monkeypatch, fused=True)
    assert "stage_update_fused" in names and "visc_flux" not in names, ", ".join(names)
    names = _kernels(case, monkeypatch, fused=False)
    assert "stage_update" in names and "visc_flux" in names, ", ".join(names)
    f0, p0 = _run(case, 3, monkeypatch, fused=False)
    f1, p1 = _run(case, 3, monkeypatch, fused=True)
    _same(f1, f0, "f after 3 steps")
    for q, nm in enumerate("duvwT"):
        _same(p1[q], p0[q], "halo'd %s after 3 steps" % nm)
    f2, p2 = _run(case, 3, monkeypatch, fused=True, graph=False)
    _same(f2, f0, "f after 3 eager steps")
    for q, nm in enumerate("duvwT"):
        _same(p2[q], p0[q], "halo'd %s after 3 eager steps" % nm)


@pytest.mark.parametrize("kw", CASES[:2], ids=[c["scheme"] + str(c["seed"]) for c in CASES[:2]])
def test_stage_fusion_routines_equal_steps(kw, monkeypatch):
    """The per-routine calls with stage fusion on reach the fused kernel too
    (du_viscous defers, time_adv_rk3 runs k_visc_fused, comput_duvwT repairs
    on the new u, v, w): bitwise equal to ocfd_step."""
    monkeypatch.delenv("OCFD_NO_FUSED", raising=False)
    case = _case(kw, SHAPES[0])
    f0, p0 = _run(case, 3, monkeypatch, fused=True)
    with native.setup_block(native.Block(case.cfg), case, use_uniform=True) as b:
        b.set_stage_fusion(True)
        b.profile_enable(True)
        b.step_by_routines(3)
        names = sorted(b.profile_read())
        f1, p1 = b.download_f(), b.download_primitives()
    print("kernels of the per-routine steps: %s" % ", ".join(names), flush=True)
    assert "stage_update_fused" in names and "visc_flux" not in names, ", ".join(names)
    _same(f1, f0, "f after 3 steps")
    for q, nm in enumerate("duvwT"):
        _same(p1[q], p0[q], "halo'd %s after 3 steps" % nm)
