"""The T <= 0 repair of the HIP path against the compiled reference.

tests/golden/ref_negt_w5_sw.npz is the reference's run of
cases.negative_t_case(): 3 to 6 cells reach T <= 0 in every RK stage of
three steps, through the fluxes of a real stage (their neighbours' T changes
within the stage too).  handle_NegativeT averages the -i/-j/-k neighbours'
new and the +i/+j/+k neighbours' previous T (OCFD_Time_Adv.f90:45-58,
OCFD_handle_NegativeT.f90:11); tests/test_negative_t.py shows on the CPU that
this fixture tells that apart from averaging the new values of all six.

Every comparison with the fixture is bitwise (the FP64 path keeps the
reference's operation order).  The run stops at the first stage with more
than 10 such cells (OCFD_handle_NegativeT.f90:13-16): ocfd_step and
ocfd_batch_step report it and leave the primitives of that stage, which are
compared with the oracle's at the same point (oracle_step returns there).
"""
import numpy as np
import pytest

import pyoracle
from ocfd import cases, native
from util import interior, load_golden

pytestmark = pytest.mark.gpu


def assert_same(got, want, what):
    nbad = int(np.count_nonzero(np.asarray(got) != np.asarray(want)))
    assert nbad == 0, "%s: %d values differ" % (what, nbad)


def _stage_counts(b, nsteps):
    """nsteps steps through the per-routine entry points; the T <= 0 count
    of every stage."""
    out = []
    for _ in range(nsteps):
        b.save_fn()
        for krk in (1, 2, 3):
            out.append(b.stage(krk))
        b.advance_clock()
        b.filtering()
    return out


@pytest.mark.parametrize("fusion", [False, True], ids=["routines", "stage_fusion"])
def test_per_routine_stages_match_reference(fusion):
    case, ref = load_golden("negt_w5_sw")
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.set_stage_fusion(fusion)
        b.save_fn()
        b.exchange_primitives()
        b.du_inviscous()
        b.du_viscous()
        b.time_adv_rk3(1)
        if not fusion:  # (stage fusion refreshes the boundary planes of f here already)
            assert_same(b.download_f(), ref["k1_f"], "f after stage 1")
        n = [b.comput_duvwt()]
        b.bc()
        assert n[0] >= 2
        assert_same(b.download_primitives()[4], ref["k1_T"], "halo'd T after stage 1")
        for krk in (2, 3):
            n.append(b.stage(krk))
        b.advance_clock()
        b.filtering()
        assert_same(b.download_f(), ref["n1_f"], "f after one step")
        assert_same(b.download_primitives()[4], ref["n1_T"], "halo'd T after one step")
        n += _stage_counts(b, 2)
        assert_same(b.download_f(), ref["nN_f"], "f after three steps")
        assert_same(b.download_primitives()[4], ref["nN_T"], "halo'd T after three steps")
        assert sum(n) == len(ref["negt_cells"])


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_whole_steps_match_reference(graph, monkeypatch):
    case, ref = load_golden("negt_w5_sw")
    if not graph:
        monkeypatch.setenv("OCFD_NO_GRAPH", "1")
    with native.setup_block(native.Block(case.cfg), case) as b:
        n1 = b.step(1)
        assert_same(b.download_f(), ref["n1_f"], "f after ocfd_step(1)")
        assert_same(b.download_primitives()[4], ref["n1_T"], "halo'd T after ocfd_step(1)")
        n2 = b.step(2)
        assert_same(b.download_f(), ref["nN_f"], "f after ocfd_step(3)")
        for q, p in enumerate(b.download_primitives()):
            if q == 4:
                assert_same(p, ref["nN_T"], "halo'd T after ocfd_step(3)")
        assert n1 + n2 == len(ref["negt_cells"]) and n1 >= 6
        assert b.get_time()[0] == 3


@pytest.mark.parametrize("mode", [0, pytest.param(1, marks=pytest.mark.skip(
    reason="the persistent queue launch is switched off: its kernel did not finish on the GPU"))],
    ids=["per_phase", "queue"])
def test_batch_box_matches_reference(mode):
    """The fixture's case as a box of a batch beside an unrelated box.  A
    batched box takes uniform metrics, which differ from the box generator's
    field metrics in the last bit on this block, so the expected values are
    the oracle's on the same uniform metrics (the oracle equals the fixture
    bit for bit on field metrics; the uniform run repairs 14 + 30 cells, as
    many as the fixture)."""
    case, _ = load_golden("negt_w5_sw")
    case.uniform_metrics = cases.uniform_box_metrics(case.cfg.nx, case.cfg.ny, case.cfg.nz)
    case.cfg.metric_mode = 1
    ob = pyoracle.setup_oracle(case)
    m1 = ob.step(1)
    want1 = (ob.download_f(), ob.download_primitives()[4])
    m2 = ob.step(2)
    want3 = (ob.download_f(), ob.download_primitives()[4])
    assert m1 >= 6 and m2 >= 6
    # (the queue launch takes non-characteristic boxes only)
    other = cases.random_case(shape=(16, 20, 24), seed=71, scheme="WENO7", split=2, char=0 if mode else 1)
    other.uniform_metrics = cases.uniform_box_metrics(other.cfg.nx, other.cfg.ny, other.cfg.nz)
    other.cfg.metric_mode = 1
    with native.Batch([case.cfg, other.cfg]) as bt:
        native.setup_block(bt.box(0), case, use_uniform=True)
        native.setup_block(bt.box(1), other, use_uniform=True)
        bt.set_launch_mode(mode)
        n1 = bt.step(1)
        assert_same(bt.box(0).download_f(), want1[0], "batch box f after one step")
        assert_same(bt.box(0).download_primitives()[4], want1[1], "batch box T after one step")
        n2 = bt.step(2)
        assert_same(bt.box(0).download_f(), want3[0], "batch box f after three steps")
        assert_same(bt.box(0).download_primitives()[4], want3[1], "batch box T after three steps")
        assert (n1, n2) == (m1, m2)


def _halting_case():
    # more than 10 cells in stage 1 of step 2 (6, 6, 8 in step 1)
    return cases.negative_t_case(amp=6.0, t_factor=1.0, dt_factor=5.0)


def _oracle_at_halt(case):
    ob = pyoracle.setup_oracle(case)
    ob.step(3)  # returns after the update and repair of the failing stage
    return ob.download_primitives()


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_run_stops_at_first_stage_over_limit(graph, monkeypatch):
    case = _halting_case()
    want = _oracle_at_halt(case)
    if not graph:
        monkeypatch.setenv("OCFD_NO_GRAPH", "1")
    with native.setup_block(native.Block(case.cfg), case) as b:
        with pytest.raises(native.OcfdError, match="stage 1 of step 2"):
            b.step(3)
        assert b.get_time()[0] == 1
        got = b.download_primitives()
        for q, nm in enumerate("duvwT"):
            assert_same(interior(got[q]), interior(want[q]), "%s at the failing stage" % nm)


def test_batch_box_stops_at_first_stage_over_limit():
    case = _halting_case()
    other = cases.random_case(shape=(16, 20, 24), seed=72, scheme="WENO5", split=1, char=0)
    for c in (case, other):
        c.uniform_metrics = cases.uniform_box_metrics(c.cfg.nx, c.cfg.ny, c.cfg.nz)
        c.cfg.metric_mode = 1
    want = _oracle_at_halt(case)  # on the batched box's uniform metrics
    with native.setup_block(native.Block(other.cfg), other, use_uniform=True) as alone:
        alone.step(3)
        f_alone = alone.download_f()
    with native.Batch([case.cfg, other.cfg]) as bt:
        native.setup_block(bt.box(0), case, use_uniform=True)
        native.setup_block(bt.box(1), other, use_uniform=True)
        with pytest.raises(native.OcfdError, match="box 0: .*stage 1 of step 2"):
            bt.step(3)
        assert bt.box(0).get_time()[0] == 1 and bt.box(1).get_time()[0] == 3
        got = bt.box(0).download_primitives()
        for q, nm in enumerate("duvwT"):
            assert_same(interior(got[q]), interior(want[q]), "batch %s at the failing stage" % nm)
        assert_same(bt.box(1).download_f(), f_alone, "the other box after three steps")
