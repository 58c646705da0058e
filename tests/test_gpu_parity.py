"""HIP path (through the C ABI) against the golden vectors of the compiled
reference and against the C oracle.

Tolerances (FP64 path, per field):
  per-routine outputs and whole steps on the golden cases: bitwise, i.e.
    relative L2 == 0 (the kernels keep the reference's operation order and
    are built without FMA contraction); the one exception is Rhybrid of the
    pressure-gradient sensor (Hybrid_para(1) = 2), which divides by the mean
    of |grad p| over the block: the device sums in parallel, not in the
    reference's i-j-k loop order, so Rhybrid may differ in the last bit
    (relative L2 <= 1e-14); Rhybrid is only compared with the thresholds, so
    everything downstream stays bitwise unless a value sits within an ulp of
    a threshold;
  the 100-step TGV run against the oracle: relative L2 <= 1e-12 per field,
    kinetic-energy / enstrophy histories <= 1e-10 relative; isotropic
    periodic boxes with uniform metrics against the oracle over a few
    steps.  The benchmark configurations (isotropic turbulence among them)
    run against the compiled reference in test_gpu_configs.py.
"""
import numpy as np
import pytest

import pyoracle
from ocfd import cases, native
from util import golden_names, interior, load_golden, rel_l2

pytestmark = pytest.mark.gpu

NAMES = golden_names()


def gpu_block(case, uniform=None):
    return native.setup_block(native.Block(case.cfg), case, use_uniform=uniform)


def assert_same(got, want, what):
    err = rel_l2(got, want)
    nbad = int(np.count_nonzero(got != want))
    assert nbad == 0, "%s: %d values differ, relative L2 %.3e" % (what, nbad, err)


@pytest.mark.parametrize("name", NAMES)
def test_golden_per_routine_bitwise(name):
    case, ref = load_golden(name)
    cfg = case.cfg
    with gpu_block(case) as b:
        if "met" in ref:
            mets = b.download_metrics()
            for q, m in enumerate((9, 10, 14, 18)):
                assert_same(mets[m], ref["met"][q], "metric %d" % m)
        b.save_fn()
        b.exchange_primitives()
        if cfg.scheme_invis.upper() == "HYBRID":
            b.comput_rhybrid()
            if cfg.hybrid_para.get(1, 1.0) == 2.0:
                assert rel_l2(interior(b.download_rhybrid()), ref["rhyb"]) <= 1e-14
            else:
                assert_same(interior(b.download_rhybrid()), ref["rhyb"], "Rhybrid")
        if cfg.if_scheme_character:
            b.du_inviscous_character()
        else:
            b.du_inviscous()
        assert_same(b.download_du(), ref["k1_duinv"], "du inviscid")
        b.du_viscous()
        assert_same(b.download_du(), ref["k1_dutot"], "du total")
        b.time_adv_rk3(1)
        assert_same(b.download_f(), ref["k1_f"], "f after stage 1")
        b.comput_duvwt()
        b.bc()
        for krk in (2, 3):
            b.stage(krk)
        b.advance_clock()
        b.filtering()
        assert_same(b.download_f(), ref["n1_f"], "f after one step")


@pytest.mark.parametrize("name", NAMES)
def test_golden_stage_fusion_bitwise(name):
    """The per-routine calls in the driver's order with stage fusion on
    (ocfd_set_stage_fusion: du_viscous forms the fluxes, time_adv_rk3 runs
    the whole-step stage update, comput_duvwT repairs and counts) against
    the reference, over three steps.  Within a stage the arrays are not
    comparable: the fused update refreshes f on the boundary planes at
    time_adv_rk3, the reference in OCFD_bc."""
    case, ref = load_golden(name)
    with gpu_block(case) as b:
        b.set_stage_fusion(True)
        b.save_fn()
        b.exchange_primitives()
        if case.cfg.scheme_invis.upper() == "HYBRID":
            b.comput_rhybrid()
        if case.cfg.if_scheme_character:
            b.du_inviscous_character()
        else:
            b.du_inviscous()
        if case.cfg.if_viscous:
            b.du_viscous()
        b.time_adv_rk3(1)
        b.comput_duvwt()
        b.bc()
        for krk in (2, 3):
            b.stage(krk)
        b.advance_clock()
        b.filtering()
        assert_same(b.download_f(), ref["n1_f"], "f after one step")
        b.step_by_routines(2)
        assert_same(b.download_f(), ref["nN_f"], "f after three steps")


@pytest.mark.parametrize("name", NAMES)
def test_golden_whole_steps_bitwise(name):
    """ocfd_step (fused stage kernels) over three steps against the reference."""
    case, ref = load_golden(name)
    with gpu_block(case) as b:
        b.step(1)
        assert_same(b.download_f(), ref["n1_f"], "f after ocfd_step(1)")
        b.step(2)
        assert_same(b.download_f(), ref["nN_f"], "f after ocfd_step(3)")
        istep, tt = b.get_time()
        assert istep == 3


@pytest.mark.parametrize("name", ["w5_sw", "hyb_c_llf", "curv_wall_fo9p", "blunt_massforce"])
def test_golden_diagnostics(name):
    case, ref = load_golden(name)
    with gpu_block(case) as b:
        for row in ref["hist"]:
            b.step(1)
            e = b.show_flow_msg()
            a = b.ana_box()
            np.testing.assert_allclose(e, row[2:5], rtol=1e-13, atol=0)
            np.testing.assert_allclose([a[12], a[13]], row[5:7], rtol=1e-12, atol=0)


MID = [
    dict(seed=31, scheme="WENO5", split=1, char=0),
    dict(seed=32, scheme="WENO7", split=2, char=1),
    dict(seed=33, scheme="HYBRID", split=2, char=1, jump=0.5, hybrid=(0.002, 0.02)),
    dict(seed=34, scheme="OMP6", split=1, char=1, noise=0.03),
    dict(seed=35, scheme="WENO5", split=1, char=1, grid="curv", iperiodic=(1, 0, 1),
         scheme_boundary=(0, 0, 1, 0, 0, 0), ghost_cell=(0, 0, 2, 0, 0, 0), bc="BC_USER_DEF",
         bc_para={1: 1.05}, filters=[[1, 1, 1, 1, 1, 40, 1, 40, 1, 40, 1, 0.2, 0.0]]),
]


@pytest.mark.parametrize("kw", MID, ids=[m["scheme"] + str(m["seed"]) for m in MID])
def test_oracle_mid_size_routines_and_steps(kw):
    """Non-cubic 40x28x20 blocks (several tiles per axis, partial tiles)."""
    case = cases.random_case(shape=(20, 28, 40), **kw)
    ob = pyoracle.setup_oracle(case)
    with gpu_block(case) as b:
        for blk in (ob, b):
            blk.save_fn()
            blk.exchange_primitives()
            if case.cfg.scheme_invis == "HYBRID":
                blk.comput_rhybrid()
            (blk.du_inviscous_character if case.cfg.if_scheme_character else blk.du_inviscous)()
        assert_same(b.download_du(), ob.download_du(), "du inviscid")
        ob.step(3)
        b.step(3)
        assert_same(b.download_f(), ob.download_f(), "f after 3 steps")
        pg, po = b.download_primitives(), ob.download_primitives()
        for q in range(5):
            assert_same(interior(pg[q]), interior(po[q]), "primitive %d" % q)


def test_fused_step_equals_routine_sequence():
    case = cases.random_case(shape=(16, 20, 24), seed=41, scheme="WENO7", split=2, char=1)
    with gpu_block(case) as a, gpu_block(case) as b:
        a.step(2)
        b.step_by_routines(2)
        assert_same(a.download_f(), b.download_f(), "fused vs per-routine f")


def test_uniform_metrics_mode_matches_oracle():
    case = cases.isotropic(32, scheme="WENO5", split=1, char=0, metric_mode=1)
    ob = pyoracle.setup_oracle(case, uniform=True)
    with gpu_block(case, uniform=True) as b:
        ob.step(3)
        b.step(3)
        assert_same(b.download_f(), ob.download_f(), "uniform-metric f after 3 steps")


# periodic box + uniform metrics selects the specialised kernels (no block-edge
# closures, metrics as constants); non-cubic 40x28x20 blocks leave partial
# tiles on every axis
UNIFORM = [
    dict(seed=51, scheme="HYBRID", split=2, char=1, jump=0.5, hybrid=(0.002, 0.02)),
    dict(seed=52, scheme="WENO7", split=2, char=1),
    dict(seed=53, scheme="OMP6", split=1, char=1, noise=0.03),
    dict(seed=54, scheme="UD7L", split=2, char=0, vis="CD8"),
    dict(seed=55, scheme="HYBRID", split=1, char=0, jump=0.5, hybrid=(0.002, 0.02)),
    dict(seed=56, scheme="WENO5", split=1, char=0, filters=[[1, 1, 1, 1, 1, 40, 1, 28, 1, 20, 1, 0.2, 0.0]]),
]


@pytest.mark.parametrize("kw", UNIFORM, ids=[m["scheme"] + str(m["seed"]) for m in UNIFORM])
def test_uniform_metrics_periodic_box_against_oracle(kw):
    case = cases.random_case(shape=(20, 28, 40), **kw)
    case.uniform_metrics = cases.uniform_box_metrics(case.cfg.nx, case.cfg.ny, case.cfg.nz)
    case.cfg.metric_mode = 1
    ob = pyoracle.setup_oracle(case, uniform=True)
    with gpu_block(case, uniform=True) as b:
        for blk in (ob, b):
            blk.save_fn()
            blk.exchange_primitives()
            if case.cfg.scheme_invis == "HYBRID":
                blk.comput_rhybrid()
            (blk.du_inviscous_character if case.cfg.if_scheme_character else blk.du_inviscous)()
            blk.du_viscous()
        assert_same(b.download_du(), ob.download_du(), "du (uniform metrics)")
        ob.step(3)
        b.step(3)
        assert_same(b.download_f(), ob.download_f(), "f after 3 steps (uniform metrics)")
        pg, po = b.download_primitives(), ob.download_primitives()
        for q in range(5):
            assert_same(interior(pg[q]), interior(po[q]), "primitive %d" % q)


def test_graph_replay_equals_eager_steps(monkeypatch):
    """ocfd_step replays a captured hipGraph of one time step; the replays,
    with eager work and synchronisations between them, equal eager launches."""
    case = cases.random_case(shape=(16, 20, 24), seed=61, scheme="HYBRID", split=2, char=1, jump=0.5,
                             hybrid=(0.002, 0.02))
    monkeypatch.setenv("OCFD_NO_GRAPH", "1")
    eager = gpu_block(case)
    monkeypatch.delenv("OCFD_NO_GRAPH")
    with eager, gpu_block(case) as g:
        for rep in range(4):
            eager.step(1)
            g.step(1)
            g.show_flow_msg()
            g.synchronize()
            assert_same(g.download_f(), eager.download_f(), "f after replay %d" % rep)
        eager.step(3)
        g.step(3)
        assert_same(g.download_f(), eager.download_f(), "f after 3 more replays")
        pg, pe = g.download_primitives(), eager.download_primitives()
        for q in range(5):
            assert_same(pg[q], pe[q], "halo'd primitive %d" % q)


def test_negative_temperature_trap_and_repair():
    """Three isolated cells get negative energy: comput_duvwT reports them and
    replaces T by the mean of the six neighbours (OCFD_handle_NegativeT.f90)."""
    case = cases.random_case(shape=(12, 14, 16), seed=43)
    ob = pyoracle.setup_oracle(case)
    with gpu_block(case) as b:
        f = ob.download_f()
        for (k, j, i) in ((2, 3, 4), (7, 9, 11), (10, 2, 14)):
            f[4, k, j, i] = -1.0
        ob.upload_f(f)
        b.upload_f(f)
        assert ob.comput_duvwt() == 3
        assert b.comput_duvwt() == 3
        for q in range(5):
            assert_same(interior(b.download_primitives()[q]), interior(ob.download_primitives()[q]), "prim %d" % q)


@pytest.mark.slow
def test_tgv64_100_steps_against_oracle():
    """Config-1 TGV, 64^3, WENO5 + Steger-Warming, CD6, CFL 0.5, 100 steps."""
    case = cases.tgv(64)
    ob = pyoracle.setup_oracle(case)
    with gpu_block(case) as b:
        ke_o, ke_g, en_o, en_g = [], [], [], []
        for _ in range(10):
            ob.step(10)
            b.step(10)
            ao, ag = ob.ana_box(), b.ana_box()
            ke_o.append(ao[12]); ke_g.append(ag[12]); en_o.append(ao[13]); en_g.append(ag[13])
        pg, po = b.download_primitives(), ob.download_primitives()
        for q, nm in enumerate("duvwT"):
            assert rel_l2(interior(pg[q]), interior(po[q])) <= 1e-12, nm
        np.testing.assert_allclose(ke_g, ke_o, rtol=1e-10)
        np.testing.assert_allclose(en_g, en_o, rtol=1e-10)


def _rhs(blk, case):
    blk.save_fn()
    blk.exchange_primitives()
    if case.cfg.scheme_invis == "HYBRID":
        blk.comput_rhybrid()
    (blk.du_inviscous_character if case.cfg.if_scheme_character else blk.du_inviscous)()
    blk.du_viscous()
    return blk.download_du()


@pytest.mark.parametrize("kw", UNIFORM[:5], ids=[m["scheme"] + str(m["seed"]) for m in UNIFORM[:5]])
def test_sweep_tile_shapes_bitwise(kw):
    """Every sweep tile shape (the autotuner's choices) gives the same bits:
    one right-hand side and two whole steps, against shape 0."""
    case = cases.random_case(shape=(20, 28, 40), **kw)
    case.uniform_metrics = cases.uniform_box_metrics(case.cfg.nx, case.cfg.ny, case.cfg.nz)
    case.cfg.metric_mode = 1
    with gpu_block(case, uniform=True) as ref:
        du0 = _rhs(ref, case)
        ref.step(2)
        f0 = ref.download_f()
    for shapes in ([1, 1, 1], [2, 2, 2], [1, 2, 0], [3, 3, 3], [4, 4, 4], [1, 3, 4], [3, 0, 4]):
        with gpu_block(case, uniform=True) as b:
            b.set_sweep_tiles(shapes)
            assert b.get_sweep_tiles() == shapes
            assert_same(_rhs(b, case), du0, "du with tiles %s" % shapes)
            b.step(2)
            assert_same(b.download_f(), f0, "f after 2 steps with tiles %s" % shapes)


@pytest.mark.parametrize("kw", [UNIFORM[0], UNIFORM[4], UNIFORM[5]], ids=["HYBRIDc", "HYBRIDnc", "WENO5"])
def test_march_shapes_wide_block(kw):
    """The marches (shapes 3, 4) on a block whose i extent needs two
    wavefronts in y and z, the second one part idle, and whose extents end
    in partial segments and runs: bitwise equal to the tile sweeps."""
    kw = dict(kw)
    kw.pop("filters", None)
    case = cases.random_case(shape=(70, 36, 45), **kw)
    case.uniform_metrics = cases.uniform_box_metrics(case.cfg.nx, case.cfg.ny, case.cfg.nz)
    case.cfg.metric_mode = 1
    with gpu_block(case, uniform=True) as ref:
        du0 = _rhs(ref, case)
        ref.step(2)
        f0 = ref.download_f()
    for shapes in ([3, 3, 4], [4, 4, 3]):
        with gpu_block(case, uniform=True) as b:
            b.set_sweep_tiles(shapes)
            assert_same(_rhs(b, case), du0, "du with tiles %s" % shapes)
            b.step(2)
            assert_same(b.download_f(), f0, "f after 2 steps with tiles %s" % shapes)


def test_autotune_keeps_results():
    case = cases.random_case(shape=(20, 28, 40), seed=57, scheme="WENO7", split=2, char=1)
    case.uniform_metrics = cases.uniform_box_metrics(case.cfg.nx, case.cfg.ny, case.cfg.nz)
    case.cfg.metric_mode = 1
    with gpu_block(case, uniform=True) as ref, gpu_block(case, uniform=True) as b:
        shapes = b.autotune(reps=2)
        assert all(s in range(5) for s in shapes) and b.get_sweep_tiles() == shapes
        ref.step(2)
        b.step(2)
        assert_same(b.download_f(), ref.download_f(), "f after autotune + 2 steps (tiles %s)" % shapes)


@pytest.mark.parametrize("name", ["flat_plate", "swept_corner"])
def test_boundary_layer_metrics_with_user_ghosts(name):
    """Metrics computed on the device for the boundary-layer cases, whose
    Ghost_Cell = -1 sides take the user ghost layers (Jac_Ghost_Userdef,
    OCFD_GhostCell_UserDef.f90:4-106), equal the reference's init_mesh output
    (tests/golden/met_<name>.npz) bit for bit, ghost layers included."""
    import os
    from util import GOLDEN
    case = getattr(cases, name)()
    ref = np.load(os.path.join(GOLDEN, "met_%s.npz" % name), allow_pickle=False)["met"]
    with native.Block(case.cfg) as b:
        b.set_grid(case.x, case.y, case.z, case.spans)
        mets = b.download_metrics()
    for q, m in enumerate((9, 10, 14, 18)):  # Ajac, Akx1, Aiy1, Asz1
        assert_same(mets[m], ref[q], "metric %d" % m)
