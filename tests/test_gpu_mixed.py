"""Mixed-precision form (OCFD_PREC_MIXED) against the FP64 path.

The mixed form forms the flux splitting, the reconstructions of the sweeps and
the viscous derivatives / stresses / fluxes in FP32, accumulates du in FP32
and stores the primitives as FP32; f, fn, the RK3 stage combination and the
metrics stay FP64, and so do the characteristic basis and projections of the
faces the hybrid sensor flags.  Its results are
compared with the FP64 path, which the other GPU tests pin bit for bit to the
compiled reference and the oracle, and with the oracle directly.

Tolerances, relative L2 per field:
  whole steps of periodic isotropic turbulence (Mt = 0.5, the mixed form's
    target case): <= 1e-5 after 20 steps with characteristic WENO7 (the
    accuracy bound the mixed form is specified to), <= 1e-4 with the hybrid
    scheme, whose face scheme is a discontinuous function of the state: a
    face whose sensor value lies within the mixed form's perturbation of a
    threshold switches between UD7L and WENO7, changing its flux by the
    difference of the two reconstructions (a few such switches on this
    strongly shocked 64^3 field raise the difference by an order of
    magnitude; the bench reports the 512^3 case);
    kinetic energy and enstrophy <= 1e-6 relative in both;
  one right-hand side du on random blocks with block-edge closures and
    curvilinear metrics: <= 2e-4 -- du is a difference of neighbouring face
    fluxes, so FP32 rounding of the fluxes (relative 6e-8) is amplified by
    |flux| / (|du| h), up to a few thousand on these blocks.
Each comparison also requires a nonzero difference, i.e. that the FP32
kernels ran.
"""
import numpy as np
import pytest

import pyoracle
from ocfd import cases, native
from ocfd.config import PREC_MIXED
from util import interior, rel_l2

pytestmark = pytest.mark.gpu


def block(case, precision=0, uniform=None):
    cfg = case.cfg.replace(precision=precision)
    return native.setup_block(native.Block(cfg), case, use_uniform=uniform)


def rhs(b, case):
    b.save_fn()
    b.exchange_primitives()
    if case.cfg.scheme_invis == "HYBRID":
        b.comput_rhybrid()
    (b.du_inviscous_character if case.cfg.if_scheme_character else b.du_inviscous)()
    if case.cfg.if_viscous == 1:
        b.du_viscous()
    return b.download_du()


RHS = [
    dict(seed=71, scheme="WENO5", split=1, char=0),
    dict(seed=72, scheme="WENO7", split=2, char=1),
    dict(seed=73, scheme="HYBRID", split=2, char=1, jump=0.5, hybrid=(0.002, 0.02)),
    dict(seed=74, scheme="OMP6", split=1, char=1, noise=0.03),
    dict(seed=75, scheme="UD7L", split=2, char=0, vis="CD8"),
    dict(seed=76, scheme="WENO5", split=1, char=1, grid="curv", iperiodic=(1, 0, 1),
         scheme_boundary=(0, 0, 1, 0, 0, 0)),
]


@pytest.mark.parametrize("kw", RHS, ids=[m["scheme"] + str(m["seed"]) for m in RHS])
def test_mixed_rhs_close_to_fp64(kw):
    case = cases.random_case(shape=(20, 28, 40), **kw)
    with block(case) as b64, block(case, PREC_MIXED) as bmx:
        d64, dmx = rhs(b64, case), rhs(bmx, case)
    for m in range(5):
        err = rel_l2(dmx[m], d64[m])
        print("du(%d) relative L2 %.3e" % (m + 1, err))
        assert err <= 2e-4, "du(%d): %.3e" % (m + 1, err)
    assert np.any(dmx != d64), "mixed du equals FP64 du bit for bit: FP32 kernels did not run"


@pytest.mark.parametrize("scheme", ["HYBRID", "WENO7"])
def test_mixed_isotropic_steps_close_to_fp64(scheme):
    case = cases.isotropic(64, scheme=scheme, split=2, char=1, hybrid_thresholds=(0.01, 0.05), metric_mode=1)
    with block(case, uniform=True) as b64, block(case, PREC_MIXED, uniform=True) as bmx:
        for _ in range(4):
            b64.step(5)
            bmx.step(5)
            a64, amx = b64.ana_box(), bmx.ana_box()
            for q, nm in ((12, "kinetic energy"), (13, "enstrophy")):
                assert abs(amx[q] - a64[q]) <= 1e-6 * abs(a64[q]), (nm, amx[q], a64[q])
        p64, pmx = b64.download_primitives(), bmx.download_primitives()
    errs = [rel_l2(interior(pmx[q]), interior(p64[q])) for q in range(5)]
    print(scheme, "relative L2 after 20 steps (d u v w T):", " ".join("%.2e" % e for e in errs))
    assert max(errs) <= (1e-4 if scheme == "HYBRID" else 1e-5), errs
    assert max(errs) > 0.0, "mixed run equals FP64 run bit for bit: FP32 kernels did not run"


def test_mixed_against_oracle():
    """The same bound against the oracle itself, on a curvilinear block with
    field metrics, non-periodic in y (closures on one side, ghost cells on
    the other): the general kernels rather than the periodic ones."""
    case = cases.random_case(shape=(20, 28, 40), seed=77, scheme="WENO7", split=2, char=1, grid="curv",
                             iperiodic=(1, 0, 1), scheme_boundary=(0, 0, 1, 0, 0, 0),
                             ghost_cell=(0, 0, 2, 0, 0, 0), bc="BC_USER_DEF", bc_para={1: 1.05})
    ob = pyoracle.setup_oracle(case)
    with block(case, PREC_MIXED) as bmx:
        ob.step(3)
        bmx.step(3)
        pg, po = bmx.download_primitives(), ob.download_primitives()
    errs = [rel_l2(interior(pg[q]), interior(po[q])) for q in range(5)]
    print("relative L2 vs oracle after 3 steps:", " ".join("%.2e" % e for e in errs))
    assert max(errs) <= 1e-5, errs


def test_mixed_primitive_storage_and_transfers(tmp_path):
    """The mixed form stores the primitives in FP32.  Downloads return their
    FP32 rounding, halos included; f formed at the upload comes from the FP64
    values given, so it equals the FP64 path's f bit for bit; a snapshot
    written from the mixed block holds the stored values, and reading it
    back into a mixed block restores them and forms f from them as the FP64
    path does from the same file."""
    from ocfd import fortran_io
    case = cases.random_case(shape=(20, 28, 40), seed=78, scheme="WENO5", split=1, char=0)
    prims = case.halo_prims()

    def bare(precision):
        b = native.Block(case.cfg.replace(precision=precision))
        if case.cfg.metric_mode == 1:
            b.set_uniform_metrics(case.uniform_metrics)
        else:
            b.set_grid(case.x, case.y, case.z, case.spans)
        b.upload_primitives(prims, 0, 0.0)
        return b

    with bare(0) as b64, bare(PREC_MIXED) as bmx:
        pmx = bmx.download_primitives()
        for q in range(5):
            assert np.array_equal(pmx[q], prims[q].astype(np.float32).astype(np.float64)), "primitive %d" % q
        assert np.array_equal(bmx.download_f(), b64.download_f()), "f formed from the FP64 upload"
        path = str(tmp_path / "mixed.dat")
        bmx.save_flow_data(path)
        _, _, filed = fortran_io.read_flow_data(path, case.cfg.shape)
        for q in range(5):
            assert np.array_equal(filed[q], interior(pmx[q])), "snapshot field %d" % q
        assert bmx.read_flow_data(path) is False and b64.read_flow_data(path) is False
        back = bmx.download_primitives()
        for q in range(5):
            assert np.array_equal(back[q], pmx[q]), "read back %d" % q
        assert np.array_equal(bmx.download_f(), b64.download_f()), "f formed at the restart"


@pytest.mark.parametrize("name", ["flat_plate", "swept_corner"])
def test_mixed_boundary_layer_conditions(name):
    """The boundary-layer / swept-corner kernels and the user ghost layers on
    the mixed form's FP32 primitives: 3 steps with the same boundary data
    (inlet and upper planes of the initial field, no wall blowing) stay
    within the mixed form's bound of the FP64 path.  (The FP64 path itself is
    pinned to the reference on these cases by tests/test_dropin.py.)"""
    case = getattr(cases, name)()
    inlet = np.stack([p[:, :, 0] for p in case.prims])   # i = 1: (5, nz, ny)
    upper = np.stack([p[:, -1, :] for p in case.prims])  # j = ny: (5, nz, nx)
    out = []
    for precision in (0, PREC_MIXED):
        with native.Block(case.cfg.replace(precision=precision)) as b:
            b.set_grid(case.x, case.y, case.z, case.spans)
            b.upload_primitives(case.halo_prims(), 0, 0.0)
            b.set_bc_boundarylayer(0.0, inlet, upper)
            b.bc()
            b.step(3)
            out.append(b.download_primitives())
    # v and w are small next to u in these flows (w is 0 in the FP64 run of
    # the flat plate): the velocity components are measured against the
    # speed, d and T against themselves
    p64, pmx = [interior(x) for x in out[0]], [interior(x) for x in out[1]]
    speed = np.sqrt(p64[1] ** 2 + p64[2] ** 2 + p64[3] ** 2)
    errs = [rel_l2(pmx[0], p64[0])]
    errs += [float(np.linalg.norm(pmx[q] - p64[q]) / np.linalg.norm(speed)) for q in (1, 2, 3)]
    errs += [rel_l2(pmx[4], p64[4])]
    print(name, "relative L2 after 3 steps (d u v w T):", " ".join("%.2e" % e for e in errs))
    assert max(errs) <= 1e-5, errs
    assert max(errs) > 0.0, "mixed run equals FP64 run bit for bit: FP32 kernels did not run"
