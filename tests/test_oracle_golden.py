"""The C oracle against the golden vectors written by the compiled reference
(tests/golden/make_golden.py).  Every comparison is bitwise: the oracle keeps
the reference's operation order and is built without FMA contraction."""
import numpy as np
import pytest

import pyoracle
from util import GOLDEN, golden_names, load_golden

NAMES = golden_names()


def test_fixtures_present():
    assert len(NAMES) >= 12


@pytest.mark.parametrize("name", NAMES)
def test_oracle_stage1_and_steps_bitwise(name):
    case, ref = load_golden(name)
    cfg = case.cfg
    ob = pyoracle.setup_oracle(case)
    if "met" in ref:
        mets = ob.download_metrics()
        for q, m in enumerate((9, 10, 14, 18)):  # Ajac, Akx1, Aiy1, Asz1
            np.testing.assert_array_equal(mets[m], ref["met"][q])
    ob.save_fn()
    ob.exchange_primitives()
    if cfg.scheme_invis.upper() == "HYBRID":
        ob.comput_rhybrid()
        np.testing.assert_array_equal(ob.download_rhybrid()[4:-4, 4:-4, 4:-4], ref["rhyb"])
    if cfg.if_scheme_character:
        ob.du_inviscous_character()
    else:
        ob.du_inviscous()
    np.testing.assert_array_equal(ob.download_du(), ref["k1_duinv"])
    ob.du_viscous()
    np.testing.assert_array_equal(ob.download_du(), ref["k1_dutot"])
    ob.time_adv_rk3(1)
    np.testing.assert_array_equal(ob.download_f(), ref["k1_f"])
    nneg = [ob.comput_duvwt()]
    ob.bc()
    if "k1_T" in ref:
        # T <= 0 cases: the repaired temperature, halos included, and the
        # number of cells the reference reported
        np.testing.assert_array_equal(ob.download_primitives()[4], ref["k1_T"])
    for krk in (2, 3):
        nneg.append(ob.stage(krk))
    ob.advance_clock()
    ob.filtering()
    np.testing.assert_array_equal(ob.download_f(), ref["n1_f"])
    nneg.append(ob.step(2))
    np.testing.assert_array_equal(ob.download_f(), ref["nN_f"])
    if "k1_T" in ref:
        np.testing.assert_array_equal(ob.download_primitives()[4], ref["nN_T"])
        assert sum(nneg) == len(ref["negt_cells"]) and nneg[0] >= 2
    else:
        assert sum(nneg) == 0


@pytest.mark.parametrize("name", ["w5_sw", "hyb_c_llf", "curv_wall_fo9p"])
def test_oracle_diagnostics_match_reference_history(name):
    """E, K, S of show_flow_msg and KE / enstrophy of ana_box after each step,
    against the reference's values (printed with 17 significant digits)."""
    case, ref = load_golden(name)
    ob = pyoracle.setup_oracle(case)
    hist = ref["hist"]
    for row in hist:
        ob.step(1)
        e = ob.show_flow_msg()
        a = ob.ana_box()
        np.testing.assert_allclose(e, row[2:5], rtol=1e-14, atol=0)
        np.testing.assert_allclose([a[12], a[13]], row[5:7], rtol=1e-13, atol=0)


def test_reference_snapshot_fixture_layout():
    """The reference's snapshot file: (Istep, tt) record, then for each of
    d, u, v, w, T one record per k plane of nx*ny doubles (OCFD_IO.f90:44-64);
    read with the host reader, it holds the w5_sw case after two steps."""
    import os
    from ocfd import fortran_io
    case, _ = load_golden("w5_sw")
    path = os.path.join(GOLDEN, "ref_snapshot_w5_sw.dat")
    nz, ny, nx = case.cfg.shape
    assert os.path.getsize(path) == (4 + 12 + 4) + 5 * nz * (4 + nx * ny * 8 + 4)
    istep, tt, prims = fortran_io.read_flow_data(path, case.cfg.shape)
    assert istep == 2 and abs(tt - 2 * case.cfg.dt) <= 1e-15 * abs(tt)
    assert all(np.all(np.isfinite(p)) for p in prims) and np.all(prims[0] > 0) and np.all(prims[4] > 0)
