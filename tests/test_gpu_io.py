"""Restart / snapshot files written and read by the library
(ocfd_save_flow_data / ocfd_read_flow_data, the reference's save_flow_data /
read_flow_data, OCFD_IO.f90:5-64).

The fixture tests/golden/ref_snapshot_w5_sw.dat is the file the unmodified
reference wrote after two steps of the w5_sw golden case
(tests/golden/make_golden.py).  Reading it and writing it back must give the
same bytes; a snapshot carrying f must restart bit for bit.  The drop-in
driver writes its snapshots through the same call and tests/test_dropin.py
compares them with the reference's byte for byte.
"""
import os

import numpy as np
import pytest

from ocfd import fortran_io, native
from util import GOLDEN, interior, load_golden

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(GOLDEN, "ref_snapshot_w5_sw.dat")


def block_for(case):
    return native.setup_block(native.Block(case.cfg), case)


def test_snapshot_read_write_round_trip(tmp_path):
    case, _ = load_golden("w5_sw")
    with block_for(case) as b:
        assert b.read_flow_data(FIXTURE) is False
        istep, tt, prims = fortran_io.read_flow_data(FIXTURE, case.cfg.shape)
        assert b.get_time() == (istep, tt) and istep == 2
        got = b.download_primitives()
        for q in range(5):
            assert np.array_equal(interior(got[q]), prims[q]), "primitive %d" % q
        out = tmp_path / "copy.dat"
        b.save_flow_data(str(out))
        assert out.read_bytes() == open(FIXTURE, "rb").read()
        # f recomputed from the primitives, as at start-up
        f = b.download_f()
        d, u = prims[0], prims[1]
        assert np.array_equal(f[0], d) and np.array_equal(f[1], d * u)


def test_snapshot_default_name(tmp_path, monkeypatch):
    case, _ = load_golden("w5_sw")
    monkeypatch.chdir(tmp_path)
    with block_for(case) as b:
        b.step(3)
        b.save_flow_data()
        assert (tmp_path / "OCFD00000003.dat").exists()
        istep, _, prims = fortran_io.read_flow_data(str(tmp_path / "OCFD00000003.dat"), case.cfg.shape)
        assert istep == 3
        got = b.download_primitives()
        for q in range(5):
            assert np.array_equal(interior(got[q]), prims[q])


@pytest.mark.parametrize("name", ["hyb_c_llf", "curv_sb_ghost"])
def test_restart_with_f_is_bit_exact(tmp_path, name):
    case, _ = load_golden(name)
    path = str(tmp_path / "restart.dat")
    with block_for(case) as a:
        a.step(2)
        a.save_flow_data(path, with_f=True)
        # the reference's reader sees an ordinary snapshot
        istep, _, _ = fortran_io.read_flow_data(path, case.cfg.shape)
        assert istep == 2
        with block_for(case) as b:
            assert b.read_flow_data(path) is True
            b.bc()  # as the reference's start-up after read_flow_data (OCFD_NS_Solver.f90:15)
            a.step(2)
            b.step(2)
            assert np.array_equal(b.download_f(), a.download_f())
            pa, pb = a.download_primitives(), b.download_primitives()
            for q in range(5):
                assert np.array_equal(interior(pb[q]), interior(pa[q])), "primitive %d" % q
