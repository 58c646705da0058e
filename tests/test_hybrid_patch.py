"""The forced-WENO patches of the shock sensor (comput_Rhybrid,
OCFD_SetHybrid.f90:87-110): Hybrid_para(4) boxes, each given by
Hybrid_para(5:10) = ib, ie, jb, je, kb, ke (global, clipped to the block),
add 100 to Rhybrid inside the box so its faces use the most robust scheme.

The reference's loop reads para(8) into ie instead of je and leaves je
unset (OCFD_SetHybrid.f90:91-93), so its result is not defined; the oracle
and the device implement the box the parameter comments describe (line 1).
No reference fixture can pin that: these properties do.  They are chosen so
that the reference's own slip (para(8) as ie), a transposed i/j range, a
missing clip or an off-by-one bound each fails one of them:
  - Rhybrid with patches equals Rhybrid without them plus 100 exactly at the
    cells of the (clipped) boxes and is unchanged everywhere else;
  - the number of changed interior cells is the box volume;
  - the periodic halo carries the patched values.
"""
import numpy as np
import pytest

import pyoracle
from ocfd import cases

L = 4


def _rhyb(case, boxes):
    hp = dict(case.cfg.hybrid_para)
    hp[4] = float(len(boxes))
    for m, b in enumerate(boxes):
        for q, v in enumerate(b):
            hp[5 + 6 * m + q] = float(v)
    case.cfg.hybrid_para = hp
    ob = pyoracle.setup_oracle(case)
    ob.exchange_primitives()
    ob.comput_rhybrid()
    return ob.download_rhybrid()


def _case():
    return cases.random_case(shape=(12, 14, 16), seed=91, scheme="HYBRID", split=2, char=1, jump=0.5,
                             hybrid=(0.002, 0.02))


# (ib, ie, jb, je, kb, ke): an interior box with distinct extents per axis
# (ie != je, so reading para(8) as ie changes the i range), and a box that
# runs past the block on every side (clipped to 1..n)
BOXES = [(3, 9, 5, 6, 2, 10), (-2, 30, 12, 40, 11, 99)]


def test_patch_boxes_add_100_exactly_inside():
    case = _case()
    nz, ny, nx = case.cfg.shape
    r0 = _rhyb(_case(), [])
    r1 = _rhyb(case, BOXES)
    inside = np.zeros((nz, ny, nx), dtype=bool)
    for (ib, ie, jb, je, kb, ke) in BOXES:
        ib, jb, kb = max(ib, 1), max(jb, 1), max(kb, 1)
        ie, je, ke = min(ie, nx), min(je, ny), min(ke, nz)
        inside[kb - 1:ke, jb - 1:je, ib - 1:ie] = True
    i0, i1 = r0[L:-L, L:-L, L:-L], r1[L:-L, L:-L, L:-L]
    # the boxes do not overlap: each patched cell got exactly one +100
    np.testing.assert_array_equal(i1[inside], i0[inside] + 100.0)
    np.testing.assert_array_equal(i1[~inside], i0[~inside])
    assert int(np.count_nonzero(i1 != i0)) == 7 * 2 * 9 + 16 * 3 * 2


def test_patched_values_reach_the_periodic_halo():
    case = _case()
    r1 = _rhyb(case, BOXES[:1])
    nx = case.cfg.nx
    # x halo on the right holds interior i = 1..4, on the left i = nx-3..nx
    np.testing.assert_array_equal(r1[L:-L, L:-L, nx + L:], r1[L:-L, L:-L, L:2 * L])
    np.testing.assert_array_equal(r1[L:-L, L:-L, :L], r1[L:-L, L:-L, nx:nx + L])


@pytest.mark.gpu
def test_gpu_patched_sensor_and_steps_match_oracle():
    """The device sensor with the same patch boxes, and two whole steps
    through them, equal the oracle's bit for bit."""
    from ocfd import native
    case = _case()
    want = _rhyb(case, BOXES)  # sets the patch parameters on case.cfg too
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.exchange_primitives()
        b.comput_rhybrid()
        np.testing.assert_array_equal(b.download_rhybrid(), want)
    ob = pyoracle.setup_oracle(case)
    ob.step(2)
    with native.setup_block(native.Block(case.cfg), case) as b:
        b.step(2)
        np.testing.assert_array_equal(b.download_f(), ob.download_f())
