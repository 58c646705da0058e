"""ocfd_show_flow_msg_noalloc, the show_flow_msg entry point that allocates
nothing (its partial sums live in a file-scope device array), against
ocfd_show_flow_msg on the same state: the same kernels sum in the same order,
so the results are bitwise equal.  Two calls in a row, the second on a
smaller block, so that the second call finds the first call's partial sums
in the shared array."""
import numpy as np
import pytest

from ocfd import cases, native
from ocfd.config import METRIC_UNIFORM, PREC_MIXED

pytestmark = pytest.mark.gpu


def _block(case, precision=None):
    cfg = case.cfg if precision is None else case.cfg.replace(precision=precision)
    return native.setup_block(native.Block(cfg), case)


def test_noalloc_flow_msg_matches_allocating_entry_point():
    big = cases.isotropic(64, scheme="WENO5", split=1, char=0, metric_mode=METRIC_UNIFORM)
    small = cases.random_case(shape=(20, 14, 10), seed=5, scheme="WENO7", split=2, char=1)
    print("no-alloc flow_msg: iso 64^3 (3 steps), then 20x14x10 (1 step)", flush=True)
    with _block(big) as b1, _block(small) as b2:
        b1.step(3)
        b2.step(1)
        got1 = b1.show_flow_msg_noalloc()
        want1 = b1.show_flow_msg()
        got2 = b2.show_flow_msg_noalloc()
        want2 = b2.show_flow_msg()
        assert np.array_equal(got1, want1), (got1, want1)
        assert np.array_equal(got2, want2), (got2, want2)
        assert not np.array_equal(got1, got2)
        # and the larger block again after the smaller one
        assert np.array_equal(b1.show_flow_msg_noalloc(), want1)


def test_noalloc_flow_msg_mixed_precision_block():
    case = cases.random_case(shape=(24, 16, 12), seed=9, scheme="HYBRID", split=2, char=1, jump=0.4)
    with _block(case, PREC_MIXED) as b:
        b.step(2)
        assert np.array_equal(b.show_flow_msg_noalloc(), b.show_flow_msg())
