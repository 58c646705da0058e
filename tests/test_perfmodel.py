"""The FLOP table of the perf model equals the counting oracle's counts."""
import pytest

import flop_table
from ocfd import perfmodel


@pytest.mark.parametrize("key,kw,which", [v for v in flop_table.VARIANTS
                                          if v[0] in ("WENO5/SW/nonchar", "WENO7/LLF/char",
                                                      "HYBRID-linear/LLF/char", "viscous/CD6")])
def test_flop_polynomial_exact_at_new_size(key, kw, which):
    n = 16
    assert flop_table.count(n, kw, which) == perfmodel.poly(key, n)


def test_table_is_current():
    # refit on the small sizes and compare every coefficient
    t = flop_table.table()
    for k, v in t.items():
        assert v == perfmodel.INVISCID_FIT[k], k
