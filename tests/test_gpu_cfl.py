"""ocfd_cfl_dt: the device max-reduction of the CFL rate
sum_a (|u . A_a| + c |A_a|) / h_a, c = sqrt(T)/Ma, over the cells.

Checked against the host's dt (cases.cfl_dt, which the case generators fix
as Para%dt) on a uniform box, where the device's metrics and the host's
grid differences are the same numbers, and against a numpy restatement of
the same expression from the downloaded state and metrics on a curvilinear
block, where the maximum over the cells must agree to rounding.
"""
import numpy as np
import pytest

from ocfd import cases, native

pytestmark = pytest.mark.gpu


def test_uniform_box_matches_host_dt():
    case = cases.isotropic(32, scheme="WENO5", split=1, char=0, metric_mode=1)
    with native.setup_block(native.Block(case.cfg), case, use_uniform=True) as b:
        dt = b.cfl_dt(0.5)
    want = cases.cfl_dt(case, 0.5)
    assert abs(dt - want) <= 1e-12 * want, (dt, want)


def test_curvilinear_block_matches_restatement():
    case = cases.random_case(shape=(20, 28, 40), seed=93, scheme="WENO5", split=1, char=1, grid="curv",
                             iperiodic=(1, 0, 1))
    with native.setup_block(native.Block(case.cfg), case) as b:
        dt = b.cfl_dt(0.7)
        d, u, v, w, T = (a[4:-4, 4:-4, 4:-4] for a in b.download_primitives())
        met = [m[4:-4, 4:-4, 4:-4] for m in b.download_metrics()]
    cfg = case.cfg
    c = np.sqrt(T) / cfg.ma
    rate = np.zeros_like(u)
    for a, n in enumerate((cfg.nx, cfg.ny, cfg.nz)):
        ax, ay, az = met[3 * a], met[3 * a + 1], met[3 * a + 2]
        hinv = 1.0 / (1.0 / (n - 1.0))
        rate = rate + (np.abs(u * ax + v * ay + w * az) + c * np.sqrt(ax * ax + ay * ay + az * az)) * hinv
    want = 0.7 / rate.max()
    assert abs(dt - want) <= 1e-14 * want, (dt, want)
