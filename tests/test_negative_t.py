"""The T <= 0 repair against the compiled reference (no GPU).

handle_NegativeT (OCFD_handle_NegativeT.f90:11) replaces T of a cell by the
mean of its six neighbours, inside the i-j-k loop of comput_duvwT
(OCFD_Time_Adv.f90:45-58) that overwrites T in place: at that moment the
-i/-j/-k neighbours hold this stage's T (repaired where they were repaired)
and the +i/+j/+k neighbours still hold the previous stage's.

tests/golden/ref_negt_w5_sw.npz is the reference's own run of
cases.negative_t_case(): strong rarefactions that make 3 to 6 cells T <= 0
in every RK stage of three steps, with the neighbours' T changing within
the stage.  These tests pin, on the CPU:
  - that restating the loop with the previous stage's +neighbours reproduces
    the reference's T after stage 1 bit for bit, while the same restatement
    with this stage's values for every neighbour (averaging after the whole
    update) does not: the fixture tells the two apart;
  - that the cells the reference reported are the cells whose T from the
    reference's own f after stage 1 is <= 0, in loop order.
The GPU path is held to the same fixture in tests/test_gpu_negative_t.py.
"""
import numpy as np

from util import interior, load_golden

L = 4


def _t_from_f(f, cfg):
    """comput_duvwT's T (OCFD_Time_Adv.f90:52) in the reference's operation
    order; f is (5, nz, ny, nx)."""
    cv = 1.0 / (cfg.gamma * (cfg.gamma - 1.0) * cfg.ma * cfg.ma)
    f1, f2, f3, f4, f5 = f
    u, v, w = f2 / f1, f3 / f1, f4 / f1
    return (f5 - (f2 * u + f3 * v + f4 * w) * 0.5) / (f1 * cv)


def _repair(t_old_halo, t_new, old_plus_neighbours):
    """The i-j-k loop of comput_duvwT on a halo'd T array (numpy, one cell at
    a time).  old_plus_neighbours=True is the reference's in-place loop;
    False averages this stage's values of every neighbour."""
    T = t_old_halo.copy()
    nz, ny, nx = t_new.shape
    if not old_plus_neighbours:
        T[L:-L, L:-L, L:-L] = t_new
    cells = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                K, J, I = k + L, j + L, i + L
                T[K, J, I] = t_new[k, j, i]
                if T[K, J, I] <= 0:
                    T[K, J, I] = (T[K, J, I + 1] + T[K, J, I - 1] + T[K, J + 1, I] + T[K, J - 1, I] +
                                  T[K + 1, J, I] + T[K - 1, J, I]) / 6.0
                    cells.append((i + 1, j + 1, k + 1))
    return T, cells


def _stage1_start_T(case):
    """T at the start of stage 1: the initial field, periodic halos filled by
    exchange_boundary_xyz (OCFD_mpi.f90:183-233, x then y then z)."""
    t = np.ones(case.cfg.halo_shape)
    t[L:-L, L:-L, L:-L] = case.prims[4]
    for ax in (2, 1, 0):
        n = t.shape[ax] - 2 * L
        idx = lambda a, b: tuple(slice(a, b) if q == ax else slice(None) for q in range(3))  # noqa: E731
        t[idx(n + L, n + 2 * L)] = t[idx(L, 2 * L)]
        t[idx(0, L)] = t[idx(n, n + L)]
    return t


def test_fixture_tells_the_neighbour_semantics_apart():
    case, ref = load_golden("negt_w5_sw")
    assert all(case.cfg.iperiodic) and case.cfg.boundary_condition == "BC_NONE"
    t_new = _t_from_f(ref["k1_f"], case.cfg)
    t0 = _stage1_start_T(case)
    got, cells = _repair(t0, t_new, True)
    wrong, _ = _repair(t0, t_new, False)
    np.testing.assert_array_equal(interior(got), interior(ref["k1_T"]))
    repaired = interior(got) != t_new
    assert 2 <= int(repaired.sum()) <= 10
    assert np.any(interior(wrong) != interior(ref["k1_T"])), "fixture does not discriminate"


def test_reported_cells_are_the_t_le_0_cells_in_loop_order():
    case, ref = load_golden("negt_w5_sw")
    t_new = _t_from_f(ref["k1_f"], case.cfg)
    _, cells = _repair(_stage1_start_T(case), t_new, True)
    rep = [tuple(c) for c in ref["negt_cells"]]
    assert len(cells) >= 2 and rep[:len(cells)] == cells
    # 9 stages, each with 3 to 6 cells
    assert 27 <= len(rep) <= 54
