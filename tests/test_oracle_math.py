"""Pins of the oracle that do not rely on the reference: closed forms,
identities and invariants the mathematics fixes, chosen so that a dropped term,
a wrong sign or a transposed index fails one of them."""
import numpy as np
import pytest

import pyoracle
from ocfd import cases
from ocfd.config import CaseConfig

rng = np.random.default_rng(12345)


def pw(name, v, center, *extra):
    return pyoracle.pointwise(name, v, center, *extra)


# ---------------------------------------------------------------- reconstructions
@pytest.mark.parametrize("name,deg", [("hh_weno5p", 2), ("hh_weno7p", 3)])
def test_weno_polynomial_reproduction(name, deg):
    """Every candidate stencil is exact for polynomials of degree <= r-1, so the
    face value equals the exact primitive-function value whatever the weights:
    v=k -> 1/2, v=k^2 -> 1/6, v=k^3 -> 0 at the face k=1/2."""
    k = np.arange(-4, 5, dtype=float)
    exact = {1: 0.5, 2: 1.0 / 6.0, 3: 0.0}
    for p in range(1, deg + 1):
        for scale, shift in ((1.0, 0.0), (3.0, 7.5)):
            v = scale * k ** p + shift
            got = pw(name, v, 4)
            assert abs(got - (scale * exact[p] + shift)) < 1e-12 * max(1.0, abs(shift) + scale)


def test_weno_small_amplitude_is_linear_upwind_scheme():
    """For smooth data with smoothness indicators << eps the weights are the
    ideal ones: WENO5 -> (2,-13,47,27,-3)/60, WENO7 -> (-3,25,-101,319,214,-38,4)/420."""
    x = np.linspace(0, 1, 9)
    v = 1e-6 * np.sin(2.3 * x + 0.4) + 2e-6 * np.cos(1.1 * x)
    w5 = np.dot([2, -13, 47, 27, -3], v[2:7]) / 60.0
    w7 = np.dot([-3, 25, -101, 319, 214, -38, 4], v[1:8]) / 420.0
    assert abs(pw("hh_weno5p", v, 4) - w5) < 1e-4 * np.abs(v).max() * 1e-2
    assert abs(pw("hh_weno7p", v, 4) - w7) < 1e-4 * np.abs(v).max() * 1e-2


def test_weno_essentially_non_oscillatory_at_a_jump():
    """A jump right of the face: the upwind smooth candidate dominates, so the
    face value stays within the left plateau's range."""
    v = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 10.0, 10.0, 10.0, 10.0])
    assert abs(pw("hh_weno5p", v, 4) - 1.0) < 1e-9
    assert abs(pw("hh_weno7p", v, 4) - 1.0) < 1e-9


@pytest.mark.parametrize("p,m", [("hh_weno5p", "hh_weno5m"), ("hh_weno7p", "hh_weno7m")])
def test_weno_mirror_symmetry_bitwise(p, m):
    """M at face i+1/2 is P applied to the mirrored data v'(k) = v(1-k)."""
    for _ in range(50):
        v = rng.standard_normal(9)
        vr = v[::-1].copy()  # vr[4+k] = v[4-k]  ->  v(1-k) at index 4+k  needs centre 3
        assert pw(m, v, 4) == pw(p, vr, 3)


def test_omp6_mirror_differs_only_through_the_limiter():
    """OMP6 P/M agree on data where the MP limiter is inactive (smooth data);
    the reference's M limiter uses a different curvature term, so on rough
    data they may differ."""
    x = np.linspace(0, 1, 9)
    v = np.sin(0.7 * x)
    assert abs(pw("hh_omp6m", v, 4) - pw("hh_omp6p", v[::-1].copy(), 3)) < 1e-14


def test_weno5_boundary_closure_linear():
    k = np.arange(-4, 5, dtype=float)
    for del_ in (1, 2):
        assert abs(pw("hh_weno5p_boundary", 2.0 * k + 1.0, 4, del_) - 2.0) < 1e-13
        assert abs(pw("hh_weno5m_boundary", 2.0 * k + 1.0, 4, del_) - 2.0) < 1e-13


# ---------------------------------------------------------------- flux splitting
def _euler_flux(gamma, c, d, u, v, w, a1, a2, a3):
    p = d * c * c / gamma
    un = a1 * u + a2 * v + a3 * w
    E = p / (gamma - 1.0) + 0.5 * d * (u * u + v * v + w * w)
    return np.array([d * un, d * u * un + a1 * p, d * v * un + a2 * p, d * w * un + a3 * p, (E + p) * un])


@pytest.mark.parametrize("which", [1, 2])
def test_split_sums_to_projected_flux(which):
    import ctypes
    n, lap = 6, 4
    npt = n + 2 * lap
    g = 1.4
    c, d = 0.8 + rng.random(npt), 0.5 + rng.random(npt)
    u, v, w = rng.standard_normal((3, npt))
    a1, a2, a3 = rng.standard_normal((3, npt))
    fp = np.zeros(5 * npt)
    fm = np.zeros(5 * npt)
    lib = pyoracle.load()
    P = ctypes.POINTER(ctypes.c_double)
    args = [np.ascontiguousarray(a) for a in (c, d, u, v, w, a1, a2, a3)]
    lib.oracle_split(which, n, lap, ctypes.c_double(g), *[a.ctypes.data_as(P) for a in args],
                     fp.ctypes.data_as(P), fm.ctypes.data_as(P))
    F = _euler_flux(g, c, d, u, v, w, a1, a2, a3)
    s = (fp + fm).reshape(5, npt)
    np.testing.assert_allclose(s, F, rtol=1e-12, atol=1e-12 * np.abs(F).max())
    if which == 1:
        # all eigenvalues positive (un > c|A|): no negative part
        u2 = np.full(npt, 5.0)
        a = [np.ascontiguousarray(x) for x in (np.full(npt, 0.5), d, u2, 0 * u2, 0 * u2, np.ones(npt),
                                               np.zeros(npt), np.zeros(npt))]
        lib.oracle_split(1, n, lap, ctypes.c_double(g), *[x.ctypes.data_as(P) for x in a],
                         fp.ctypes.data_as(P), fm.ctypes.data_as(P))
        assert np.all(fm == 0.0)


def test_characteristic_matrices_are_inverse():
    import ctypes
    lib = pyoracle.load()
    P = ctypes.POINTER(ctypes.c_double)
    for _ in range(20):
        st = np.concatenate([rng.standard_normal(3), [0.5 + rng.random()], rng.standard_normal(3)])
        S = np.zeros(25)
        S1 = np.zeros(25)
        lib.oracle_char_matrices(ctypes.c_double(1.4), st.ctypes.data_as(P), S.ctypes.data_as(P),
                                 S1.ctypes.data_as(P))
        np.testing.assert_allclose(S.reshape(5, 5) @ S1.reshape(5, 5), np.eye(5), atol=1e-12)


# ---------------------------------------------------------------- central derivatives
def _line_block(vis, n=20):
    cfg = CaseConfig(nx=n, ny=8, nz=8, iperiodic=(0, 1, 1), scheme_vis=vis)
    return pyoracle.OracleBlock(cfg)


@pytest.mark.parametrize("vis,deg", [("CD6", 6), ("CD8", 8)])
def test_central_difference_polynomial_exactness(vis, deg):
    n = 20
    b = _line_block(vis, n)
    L = 4
    h = 1.0 / (n - 1)
    xi = (np.arange(1 - L, n + L + 1) - 1) * h
    for p in range(deg + 1):
        line = xi ** p
        f = np.ones(b.halo_shape)
        f[:, :, :] = line[None, None, :]
        d = b.dxyz0(0, f)[0, 0, :]
        exact = p * xi[L:L + n] ** (p - 1) if p else np.zeros(n)
        inner = slice(4 if vis == "CD8" else 3, n - (4 if vis == "CD8" else 3))
        np.testing.assert_allclose(d[inner], exact[inner], rtol=1e-9, atol=1e-9)
        if p <= 2:  # one-sided / 2nd-order closures exact for quadratics
            np.testing.assert_allclose(d, exact, rtol=1e-9, atol=1e-9)


def test_cd6_convergence_rate_on_cosine():
    """Known-answer pattern of the reference's stale unit driver
    (src/test-dxyz0.f90:29-80): d/dx cos -> -sin, error O(h^6) (periodic)."""
    errs = []
    for n in (16, 32):
        cfg = CaseConfig(nx=n, ny=8, nz=8, iperiodic=(1, 1, 1), scheme_vis="CD6")
        b = pyoracle.OracleBlock(cfg)
        L = 4
        th = 2 * np.pi * (np.arange(1 - L, n + L + 1) - 1) / n
        f = np.ones(b.halo_shape)
        f[:, :, :] = np.cos(th)[None, None, :]
        d = b.dxyz0(0, f)[0, 0, :]
        # d/dxi with h = 1/(n-1): dtheta/dxi = 2*pi*(n-1)/n
        exact = -np.sin(th[L:L + n]) * 2 * np.pi * (n - 1) / n
        errs.append(np.abs(d - exact).max() / (2 * np.pi))
    rate = np.log2(errs[0] / errs[1])
    assert rate > 5.7


# ---------------------------------------------------------------- time advance
def test_rk3_amplification_factor():
    cfg = CaseConfig(nx=8, ny=8, nz=8, iperiodic=(1, 1, 1), dt=0.1)
    b = pyoracle.OracleBlock(cfg)
    lam = -1.3
    f0 = 1.0 + rng.random(b.f_shape)
    b.upload_f(f0)
    b.save_fn()
    for krk in (1, 2, 3):
        b.upload_du(lam * b.download_f())
        b.time_adv_rk3(krk)
    z = lam * cfg.dt
    np.testing.assert_allclose(b.download_f(), f0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14)


# ---------------------------------------------------------------- filter
def test_fo9p_transfer_function():
    """f = c + a(-1)^i: the 2-point wave is removed at s0 = 1 (D(pi) = 1) and
    the constant kept (D(0) = 0)."""
    n = 16
    cfg = CaseConfig(nx=n, ny=8, nz=8, iperiodic=(1, 1, 1),
                     filters=[[1, 1, 0, 0, 1, n, 1, 8, 1, 8, 1, 1.0, 0.0]])
    b = pyoracle.OracleBlock(cfg)
    i = np.arange(n)
    f = np.empty(b.f_shape)
    for m in range(5):
        f[m] = (2.0 + m) + 0.3 * (-1.0) ** i[None, None, :]
    b.upload_f(f)
    b.advance_clock()  # Istep = 1, filter period 1
    b.filtering()
    g = b.download_f()
    for m in range(5):
        np.testing.assert_allclose(g[m], 2.0 + m, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- whole steps
def test_uniform_flow_is_preserved():
    """Constant state: with constant metrics every face flux is the same
    number, so du is exactly zero; whole steps keep f to round-off (f is
    rebuilt from the primitives on the block faces)."""
    c = cases.random_case(shape=(8, 10, 12), seed=2)
    c.prims = [np.full_like(a, v) for a, v in zip(c.prims, (1.2, 0.3, -0.2, 0.1, 0.9))]
    c.uniform_metrics = cases.uniform_box_metrics(12, 10, 8)
    b = pyoracle.setup_oracle(c, uniform=True)
    f0 = b.download_f()
    b.exchange_primitives()
    b.du_inviscous()
    b.du_viscous()
    assert np.all(b.download_du() == 0.0)
    b.step(2)
    np.testing.assert_allclose(b.download_f(), f0, rtol=1e-15, atol=0)


@pytest.mark.parametrize("scheme,char", [("WENO5", 0), ("WENO7", 1), ("HYBRID", 1)])
def test_periodic_box_conserves_mass_momentum_energy(scheme, char):
    c = cases.random_case(shape=(10, 12, 14), seed=21, scheme=scheme, char=char, split=2)
    b = pyoracle.setup_oracle(c)
    f0 = b.download_f().reshape(5, -1).sum(axis=1)
    b.step(3)
    f1 = b.download_f().reshape(5, -1).sum(axis=1)
    scale = np.abs(b.download_f()).reshape(5, -1).sum(axis=1)
    assert np.all(np.abs(f1 - f0) <= 1e-12 * scale)


def test_tgv_initial_kinetic_energy_and_enstrophy():
    c = cases.tgv(32)
    b = pyoracle.setup_oracle(c)
    a = b.ana_box()
    assert abs(a[12] - 0.125) < 1e-14        # KE = 1/8 (exact on the uniform grid)
    assert abs(a[13] - 0.375) / 0.375 < 2e-4  # enstrophy = 3/8 up to CD6 error


def test_tgv_early_decay_matches_viscous_dissipation():
    """dK/dt = -2 nu Omega at t = 0 (incompressible limit, Ma = 0.1)."""
    c = cases.tgv(32, re=100.0)
    b = pyoracle.setup_oracle(c)
    a0 = b.ana_box()
    b.step(4)
    a1 = b.ana_box()
    dkdt = (a1[12] - a0[12]) / (4 * c.cfg.dt)
    expect = -2.0 / 100.0 * 0.5 * (a0[13] + a1[13])
    assert abs(dkdt - expect) / abs(expect) < 0.05
