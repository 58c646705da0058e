"""Initial conditions and grids for the benchmark and test cases.

A :class:`Case` bundles the run configuration, the grid (interior coordinates
and periodic spans, as the reference's init_mesh sees them) and the initial
primitive variables, and can write the three input files the reference solver
reads (opencfd2.in, opencfd.dat, OCFD-grid.dat).

Nondimensionalisation follows the reference: free-stream density, velocity and
temperature are 1, p = d*T/(gamma*Ma^2) (OCFD_SetHybrid.f90:32-41), the sound
speed is sqrt(T)/Ma (OCFD_inviscous.f90:67) and Cv = 1/(gamma(gamma-1)Ma^2)
(OCFD_readpara.f90:269).

Arrays are float64 with shape (nz, ny, nx) (Fortran order of (nx, ny, nz)).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import fortran_io
from .config import (GRID1D, GRID2D_PLANE, GRID3D, GRID_AND_JACOBIAN3D, GRID_BOX,
                     METRIC_UNIFORM, CaseConfig)

PI = 3.1415926535897932  # the reference's PI literal (OCFD_mesh.f90:330)


@dataclass
class Case:
    cfg: CaseConfig
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    spans: np.ndarray
    prims: list                              # [d, u, v, w, T] interior
    grid_payload: tuple = ()                 # what OCFD-grid.dat holds
    uniform_metrics: np.ndarray | None = None  # 10 values for METRIC_UNIFORM
    name: str = "case"
    info: dict = field(default_factory=dict)

    def halo_prims(self):
        """Primitives in halo'd arrays; halos hold 1.0 as after the reference's
        allocate_flow_data (OCFD_NS_Solver.f90:116)."""
        L = self.cfg.lap
        out = []
        for a in self.prims:
            h = np.ones(self.cfg.halo_shape)
            h[L:-L, L:-L, L:-L] = a
            out.append(h)
        return out

    def write_reference_inputs(self, directory, nsteps: int, ana_box_every: int = 0,
                               istep_save: int = 1000000):
        os.makedirs(directory, exist_ok=True)
        cfg = self.cfg
        end_time = nsteps * cfg.dt
        fortran_io.write_namelist(os.path.join(directory, "opencfd2.in"),
                                  cfg.namelist(end_time, istep_save=istep_save,
                                               ana_box_every=ana_box_every))
        fortran_io.write_flow_data(os.path.join(directory, "opencfd.dat"), 0, 0.0, self.prims)
        gp = os.path.join(directory, "OCFD-grid.dat")
        if cfg.grid_type == GRID1D:
            fortran_io.write_grid_1d(gp, *self.grid_payload)
        elif cfg.grid_type == GRID2D_PLANE:
            fortran_io.write_grid_2d_plane(gp, *self.grid_payload)
        elif cfg.grid_type in (GRID3D, GRID_AND_JACOBIAN3D):
            fields = [self.x, self.y, self.z]
            if cfg.grid_type == GRID_AND_JACOBIAN3D:
                fields += [np.full(cfg.shape, v) for v in self.uniform_metrics]
            fortran_io.write_grid_3d(gp, *fields)
        if "inlet1d" in self.info:
            # 1-d inlet profile of the boundary-layer conditions: a header line,
            # then y, d, u, v, T per j (init_bcdata_boundarylayer,
            # OCFD_BC_Boundarylayers.f90:186-196)
            with open(os.path.join(directory, "flow1d-inlet.dat"), "w") as fh:
                fh.write("variables = y d u v T\n")
                for row in self.info["inlet1d"]:
                    fh.write(" ".join("%.17e" % v for v in row) + "\n")


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
def box_coords(nx, ny, nz):
    """mesh_gen('box') (OCFD_mesh.f90:335-347): x_i = 2*PI/N * i, i = 1..N."""
    xs = 2.0 * PI / float(nx) * np.arange(1, nx + 1, dtype=np.float64)
    ys = 2.0 * PI / float(ny) * np.arange(1, ny + 1, dtype=np.float64)
    zs = 2.0 * PI / float(nz) * np.arange(1, nz + 1, dtype=np.float64)
    shape = (nz, ny, nx)
    # read-only broadcast views: consumers that need memory copy them
    X = np.broadcast_to(xs[None, None, :], shape)
    Y = np.broadcast_to(ys[None, :, None], shape)
    Z = np.broadcast_to(zs[:, None, None], shape)
    return X, Y, Z, (xs, ys, zs)


def uniform_box_metrics(nx, ny, nz, lx=2 * PI, ly=2 * PI, lz=2 * PI):
    """Exact metrics of a uniform Cartesian box on the computational grid
    h = 1/(n-1) (OCFD_readpara.f90:265-267): x_xi = (Lx/nx)*(nx-1) ...
    Returns Akx,Aky,Akz,Aix,Aiy,Aiz,Asx,Asy,Asz,Ajac."""
    xi = lx / nx * (nx - 1)
    yj = ly / ny * (ny - 1)
    zk = lz / nz * (nz - 1)
    jac = 1.0 / (xi * yj * zk)
    return np.array([jac * (yj * zk), 0.0, 0.0, 0.0, jac * (zk * xi), 0.0, 0.0, 0.0,
                     jac * (xi * yj), jac])


def _periodic_spans(iperiodic, lx, ly, lz):
    sp = np.zeros(9)
    if iperiodic[0]:
        sp[0] = lx
    if iperiodic[1]:
        sp[4] = ly
    if iperiodic[2]:
        sp[8] = lz
    return sp


# ---------------------------------------------------------------------------
# counter-based random numbers (so that a field drawn on N^3 and on (2N)^3
# share every common Fourier mode)
# ---------------------------------------------------------------------------
def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _keyed_normals(seed, kx, ky, kz, stream):
    """Standard normals keyed by (seed, kx, ky, kz, stream) via splitmix64 and
    Box-Muller."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)
               ^ (kx.astype(np.int64).astype(np.uint64) * np.uint64(0x9E3779B1))
               ^ (ky.astype(np.int64).astype(np.uint64) * np.uint64(0x85EBCA77))
               ^ (kz.astype(np.int64).astype(np.uint64) * np.uint64(0xC2B2AE3D))
               ^ (np.uint64(stream) * np.uint64(0x27D4EB2F165667C5)))
        a = _splitmix64(key)
        b = _splitmix64(a ^ np.uint64(0xD6E8FEB86659FD93))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u2 = ((b >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)


def _ifftn_real(a):
    """Real part of the inverse 3-D FFT, on all host cores."""
    try:
        import scipy.fft as sfft
        return np.real(sfft.ifftn(a, workers=os.cpu_count() or 1))
    except ImportError:
        return np.real(np.fft.ifftn(a))


def solenoidal_field(n, seed, k0=8.0, kmax=None, kcut=None):
    """Random solenoidal velocity on an n^3 periodic box with energy spectrum
    E(k) ~ k^4 exp(-2 k^2 / k0^2); normalised to <|u|^2> = 1.  Modes are drawn
    for 0 < |k| <= min(kmax, kcut) (kmax = n/2 - 1; kcut = 6 k0, where the
    amplitude has fallen to ~1e-15 of its peak), so the draw costs the same
    at any n and a field on N^3 and on (2N)^3 share every drawn mode."""
    if kmax is None:
        kmax = n // 2 - 1
    if kcut is None:
        kcut = 6.0 * k0
    kc = min(float(kmax), float(kcut))
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    sel = np.nonzero(np.abs(k1) <= kc)[0]          # the sub-cube holding the sphere
    ks = k1[sel]
    kz, ky, kx = np.meshgrid(ks, ks, ks, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    mask = (kk > 0) & (kk <= kc)
    comps = []
    for c in range(3):
        re, im = _keyed_normals(seed, kx, ky, kz, c)
        comps.append(re + 1j * im)
    ax, ay, az = comps
    k2 = np.where(kk > 0, kk * kk, 1.0)
    kd = (kx * ax + ky * ay + kz * az) / k2
    ax, ay, az = ax - kx * kd, ay - ky * kd, az - kz * kd
    amp = np.where(mask, kk * np.exp(-kk * kk / (k0 * k0)), 0.0)
    vel = []
    for a in (ax, ay, az):
        full = np.zeros((n, n, n), dtype=np.complex128)
        full[np.ix_(sel, sel, sel)] = amp * a
        vel.append(_ifftn_real(full))
        del full
    e = np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    s = 1.0 / np.sqrt(e)
    return [v * s for v in vel]


# ---------------------------------------------------------------------------
# cases
# ---------------------------------------------------------------------------
def tgv(n=64, ma=0.1, re=1600.0, pr=0.7, gamma=1.4, cfl=0.5, scheme="WENO5", split=1,
        char=0, scheme_vis="CD6", ny=None, nz=None, metric_mode=0) -> Case:
    """Taylor-Green vortex: u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
    rho = 1, p = p0 + (rho/16)(cos 2x + cos 2y)(cos 2z + 2), p0 = 1/(gamma Ma^2).
    Grid: the reference's generated box (Iflag_Gridtype = 0)."""
    ny = ny or n
    nz = nz or n
    X, Y, Z, _ = box_coords(n, ny, nz)
    d = np.ones_like(X)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    p0 = 1.0 / (gamma * ma * ma)
    p = p0 + (d / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    T = gamma * ma * ma * p / d
    cfg = CaseConfig(nx=n, ny=ny, nz=nz, iperiodic=(1, 1, 1), scheme_invis=scheme,
                     scheme_vis=scheme_vis, flux_splitting=split, if_scheme_character=char,
                     ma=ma, re=re, pr=pr, gamma=gamma, grid_type=GRID_BOX,
                     periodic_span=tuple(_periodic_spans((1, 1, 1), 2 * PI, 2 * PI, 2 * PI)),
                     metric_mode=metric_mode)
    case = Case(cfg=cfg, x=X, y=Y, z=Z, spans=np.array(cfg.periodic_span),
                prims=[d, u, v, w, T], name="tgv%d" % n)
    if metric_mode == METRIC_UNIFORM:
        case.uniform_metrics = uniform_box_metrics(n, ny, nz)
        cfg.grid_type = GRID_AND_JACOBIAN3D
    cfg.dt = cfl_dt(case, cfl)
    return case


def isotropic(n=256, seed=20261202, k0=8.0, mt=0.5, re_lambda=72.0, cfl=0.5,
              scheme="HYBRID", split=2, char=1, hybrid_thresholds=(0.02, 0.1),
              metric_mode=0, re=None) -> Case:
    """Decaying compressible isotropic turbulence: random solenoidal velocity
    with E(k) ~ k^4 exp(-2k^2/k0^2), uniform rho = T = 1, scaled so that
    <|u|^2> = 1 and M_t = sqrt(<|u|^2>)/<c> = mt (so Ma = mt), and Re chosen so
    that Re_lambda (ana_box definition, OCFD_ANA.f90:265-272) equals re_lambda
    (or Re = re when given).
    Grid: uniform box (Iflag_Gridtype = 0)."""
    X, Y, Z, _ = box_coords(n, n, n)
    u, v, w = solenoidal_field(n, seed, k0)
    ma = mt
    # Re_lambda = VV_rms^2 / Ux_rms * Re (mu(T=1) = 1/Re, rho = 1)
    kx = np.fft.fftfreq(n, d=1.0 / n)[None, None, :]
    try:
        import scipy.fft as sfft
        uh = sfft.fftn(u, workers=os.cpu_count() or 1)
    except ImportError:
        uh = np.fft.fftn(u)
    ux = _ifftn_real(1j * kx * uh)
    del uh
    ux_rms = np.sqrt(np.mean(ux * ux))
    vv2 = np.mean(u * u + v * v + w * w) / 3.0
    if re is None:
        re = re_lambda * ux_rms / vv2
    d = np.ones_like(X)
    T = np.ones_like(X)
    cfg = CaseConfig(nx=n, ny=n, nz=n, iperiodic=(1, 1, 1), scheme_invis=scheme,
                     scheme_vis="CD6", flux_splitting=split, if_scheme_character=char,
                     ma=ma, re=re, grid_type=GRID_BOX,
                     hybrid_para={1: 1.0, 2: hybrid_thresholds[0], 3: hybrid_thresholds[1], 4: 0.0},
                     periodic_span=tuple(_periodic_spans((1, 1, 1), 2 * PI, 2 * PI, 2 * PI)),
                     metric_mode=metric_mode)
    case = Case(cfg=cfg, x=X, y=Y, z=Z, spans=np.array(cfg.periodic_span),
                prims=[d, u, v, w, T], name="iso%d" % n,
                info={"seed": seed, "k0": k0, "Re": re, "Ux_rms": ux_rms})
    if metric_mode == METRIC_UNIFORM:
        case.uniform_metrics = uniform_box_metrics(n, n, n)
        cfg.grid_type = GRID_AND_JACOBIAN3D
    cfg.dt = cfl_dt(case, cfl)
    return case


def _smooth_random(shape, rng, nmodes=6, lens=(2 * PI,) * 3, periodic=(1, 1, 1)):
    nz, ny, nx = shape
    zs = np.arange(nz)[:, None, None] / nz
    ys = np.arange(ny)[None, :, None] / ny
    xs = np.arange(nx)[None, None, :] / nx
    out = np.zeros(shape)
    for _ in range(nmodes):
        k = rng.integers(1, 4, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        a = rng.uniform(-1, 1)
        out += a * np.cos(2 * np.pi * k[0] * xs + ph[0]) * np.cos(2 * np.pi * k[1] * ys + ph[1]) \
            * np.cos(2 * np.pi * k[2] * zs + ph[2])
    return out / nmodes


def random_case(shape=(10, 12, 16), seed=1, scheme="WENO5", split=1, char=0, vis="CD6",
                iperiodic=(1, 1, 1), ma=0.8, re=300.0, noise=2e-3, jump=0.0,
                grid="box", scheme_boundary=(0, 0, 0, 0, 0, 0), ghost_cell=(0,) * 6,
                bc="BC_NONE", bc_para=None, filters=(), hybrid=(0.02, 0.1),
                mass_force=None, dt=None, sensor=1) -> Case:
    """Seeded, non-cubic test case exercising nonlinear weights (smooth field
    plus white noise), optionally a density/temperature jump (shock sensor),
    a curvilinear grid and non-periodic boundaries."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    if grid == "box":
        X, Y, Z, axes = box_coords(nx, ny, nz)
        gtype, payload = GRID_BOX, ()
        lens = (2 * PI, 2 * PI, 2 * PI)
    elif grid == "1d":
        lens = (2.0, 1.5, 1.0)
        xs = lens[0] * np.arange(nx) / (nx if iperiodic[0] else nx - 1)
        if not iperiodic[0]:
            xs = xs * (1.0 + 0.05 * np.sin(np.pi * np.arange(nx) / nx))
        ys = lens[1] * np.arange(ny) / (ny if iperiodic[1] else ny - 1)
        zs = lens[2] * np.arange(nz) / (nz if iperiodic[2] else nz - 1)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        X, Y, Z = X.copy(), Y.copy(), Z.copy()
        gtype, payload = GRID1D, (xs, ys, zs)
    elif grid == "curv":
        # smooth curvilinear mapping, periodic where requested
        lens = (2.0, 1.5, 1.0)
        xi = np.arange(nx) / (nx if iperiodic[0] else nx - 1)
        et = np.arange(ny) / (ny if iperiodic[1] else ny - 1)
        ze = np.arange(nz) / (nz if iperiodic[2] else nz - 1)
        ZE, ET, XI = np.meshgrid(ze, et, xi, indexing="ij")
        X = lens[0] * XI + 0.04 * np.sin(2 * np.pi * ET) * np.cos(2 * np.pi * ZE)
        Y = lens[1] * ET + 0.03 * np.sin(2 * np.pi * XI) + 0.02 * np.sin(2 * np.pi * ZE)
        Z = lens[2] * ZE + 0.02 * np.sin(2 * np.pi * XI) * np.sin(2 * np.pi * ET)
        gtype, payload = GRID3D, ()
    else:
        raise ValueError(grid)
    s = [_smooth_random(shape, rng) for _ in range(5)]
    wn = [rng.uniform(-1, 1, size=shape) for _ in range(5)]
    d = 1.0 + 0.2 * s[0] + noise * wn[0]
    u = 0.6 + 0.5 * s[1] + noise * wn[1]
    v = 0.5 * s[2] + noise * wn[2]
    w = 0.5 * s[3] + noise * wn[3]
    T = 1.0 + 0.2 * s[4] + noise * wn[4]
    if jump:
        step = 0.5 * (1.0 + np.tanh((np.arange(nx)[None, None, :] - nx / 2.0) / 0.7))
        d = d * (1.0 + jump * step)
        T = T * (1.0 + 0.5 * jump * step)
    cfg = CaseConfig(nx=nx, ny=ny, nz=nz, iperiodic=tuple(iperiodic), scheme_invis=scheme,
                     scheme_vis=vis, flux_splitting=split, if_scheme_character=char,
                     ma=ma, re=re, grid_type=gtype, scheme_boundary=tuple(scheme_boundary),
                     ghost_cell=tuple(ghost_cell), boundary_condition=bc,
                     bc_para=dict(bc_para or {}), filters=[list(f) for f in filters],
                     hybrid_para={1: float(sensor), 2: hybrid[0], 3: hybrid[1], 4: 0.0},
                     periodic_span=tuple(_periodic_spans(iperiodic, *lens)))
    if mass_force is not None:
        cfg.if_mass_force = 1
        cfg.mass_force = tuple(mass_force)
    case = Case(cfg=cfg, x=X, y=Y, z=Z, spans=np.array(cfg.periodic_span),
                prims=[d, u, v, w, T], grid_payload=payload, name="rand%d" % seed)
    cfg.dt = dt if dt is not None else 0.2 * cfl_dt(case, 1.0)
    return case


def negative_t_case(shape=(8, 10, 14), seed=44, amp=4.0, t_factor=0.1, dt_factor=4.0,
                    centres=((2, 3, 4), (5, 6, 9), (6, 2, 11)), scheme="WENO5", split=1, char=0) -> Case:
    """A seeded periodic test case (random_case) with strong local rarefactions:
    at each centre (k, j, i) u drops by `amp` at i-1, i-2 and rises by `amp`
    at i+1, i+2, and T is scaled by `t_factor`; dt is `dt_factor` times
    random_case's.  The RK stages then produce a few cells with T <= 0 in
    the fluid (not planted in f), whose neighbours also change within the
    stage, which is what handle_NegativeT's treatment of the +i/+j/+k
    neighbours depends on (OCFD_handle_NegativeT.f90:11).  With the defaults
    every stage of the first three steps repairs 3 to 6 cells; amp=6,
    t_factor=1, dt_factor=5 exceeds 10 cells in the first stage of step 2."""
    case = random_case(shape=shape, seed=seed, scheme=scheme, split=split, char=char)
    d, u, v, w, T = case.prims
    nx = shape[2]
    for (k, j, i) in centres:
        for s in (1, 2):
            u[k, j, (i - s) % nx] -= amp
            u[k, j, (i + s) % nx] += amp
        T[k, j, i] *= t_factor
    case.cfg.dt *= dt_factor
    case.name = "negt%d" % seed
    return case


def blunt_ogrid(nx=512, ny=256, nz=8, r0=1.0, dr0=None, ratio=1.05, zspan=0.1, ma=6.0,
                re=1.0e5, t_inf=79.0, t_wall=None, cfl=0.3, filter_every=1, filter_s0=0.2,
                wake_at_rest=True, layer_cells=24.0) -> Case:
    """Mach-6 flow over a cylinder: O-grid (i around the body, periodic; j
    radial, wall at j=1 with ratio-`ratio` stretching; z periodic).
    Isothermal wall through the BC_USER_DEF hook with ghost cells at j-,
    WENO5 characteristic + Steger-Warming, Fo9p filtering.

    Grid: first radial spacing dr0 (1e-3 r0), growing by `ratio` per cell up
    to the circumferential spacing at 2 r0, then uniform (outer radius about
    6.1 r0 at 512 x 256); z span `zspan`.  Wall temperature t_wall = 300 K /
    T_inf.  Fo9p filter with s0 = filter_s0 every `filter_every` steps.

    Initial state: free stream (u = 1) upstream of the body and fluid at rest
    behind it, blended by 0.5 (1 - tanh(x / r0)), inside a wall layer (u and
    T rise from 0 and t_wall over about `layer_cells` radial cells, uniform
    pressure).  A uniform Mach-6 start (wake_at_rest=False) is not runnable
    on the full O-grid: behind the body the flow leaves the wall faster than
    the escape speed 2c/(gamma - 1) = 5c of an unsteady expansion, so T
    reaches 0 in the first stage (the reference stops there too).  With
    these defaults (layer of 24 cells, CFL 0.3) the reference runs the first
    100 steps without a T <= 0 cell, so the run does not depend on how the
    reference's per-rank T <= 0 count and repair see its MPI decomposition."""
    if dr0 is None:
        dr0 = 1.0e-3 * r0
    # radial spacing: dr0 at the wall, growing by `ratio` per cell up to the
    # circumferential spacing at 2 r0, then uniform
    dr_max = 2.0 * PI * (2.0 * r0) / nx
    th = 2.0 * PI * np.arange(nx) / nx
    r = np.empty(ny)
    r[0] = r0
    for j in range(1, ny):
        r[j] = r[j - 1] + min(dr0 * ratio ** (j - 1), dr_max)
    x2 = -(r[:, None] * np.cos(th[None, :]))
    y2 = r[:, None] * np.sin(th[None, :])
    z1 = zspan * np.arange(nz) / nz
    X = np.broadcast_to(x2, (nz, ny, nx)).copy()
    Y = np.broadcast_to(y2, (nz, ny, nx)).copy()
    Z = np.broadcast_to(z1[:, None, None], (nz, ny, nx)).copy()
    if t_wall is None:
        t_wall = 300.0 / t_inf
    shape = (nz, ny, nx)
    v = np.zeros(shape)
    w = np.zeros(shape)
    if wake_at_rest:
        delta = layer_cells * dr0
        layer = np.tanh((r - r0) / delta)[:, None] * np.ones((1, nx))
        blend = 0.5 * (1.0 - np.tanh(x2 / r0))
        T2 = t_wall + (1.0 - t_wall) * layer
        u = np.broadcast_to(layer * blend, shape).copy()
        T = np.broadcast_to(T2, shape).copy()
        d = 1.0 / T
    else:
        d = np.ones(shape)
        u = np.ones(shape)
        T = np.ones(shape)
        # start from free stream with a no-slip, isothermal layer at the wall
        u[:, 0, :] = 0.0
        T[:, 0, :] = t_wall
    cfg = CaseConfig(nx=nx, ny=ny, nz=nz, iperiodic=(1, 0, 1), scheme_invis="WENO5",
                     scheme_vis="CD6", flux_splitting=1, if_scheme_character=1, ma=ma, re=re,
                     ref_amu_t0=t_inf, grid_type=GRID2D_PLANE, boundary_condition="BC_USER_DEF",
                     bc_para={1: t_wall}, scheme_boundary=(0, 0, 1, 0, 0, 0),
                     ghost_cell=(0, 0, 2, 0, 0, 0),
                     filters=[[filter_every, 1, 1, 1, 1, nx, 1, ny, 1, nz, 1, filter_s0, 0.0]],
                     periodic_span=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, zspan))
    case = Case(cfg=cfg, x=X, y=Y, z=Z, spans=np.array(cfg.periodic_span), prims=[d, u, v, w, T],
                grid_payload=(x2, y2, z1), name="blunt%dx%dx%d" % (nx, ny, nz),
                info={"dr0": dr0, "dr_max": dr_max, "r_outer": float(r[-1]), "ratio": ratio, "t_wall": t_wall,
                      "zspan": zspan})
    cfg.dt = cfl_dt(case, cfl)
    return case


def _wall_layer(yw, delta, t_wall):
    """Laminar-like profile over a wall at distance yw: u = tanh(y/delta),
    T from the wall temperature to 1, uniform pressure (d = 1/T)."""
    u = np.tanh(yw / delta)
    T = t_wall + (1.0 - t_wall) * u
    return 1.0 / T, u, T


def flat_plate(nx=24, ny=20, nz=8, x0=0.1, lx=1.0, y_top=0.25, rise=0.15, zspan=0.2, ma=2.25, re=1.0e4,
               t_wall=1.5, delta=0.03, cfl=0.4, scheme="WENO5", split=1, char=1) -> Case:
    """Supersonic flat-plate boundary layer with BoundaryCondition =
    BC_BoundaryLayer (ocfd_bc_BoundaryLayer, OCFD_BC_Boundarylayers.f90:17-167):
    1-d inlet profile file, free-stream upper boundary whose non-reflecting
    points are chosen from its slope (flat over the first half, rising by
    `rise` over the second), isothermal wall with multi-frequency
    blow-and-suction (random phases, drawn by the reference), second-order
    outlet, user ghost layers at j- (Ghost_Cell(3) = -1), periodic in z.
    Grid: GRID2D_PLANE, wall-clustered in y."""
    xs = x0 + lx * np.arange(nx) / (nx - 1)
    s = np.arange(ny) / (ny - 1)
    eta = (np.exp(2.0 * s) - 1.0) / (np.exp(2.0) - 1.0)          # wall clustering
    top = y_top + rise * np.clip((xs - x0 - 0.5 * lx) / (0.5 * lx), 0.0, None)
    x2 = np.broadcast_to(xs[None, :], (ny, nx)).copy()
    y2 = eta[:, None] * top[None, :]
    z1 = zspan * np.arange(nz) / nz
    shape = (nz, ny, nx)
    X = np.broadcast_to(x2, shape).copy()
    Y = np.broadcast_to(y2, shape).copy()
    Z = np.broadcast_to(z1[:, None, None], shape).copy()
    d, u, T = _wall_layer(Y, delta, t_wall)
    v = np.zeros(shape)
    w = np.zeros(shape)
    inlet = np.stack([y2[:, 0], d[0, :, 0], u[0, :, 0], v[0, :, 0], T[0, :, 0]], axis=1)
    bc = {1: 1, 2: 0, 3: 0, 4: 2, 5: t_wall, 6: 0.0, 7: 1, 8: 0.02, 9: x0 + 0.2 * lx, 10: x0 + 0.7 * lx,
          11: 3, 12: 2, 13: zspan, 14: 40.0}
    cfg = CaseConfig(nx=nx, ny=ny, nz=nz, iperiodic=(0, 0, 1), scheme_invis=scheme, scheme_vis="CD6",
                     flux_splitting=split, if_scheme_character=char, ma=ma, re=re,
                     grid_type=GRID2D_PLANE, boundary_condition="BC_BOUNDARYLAYER", bc_para=bc,
                     scheme_boundary=(0, 0, 1, 0, 0, 0), ghost_cell=(0, 0, -1, 0, 0, 0),
                     periodic_span=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, zspan))
    case = Case(cfg=cfg, x=X, y=Y, z=Z, spans=np.array(cfg.periodic_span), prims=[d, u, v, w, T],
                grid_payload=(x2, y2, z1), name="flatplate%dx%dx%d" % (nx, ny, nz),
                info={"inlet1d": inlet, "t_wall": t_wall})
    cfg.dt = cfl_dt(case, cfl)
    return case


def swept_corner(nx=24, ny=18, nz=12, x0=0.0, lx=1.0, y_top=0.3, ramp=0.15, zspan=0.3, wall_zl=0.18,
                 ma=2.5, re=1.0e4, t_wall=1.4, delta=0.03, cfl=0.4, scheme="WENO5", split=1, char=1) -> Case:
    """Swept compression ramp with BoundaryCondition = BC_SweptCorner
    (ocfd_bc_swept_corner, OCFD_BC_Boundarylayers.f90:443-617): 1-d inlet
    profile file, free-stream upper boundary, a wall that is no-slip for
    z <= wall_zl or where it is flat and extrapolated elsewhere, random
    wall blowing (type 2, drawn by the reference), second-order outlet,
    first-order extrapolation at both spanwise sides, symmetry ghost plane at
    k- (Ghost_Cell(5) = -1).  Grid: GRID3D, the ramp starting mid-plate."""
    xs = x0 + lx * np.arange(nx) / (nx - 1)
    s = np.arange(ny) / (ny - 1)
    eta = (np.exp(2.0 * s) - 1.0) / (np.exp(2.0) - 1.0)
    zs = zspan * np.arange(nz) / (nz - 1)
    yw = ramp * np.clip((xs - x0 - 0.5 * lx) / (0.5 * lx), 0.0, None)
    shape = (nz, ny, nx)
    X = np.broadcast_to(xs[None, None, :], shape).copy()
    Y = np.broadcast_to(yw[None, None, :] + eta[None, :, None] * (y_top - yw[None, None, :]), shape).copy()
    Z = np.broadcast_to(zs[:, None, None], shape).copy()
    d, u, T = _wall_layer(Y - yw[None, None, :], delta, t_wall)
    v = np.zeros(shape)
    w = np.zeros(shape)
    inlet = np.stack([Y[0, :, 0], d[0, :, 0], u[0, :, 0], v[0, :, 0], T[0, :, 0]], axis=1)
    bc = {1: 2, 2: t_wall, 3: wall_zl, 4: 2, 5: 0.01, 6: x0 + 0.1 * lx, 7: x0 + 0.6 * lx, 8: 2, 9: 0, 10: 30.0,
          11: 1, 12: 1}
    cfg = CaseConfig(nx=nx, ny=ny, nz=nz, iperiodic=(0, 0, 0), scheme_invis=scheme, scheme_vis="CD6",
                     flux_splitting=split, if_scheme_character=char, ma=ma, re=re,
                     grid_type=GRID3D, boundary_condition="BC_SWEPTCORNER", bc_para=bc,
                     scheme_boundary=(0, 0, 0, 0, 1, 0), ghost_cell=(0, 0, 0, 0, -1, 0))
    case = Case(cfg=cfg, x=X, y=Y, z=Z, spans=np.zeros(9), prims=[d, u, v, w, T],
                name="swept%dx%dx%d" % (nx, ny, nz), info={"inlet1d": inlet, "t_wall": t_wall})
    cfg.dt = cfl_dt(case, cfl)
    return case


# ---------------------------------------------------------------------------
def cfl_dt(case: Case, cfl: float) -> float:
    """dt = cfl / max_cells sum_axes (|u.grad xi| + c |grad xi|) / h_xi, with the
    grid derivatives by second-order differences (only the time-step size
    depends on it; both codes use the same fixed dt)."""
    cfg = case.cfg
    d, u, v, w, T = case.prims
    c = np.sqrt(T) / cfg.ma
    X, Y, Z = case.x, case.y, case.z
    per = cfg.iperiodic
    spans = np.asarray(case.spans).reshape(3, 3)

    def deriv(F, axis, comp):
        # axis in the (z, y, x) array order: 2 -> i, 1 -> j, 0 -> k
        a = {0: 2, 1: 1, 2: 0}[axis]
        n = F.shape[a]
        if per[axis]:
            fp = np.roll(F, -1, axis=a)
            fm = np.roll(F, 1, axis=a)
            idx = [slice(None)] * 3
            idx[a] = slice(n - 1, n)
            fp[tuple(idx)] += spans[axis][comp]
            idx[a] = slice(0, 1)
            fm[tuple(idx)] -= spans[axis][comp]
            return (fp - fm) / 2.0 * (n - 1)
        return np.gradient(F, axis=a) * (n - 1)

    hs = [1.0 / (cfg.nx - 1), 1.0 / (cfg.ny - 1), 1.0 / (cfg.nz - 1)]
    if _separable(X, Y, Z):
        # x = x(i), y = y(j), z = z(k): the metric tensor is diagonal
        gx = 1.0 / deriv(X[:1, :1, :], 0, 0)
        gy = 1.0 / deriv(Y[:1, :, :1], 1, 1)
        gz = 1.0 / deriv(Z[:, :1, :1], 2, 2)
        rate = ((np.abs(u * gx) + c * np.abs(gx)) / hs[0] + (np.abs(v * gy) + c * np.abs(gy)) / hs[1]
                + (np.abs(w * gz) + c * np.abs(gz)) / hs[2])
        return float(cfl / rate.max())
    J = np.empty(X.shape + (3, 3))
    for comp, F in enumerate((X, Y, Z)):
        for axis in range(3):
            J[..., comp, axis] = deriv(F, axis, comp)
    G = np.linalg.inv(J)          # G[..., axis, comp] = d xi_axis / d x_comp
    rate = np.zeros(X.shape)
    for axis in range(3):
        gx, gy, gz = G[..., axis, 0], G[..., axis, 1], G[..., axis, 2]
        un = np.abs(u * gx + v * gy + w * gz)
        gn = np.sqrt(gx * gx + gy * gy + gz * gz)
        rate += (un + c * gn) / hs[axis]
    return float(cfl / rate.max())


def _separable(X, Y, Z):
    if all(a.base is not None and 0 in a.strides for a in (X, Y, Z)) and \
            X.strides[:2] == (0, 0) and Y.strides[::2] == (0, 0) and Z.strides[1:] == (0, 0):
        return True
    return (np.array_equal(X, np.broadcast_to(X[:1, :1, :], X.shape))
            and np.array_equal(Y, np.broadcast_to(Y[:1, :, :1], Y.shape))
            and np.array_equal(Z, np.broadcast_to(Z[:, :1, :1], Z.shape)))


# ---------------------------------------------------------------------------
# batched ensemble of independent periodic boxes
# ---------------------------------------------------------------------------
BATCH_SIZES = (32, 48, 64, 96, 128)


def batch_ensemble(nbox: int = 96, precision: int = 0, seed: int = 5):
    """nbox boxes: n^3 with n uniform over BATCH_SIZES, a Taylor-Green or a
    seeded isotropic initial field (seed 20261202 + box), Ma uniform in
    [0.1, 0.5], Re uniform in [400, 3200], WENO5 or OMP6 with Steger-Warming
    splitting (component-wise), CD6, CFL 0.5, uniform-box metrics."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nbox):
        n = int(rng.choice(BATCH_SIZES))
        kind = "tgv" if rng.random() < 0.5 else "iso"
        ma = float(rng.uniform(0.1, 0.5))
        re = float(rng.uniform(400.0, 3200.0))
        sch = "WENO5" if rng.random() < 0.5 else "OMP6"
        if kind == "tgv":
            c = tgv(n, ma=ma, re=re, scheme=sch, split=1, metric_mode=METRIC_UNIFORM)
        else:
            c = isotropic(n, seed=20261202 + i, mt=ma, re=re, scheme=sch, split=1, char=0,
                          metric_mode=METRIC_UNIFORM)
        c.cfg = c.cfg.replace(precision=precision)
        c.name = "box%d_%s%d_%s" % (i, kind, n, sch)
        out.append(c)
    return out
