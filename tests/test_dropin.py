"""The reference solver with its hot-path routines served by the GPU library
(opencfd-sc_amd/fortran: the reference's own objects, NS_solver unchanged
but for its host fn=f loop, routine calls redirected through the
iso_c_binding layer) against the unmodified reference executable on the same
input files: every flow-field file, the ana_box statistics and the
show_flow_msg sums must be identical.  The stage routines run in stage
fusion (the default; OCFD_HIP_STAGE_FUSION=0 is run too).  show_flow_msg and
ana_box run on the host after a download (OCFD_HIP_DIAG=host) where their
printouts are compared byte for byte; from the device sums (the default)
they are compared as numbers."""
import glob
import os
import re
import subprocess

import pytest

from util import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "opencfd-sc_amd", "fortran", "_build", "opencfd_hip")
# the same with the wall hook of oracle/ref_user_hooks.f90 as the user's ocfd_bc_User
DROPIN_WALL = os.path.join(ROOT, "opencfd-sc_amd", "fortran", "_build", "opencfd_hip_wall")
FC_MOD = os.path.join(ROOT, "opencfd-sc_amd", "fortran", "ocfd_hip_mod.f90")

# (case name, steps): periodic hybrid, curvilinear closures, filters,
# blunt-body BC with mass force, pressure-gradient sensor (the golden cases),
# then the boundary-layer and swept-corner conditions with their inlet files,
# time-dependent random wall blowing and user ghost layers
CASES = [("hyb_c_llf", 3), ("curv_w5c", 3), ("grid1d_shockfilter", 4), ("blunt_massforce", 3),
         ("hyb_sensor2_curv", 3), ("flat_plate", 4), ("swept_corner", 4)]


def _case(name):
    """The seeded case behind golden vector `name` (same generator), or a
    generated boundary-layer case."""
    import sys
    from ocfd import cases
    if name == "flat_plate":
        return cases.flat_plate()
    if name == "swept_corner":
        return cases.swept_corner()
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import golden_cases
    return golden_cases()[name]


HOST_DIAG = {"OCFD_HIP_DIAG": "host"}


def _run(exe, workdir, case, nsteps, env=None):
    from pyoracle import MPIEXEC
    os.makedirs(workdir, exist_ok=True)
    case.cfg.istep_show = 1
    case.write_reference_inputs(workdir, nsteps=nsteps, ana_box_every=1, istep_save=1)
    r = subprocess.run([MPIEXEC, "-n", "1", exe], cwd=workdir, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, **(env or {})))
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "STOP" not in out and "failed" not in out, out[-3000:]
    return out


def _sums(stdout):
    """The E, K, S lines printed by show_flow_msg and the hybrid shares."""
    lines = stdout.splitlines()
    keep = []
    for i, l in enumerate(lines):
        if "Averaged Total-energy" in l and i + 1 < len(lines):
            keep.append(lines[i + 1].strip())
        if "the percent of 3 schemes" in l:
            keep.append(re.sub(r"\s+", " ", l.strip()))
    return keep


@pytest.mark.gpu
@pytest.mark.parametrize("name,nsteps", CASES)
def test_dropin_matches_reference_files(tmp_path, name, nsteps):
    import pyoracle
    if not gpu_available():
        pytest.skip("no GPU")
    if not os.path.exists(DROPIN):
        pytest.skip("drop-in executable not built (make -C opencfd-sc_amd/fortran)")
    if not os.path.exists(pyoracle.REF_EXE):
        pytest.skip("reference executable not built (make -C oracle ref)")
    case = _case(name)
    print("drop-in case", name, case.cfg.shape, flush=True)
    out_ref = _run(pyoracle.REF_EXE, str(tmp_path / "ref"), case, nsteps)
    out_gpu = _run(DROPIN, str(tmp_path / "gpu"), case, nsteps, env=HOST_DIAG)
    assert "libocfd_hip: block on GPU" in out_gpu
    files = sorted(os.path.basename(p) for p in glob.glob(str(tmp_path / "ref" / "OCFD[0-9]*.dat")))
    assert len(files) == nsteps  # one per step; the final save rewrites the last
    for fn in files + ["statistics.dat"]:
        a = open(tmp_path / "ref" / fn, "rb").read()
        b = open(tmp_path / "gpu" / fn, "rb").read()
        assert a == b, fn
    assert _sums(out_ref) == _sums(out_gpu)
    assert len(_sums(out_ref)) >= nsteps


def _compare_files(tmp_path, ref, gpu, nsteps):
    files = sorted(os.path.basename(p) for p in glob.glob(str(tmp_path / ref / "OCFD[0-9]*.dat")))
    assert len(files) == nsteps
    for fn in files + ["statistics.dat"]:
        a = open(tmp_path / ref / fn, "rb").read()
        b = open(tmp_path / gpu / fn, "rb").read()
        assert a == b, fn


def _small_blunt():
    from ocfd import cases
    return cases.blunt_ogrid(nx=48, ny=24, nz=8)


# BoundaryCondition = BC_User_Def: the user's ocfd_bc_User runs on the host by
# default (OCFD_BC.f90:18-19).  With the reference's own, empty, hook the
# drop-in must reproduce the reference (no wall appears); with the wall hook
# of ref_user_hooks.f90 linked into both executables, the host pass-through
# and the library's device hook (OCFD_HIP_USER_BC=device) must both
# reproduce the reference built with that hook.
USER_BC = [("curv_sb_ghost", "stub", "host"), ("curv_wall_fo9p", "wall", "host"),
           ("curv_wall_fo9p", "wall", "device"), ("blunt_small", "wall", "host"),
           ("blunt_small", "wall", "device")]


@pytest.mark.gpu
@pytest.mark.parametrize("name,hook,mode", USER_BC)
def test_dropin_user_bc(tmp_path, name, hook, mode):
    import pyoracle
    if not gpu_available():
        pytest.skip("no GPU")
    exe, ref = (DROPIN, pyoracle.REF_EXE) if hook == "stub" else (DROPIN_WALL, pyoracle.REF_WALL_EXE)
    for p in (exe, ref):
        if not os.path.exists(p):
            pytest.skip("%s not built" % p)
    case = _small_blunt() if name == "blunt_small" else _case(name)
    assert case.cfg.boundary_condition.upper() == "BC_USER_DEF"
    print("drop-in user BC", name, hook, mode, case.cfg.shape, flush=True)
    nsteps = 3
    out_ref = _run(ref, str(tmp_path / "ref"), case, nsteps)
    out_gpu = _run(exe, str(tmp_path / "gpu"), case, nsteps, env=dict(HOST_DIAG, OCFD_HIP_USER_BC=mode))
    want = "device wall hook" if mode == "device" else "host routine ocfd_bc_User"
    assert want in out_gpu
    _compare_files(tmp_path, "ref", "gpu", nsteps)
    assert _sums(out_ref) == _sums(out_gpu)


def _stats(path):
    import numpy as np
    import pyoracle
    return np.asarray(pyoracle.read_statistics(path))


def _sum_values(stdout):
    """E, K, S of every show_flow_msg printout, as numbers."""
    return [[float(x) for x in l.split()] for l in _sums(stdout) if "percent" not in l]


@pytest.mark.gpu
@pytest.mark.parametrize("name,nsteps", [("hyb_c_llf", 3), ("flat_plate", 4)])
def test_dropin_device_diagnostics_and_per_routine_kernels(tmp_path, name, nsteps):
    """Defaults (stage fusion, show_flow_msg / ana_box from the device sums)
    and the per-routine kernel sequence (OCFD_HIP_STAGE_FUSION=0) against the
    reference: flow files byte for byte in both; the device diagnostics
    within 1e-12 of the reference's printed E, K, S and 1e-10 of its
    statistics.dat (both sum in another order; 1e-6 for the odd moments)."""
    import numpy as np
    import pyoracle
    if not gpu_available():
        pytest.skip("no GPU")
    if not os.path.exists(DROPIN) or not os.path.exists(pyoracle.REF_EXE):
        pytest.skip("drop-in or reference executable not built")
    case = _case(name)
    print("drop-in device diagnostics", name, case.cfg.shape, flush=True)
    out_ref = _run(pyoracle.REF_EXE, str(tmp_path / "ref"), case, nsteps)
    out_dev = _run(DROPIN, str(tmp_path / "gpu"), case, nsteps)
    _compare_files_only(tmp_path, "ref", "gpu", nsteps)
    e_ref, e_dev = _sum_values(out_ref), _sum_values(out_dev)
    assert len(e_ref) == len(e_dev) >= nsteps
    np.testing.assert_allclose(e_dev, e_ref, rtol=1e-12, atol=0)
    s_ref, s_dev = _stats(str(tmp_path / "ref" / "statistics.dat")), _stats(str(tmp_path / "gpu" / "statistics.dat"))
    assert s_ref.shape == s_dev.shape
    # odd moments (u and ux skewness, columns 9 and 11) are sums of terms of
    # both signs: their relative rounding is not bounded by the sum's size
    rtol = np.full(20, 1e-10)
    rtol[[8, 10]] = 1e-6
    rel = np.abs(s_dev - s_ref) / np.abs(s_ref)
    assert np.all(rel <= rtol), rel.max(axis=0)
    out_pr = _run(DROPIN, str(tmp_path / "routines"), case, nsteps,
                  env=dict(HOST_DIAG, OCFD_HIP_STAGE_FUSION="0"))
    _compare_files(tmp_path, "ref", "routines", nsteps)
    assert _sums(out_ref) == _sums(out_pr)


def _compare_files_only(tmp_path, ref, gpu, nsteps):
    files = sorted(os.path.basename(p) for p in glob.glob(str(tmp_path / ref / "OCFD[0-9]*.dat")))
    assert len(files) == nsteps
    for fn in files:
        assert open(tmp_path / ref / fn, "rb").read() == open(tmp_path / gpu / fn, "rb").read(), fn


def test_fortran_binding_config_layout(tmp_path):
    """The bind(C) mirror of ocfd_config in the Fortran module has the C
    struct's size and member offsets."""
    import ctypes
    import shutil
    from ocfd.config import OcfdConfigC
    fc = shutil.which("amdflang")
    if fc is None:
        pytest.skip("amdflang not available")
    prog = tmp_path / "layout.f90"
    prog.write_text("""program layout
   use ocfd_hip
   implicit none
   type(ocfd_config), target :: c
   integer(c_intptr_t) :: b
   b = transfer(c_loc(c), b)
   print *, c_sizeof(c)
   print *, transfer(c_loc(c%iperiodic), b) - b, transfer(c_loc(c%ma), b) - b
   print *, transfer(c_loc(c%bc_para), b) - b, transfer(c_loc(c%filter_para), b) - b
   print *, transfer(c_loc(c%reserved_d), b) - b
end program layout
""")
    libdir = os.path.join(ROOT, "opencfd-sc_amd")
    r = subprocess.run([fc, "-J", str(tmp_path), "-o", str(tmp_path / "layout"), FC_MOD, str(prog),
                        "-L" + libdir, "-locfd_hip", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    vals = [int(v) for v in subprocess.run([str(tmp_path / "layout")], capture_output=True,
                                           text=True).stdout.split()]
    C = OcfdConfigC
    assert vals == [ctypes.sizeof(C), C.iperiodic.offset, C.ma.offset, C.bc_para.offset,
                    C.filter_para.offset, C.reserved_d.offset]


def test_dropin_timer_parsing():
    """The bench's drop-in leg reads the reference's own timers from the run's
    output (list-directed prints of show_flow_msg and NS_solver)."""
    from ocfd import dropin
    out = (" Istep=          20 tt=  1.1440000000000001E-002\n"
           " CPU time per          20 step is:      0.31456789000000001\n"
           " Istep=          40 tt=  2.2880000000000001E-002\n"
           " CPU time per          20 step is:      0.28571234500000000\n"
           " total wtime cost:  0.60245000000000004\n")
    iv, total = dropin.parse_timers(out)
    assert iv == [0.31456789000000001, 0.285712345] and total == 0.60245000000000004
    assert dropin.parse_timers("no timers") == ([], None)
