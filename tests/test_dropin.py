"""The reference solver with its hot-path routines served by the GPU library
(opencfd-sc_amd/fortran: the reference's own objects, NS_solver unchanged,
routine calls redirected through the iso_c_binding layer) against the
unmodified reference executable on the same input files: every flow-field
file, the ana_box statistics and the show_flow_msg sums must be identical."""
import glob
import os
import re
import subprocess

import pytest

from util import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "opencfd-sc_amd", "fortran", "_build", "opencfd_hip")
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


def _run(exe, workdir, case, nsteps):
    from pyoracle import MPIEXEC
    os.makedirs(workdir, exist_ok=True)
    case.cfg.istep_show = 1
    case.write_reference_inputs(workdir, nsteps=nsteps, ana_box_every=1, istep_save=1)
    r = subprocess.run([MPIEXEC, "-n", "1", exe], cwd=workdir, capture_output=True, text=True, timeout=300)
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
    out_gpu = _run(DROPIN, str(tmp_path / "gpu"), case, nsteps)
    assert "libocfd_hip: block on GPU" in out_gpu
    files = sorted(os.path.basename(p) for p in glob.glob(str(tmp_path / "ref" / "OCFD[0-9]*.dat")))
    assert len(files) == nsteps  # one per step; the final save rewrites the last
    for fn in files + ["statistics.dat"]:
        a = open(tmp_path / "ref" / fn, "rb").read()
        b = open(tmp_path / "gpu" / fn, "rb").read()
        assert a == b, fn
    assert _sums(out_ref) == _sums(out_gpu)
    assert len(_sums(out_ref)) >= nsteps


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
