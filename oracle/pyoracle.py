"""ctypes driver for the C oracle (liboracle.so) and the compiled reference
(_ref/ocfd_probe, _ref/opencfd_ref).  TEST INFRASTRUCTURE ONLY: used by tests/,
the smoke check and the CPU-baseline leg of bench.py, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "opencfd-sc_amd"))

from ocfd.config import CaseConfig, OcfdConfigC  # noqa: E402
from ocfd.native import BlockBase, OcfdError, _ptr  # noqa: E402

LIB_PATH = os.path.join(HERE, "liboracle.so")
COUNT_LIB_PATH = os.path.join(HERE, "liboracle_count.so")
REF_DIR = os.path.join(HERE, "_ref")
PROBE = os.path.join(REF_DIR, "ocfd_probe")
REF_EXE = os.path.join(REF_DIR, "opencfd_ref")
# the reference with its empty ocfd_bc_User replaced by ref_user_hooks.f90
REF_WALL_EXE = os.path.join(REF_DIR, "opencfd_ref_wall")
MPIEXEC = "/opt/conda/bin/mpiexec"

_lib = None
_count_lib = None


def load_counting():
    """The operation-counting build of the oracle (oracle_count.cpp)."""
    global _count_lib
    if _count_lib is None:
        if not os.path.exists(COUNT_LIB_PATH):
            raise OcfdError("counting oracle not built: %s" % COUNT_LIB_PATH)
        lib = ctypes.CDLL(COUNT_LIB_PATH)
        lib.oracle_create.argtypes = [ctypes.POINTER(OcfdConfigC), ctypes.POINTER(ctypes.c_void_p)]
        _count_lib = lib
    return _count_lib


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OcfdError("oracle not built: %s" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_create.argtypes = [ctypes.POINTER(OcfdConfigC), ctypes.POINTER(ctypes.c_void_p)]
        for nm in ("hh_weno5p", "hh_weno5m", "hh_weno7p", "hh_weno7m", "hh_omp6p", "hh_omp6m"):
            f = getattr(lib, "oracle_" + nm)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.POINTER(ctypes.c_double)]
        for nm in ("hh_weno5p_boundary", "hh_weno5m_boundary"):
            f = getattr(lib, "oracle_" + nm)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        _lib = lib
    return _lib


class OracleBlock(BlockBase):
    prefix = "oracle"

    def __init__(self, cfg: CaseConfig, counting: bool = False):
        super().__init__(load_counting() if counting else load(), cfg)
        h = ctypes.c_void_p()
        if self.lib.oracle_create(ctypes.byref(self._c), ctypes.byref(h)) != 0:
            raise OcfdError("oracle_create failed")
        self.h = h

    def set_grid_and_jacobian(self, x, y, z, spans, m10):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        sp = np.ascontiguousarray(spans, dtype=np.float64)
        fields = [np.ascontiguousarray(np.broadcast_to(np.float64(v), x.shape)) for v in m10]
        arr = (ctypes.POINTER(ctypes.c_double) * 10)(*[_ptr(m) for m in fields])
        return self._call("set_grid_and_jacobian", _ptr(x), _ptr(y), _ptr(z), _ptr(sp), arr)

    def set_uniform_metrics(self, m10):
        raise OcfdError("oracle: use set_grid_and_jacobian")

    def dxyz0(self, axis, fh):
        fh = np.ascontiguousarray(fh, dtype=np.float64)
        out = np.empty((self.cfg.nz, self.cfg.ny, self.cfg.nx))
        self.lib.oracle_dxyz0(self.h, axis, _ptr(fh), _ptr(out))
        return out

    def close(self):
        if self.h is not None:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def setup_oracle(case, uniform=None, counting=False):
    """Oracle block in the state the reference reaches after init3d + OCFD_bc."""
    cfg = case.cfg
    if uniform is None:
        uniform = cfg.metric_mode == 1
    b = OracleBlock(cfg, counting=counting)
    if uniform:
        b.set_grid_and_jacobian(case.x, case.y, case.z, case.spans, case.uniform_metrics)
    else:
        b.set_grid(case.x, case.y, case.z, case.spans)
    b.upload_primitives(case.halo_prims(), 0, 0.0)
    b.bc()
    return b


def pointwise(name, v, center, *extra):
    """Evaluate an oracle point-wise reconstruction on array v with v(0) at
    index `center`."""
    lib = load()
    v = np.ascontiguousarray(v, dtype=np.float64)
    p = ctypes.cast(v.ctypes.data + 8 * center, ctypes.POINTER(ctypes.c_double))
    return getattr(lib, "oracle_" + name)(p, *extra)


# ---------------------------------------------------------------------------
# the compiled reference
# ---------------------------------------------------------------------------
def reference_available():
    return os.path.exists(PROBE) and os.path.exists(REF_EXE) and os.path.exists(MPIEXEC)


def run_probe(case, workdir, nextra=0, timeout=600, allow_negative_t=False):
    """Run the probe driver on `case` in `workdir`; returns dict of arrays.
    With allow_negative_t the cells the reference reports as "Negative T
    found" (handle_NegativeT, OCFD_handle_NegativeT.f90:9, 1-based global
    i, j, k in the order repaired) come back as res["negt_cells"]; a run the
    reference stops (more than 10 in one stage) is an error either way."""
    os.makedirs(workdir, exist_ok=True)
    case.write_reference_inputs(workdir, nsteps=1 + nextra)
    out = os.path.join(workdir, "probe_out")
    os.makedirs(out, exist_ok=True)
    r = subprocess.run([MPIEXEC, "-n", "1", PROBE, out, str(nextra)], cwd=workdir,
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0 or "STOP" in r.stdout or "Limit" in r.stdout or \
            ("Negative T" in r.stdout and not allow_negative_t):
        raise OcfdError("probe failed: %s\n%s" % (r.stdout[-2000:], r.stderr[-2000:]))
    cfg = case.cfg
    L = cfg.lap
    hs = (cfg.nz + 2 * L, cfg.ny + 2 * L, cfg.nx + 2 * L)
    fs = (5, cfg.nz, cfg.ny, cfg.nx)
    res = {}
    for fn in os.listdir(out):
        if not fn.endswith(".bin"):
            continue
        a = np.fromfile(os.path.join(out, fn), dtype="<f8")
        name = fn[:-4]
        if a.size == np.prod(hs):
            res[name] = a.reshape(hs)
        elif a.size == np.prod(fs):
            res[name] = a.reshape(fs)
        else:
            res[name] = a
    hist = os.path.join(out, "hist.txt")
    if os.path.exists(hist):
        res["hist"] = np.loadtxt(hist, ndmin=2)
    res["stdout"] = r.stdout
    if allow_negative_t:
        cells = [[int(x) for x in line.split("!")[1].split()] for line in r.stdout.splitlines()
                 if "Negative T found" in line]
        res["negt_cells"] = np.array(cells, dtype=np.int32).reshape(-1, 3)
    return res


def run_reference(case, workdir, nsteps, nranks=1, decomposition=(1, 1, 1), timeout=3600,
                  ana_box_every=0, exe=None):
    """Run the reference solver (exe: REF_EXE, the unmodified one, by default;
    REF_WALL_EXE for BC_User_Def cases that need the wall hook); returns (wall
    seconds reported by the solver, final primitives, stdout)."""
    from ocfd import fortran_io
    os.makedirs(workdir, exist_ok=True)
    case.write_reference_inputs(workdir, nsteps=nsteps, ana_box_every=ana_box_every)
    if decomposition != (1, 1, 1):
        nl = os.path.join(workdir, "opencfd2.in")
        s = open(nl).read()
        for key, val in zip(("npx0", "npy0", "npz0"), decomposition):
            s = s.replace(" %s=1\n" % key, " %s=%d\n" % (key, val))
        open(nl, "w").write(s)
    r = subprocess.run([MPIEXEC, "-n", str(nranks), exe or REF_EXE], cwd=workdir, capture_output=True,
                       text=True, timeout=timeout)
    if r.returncode != 0:
        raise OcfdError("reference failed: %s\n%s" % (r.stdout[-2000:], r.stderr[-2000:]))
    wall = None
    for line in r.stdout.splitlines():
        if "total wtime cost" in line:
            wall = float(line.split(":")[-1])
    fn = os.path.join(workdir, "OCFD%08d.dat" % nsteps)
    prims = None
    if os.path.exists(fn):
        _, _, prims = fortran_io.read_flow_data(fn, case.cfg.shape)
    return wall, prims, r.stdout


def read_statistics(path):
    """statistics.dat as the reference's ana_box appends it (OCFD_ANA.f90:305-308):
    one row of 20 values per call (E20.13 each), in the column order of
    ocfd_ana_box's output."""
    rows = []
    for line in open(path):
        line = line.rstrip("\n")
        if not line.strip():
            continue
        rows.append([float(line[20 * q:20 * (q + 1)].replace("D", "E")) for q in range(20)])
    return np.array(rows)


def count_ops(block, fn):
    """Floating-point operations (add/sub, mul, div, sqrt, pow) executed by
    fn(block) on a counting oracle block."""
    lib = block.lib
    lib.oracle_count_reset()
    fn(block)
    out = (ctypes.c_int64 * 5)()
    lib.oracle_count_read(out)
    return list(out)
