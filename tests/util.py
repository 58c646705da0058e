"""Helpers shared by the tests: golden-fixture loading and error norms."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from ocfd import cases
from ocfd.config import CaseConfig

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    return sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(GOLDEN, "ref_*.npz")))


def load_golden(name):
    """Return (Case, dict of reference outputs) for tests/golden/ref_<name>.npz."""
    z = np.load(os.path.join(GOLDEN, "ref_%s.npz" % name), allow_pickle=False)
    d = json.loads(str(z["config"]))
    d["bc_para"] = {int(k): v for k, v in d["bc_para"].items()}
    d["hybrid_para"] = {int(k): v for k, v in d["hybrid_para"].items()}
    for key in ("iperiodic", "scheme_boundary", "ghost_cell", "mass_force", "periodic_span"):
        d[key] = tuple(d[key])
    cfg = CaseConfig(**d)
    prims = [np.ascontiguousarray(a) for a in z["prims"]]
    case = cases.Case(cfg=cfg, x=z["x"], y=z["y"], z=z["z"], spans=z["spans"], prims=prims, name=name)
    ref = {k: z[k] for k in z.files if k not in ("config", "x", "y", "z", "spans", "prims")}
    return case, ref


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def interior(a, lap=4):
    return a[lap:-lap, lap:-lap, lap:-lap]


def gpu_available():
    try:
        import ctypes
        lib = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return lib.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False
