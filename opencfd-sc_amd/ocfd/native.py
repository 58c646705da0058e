"""ctypes binding of the C ABI (include/ocfd_hip.h).

:class:`Block` mirrors the reference's per-routine interface: one method per
reference routine (du_inviscous, du_viscous, OCFD_time_adv_RK3, ...) plus
whole-step and transfer helpers.  Arrays are numpy float64 in the reference's
Fortran memory order: halo'd fields have shape (nz+2L, ny+2L, nx+2L) and the
conserved-variable arrays f, fn, du have shape (5, nz, ny, nx).

The library is loaded from the package directory (built in-tree by
``build.py``).  Nothing here falls back to another implementation: a missing
or unloadable library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import CaseConfig, OcfdConfigC

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libocfd_hip.so")

_DP = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected C-contiguous float64 array")
    return a.ctypes.data_as(_DP)


class OcfdError(RuntimeError):
    pass


class BlockBase:
    """Common driver for any library exposing the ocfd_* call set under a
    prefix (the HIP library uses ``ocfd``)."""

    prefix = "ocfd"

    def __init__(self, lib, cfg: CaseConfig):
        self.lib = lib
        self.cfg = cfg
        self._c = cfg.to_c()
        self.h = None
        L = cfg.lap
        self.halo_shape = (cfg.nz + 2 * L, cfg.ny + 2 * L, cfg.nx + 2 * L)
        self.f_shape = (5, cfg.nz, cfg.ny, cfg.nx)

    # -- plumbing ----------------------------------------------------------
    def _fn(self, name):
        return getattr(self.lib, "%s_%s" % (self.prefix, name))

    def _err(self):
        return "error"

    def _call(self, name, *args, allow_positive=True):
        rc = self._fn(name)(self.h, *args)
        if rc < 0 or (rc > 0 and not allow_positive):
            raise OcfdError("%s_%s failed (%d): %s" % (self.prefix, name, rc, self._err()))
        return rc

    # -- grid ----------------------------------------------------------------
    def set_grid(self, x, y, z, spans):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        sp = np.ascontiguousarray(spans, dtype=np.float64)
        return self._call("set_grid", _ptr(x), _ptr(y), _ptr(z), _ptr(sp))

    def upload_metrics(self, mets):
        mets = [np.ascontiguousarray(m, dtype=np.float64) for m in mets]
        arr = (_DP * 19)(*[_ptr(m) for m in mets])
        return self._call("upload_metrics", arr)

    def download_metrics(self):
        out = [np.empty(self.halo_shape) for _ in range(19)]
        arr = (_DP * 19)(*[_ptr(m) for m in out])
        self._call("download_metrics", arr)
        return out

    # -- state ---------------------------------------------------------------
    def upload_primitives(self, prims, istep=0, tt=0.0):
        ps = [np.ascontiguousarray(p, dtype=np.float64) for p in prims]
        for p in ps:
            if p.shape != self.halo_shape:
                raise ValueError("primitive arrays must be halo'd %s" % (self.halo_shape,))
        return self._call("upload_primitives", *[_ptr(p) for p in ps], ctypes.c_int32(istep),
                          ctypes.c_double(tt))

    def _down5(self, name):
        a = np.empty(self.f_shape)
        self._call(name, _ptr(a))
        return a

    def download_f(self):
        return self._down5("download_f")

    def download_fn(self):
        return self._down5("download_fn")

    def download_du(self):
        return self._down5("download_du")

    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.f_shape)
        return self._call("upload_f", _ptr(f))

    def upload_du(self, du):
        du = np.ascontiguousarray(du, dtype=np.float64).reshape(self.f_shape)
        return self._call("upload_du", _ptr(du))

    def download_primitives(self):
        out = [np.empty(self.halo_shape) for _ in range(5)]
        self._call("download_primitives", *[_ptr(a) for a in out])
        return out

    def download_rhybrid(self):
        a = np.empty(self.halo_shape)
        self._call("download_rhybrid", _ptr(a))
        return a

    # -- restart / snapshot files (save_flow_data / read_flow_data) ---------------
    def save_flow_data(self, path=None, with_f=False):
        p = path.encode() if path is not None else None
        return self._call("save_flow_data", ctypes.c_char_p(p), ctypes.c_int32(1 if with_f else 0))

    def read_flow_data(self, path=None):
        """Returns True when the file carried f (bit-exact restart)."""
        has_f = ctypes.c_int32()
        p = path.encode() if path is not None else None
        self._call("read_flow_data", ctypes.c_char_p(p), ctypes.byref(has_f))
        return bool(has_f.value)

    def get_time(self):
        i = ctypes.c_int32()
        t = ctypes.c_double()
        self._call("get_time", ctypes.byref(i), ctypes.byref(t))
        return i.value, t.value

    # -- one reference routine each -------------------------------------------
    def exchange_primitives(self):
        return self._call("exchange_primitives")

    def comput_rhybrid(self, counts=False):
        if counts:
            c = (ctypes.c_int64 * 4)()
            self._call("comput_rhybrid", c)
            return list(c)
        self._call("comput_rhybrid", None)
        return None

    def du_inviscous(self):
        return self._call("du_inviscous")

    def du_inviscous_character(self):
        return self._call("du_inviscous_character")

    def du_viscous(self):
        return self._call("du_viscous")

    def time_adv_rk3(self, krk):
        return self._call("time_adv_rk3", ctypes.c_int32(krk))

    def save_fn(self):
        return self._call("save_fn")

    def comput_duvwt(self):
        return self._call("comput_duvwt")

    def bc(self):
        return self._call("bc")

    def set_stage_fusion(self, on=True):
        """ocfd_set_stage_fusion: the per-routine calls of one stage, issued
        in NS_solver's order, run the whole-step stage kernels."""
        return self._call("set_stage_fusion", ctypes.c_int32(1 if on else 0))

    def set_bc_boundarylayer(self, aoa=0.0, inlet=None, upper=None, nonref=None, pertb=None, wall=None):
        """ocfd_set_bc_boundarylayer: the data of the boundary-layer /
        swept-corner conditions -- inlet (5, nz, ny) and upper (5, nz, nx)
        primitives, nonref and wall (nz, nx) flags, pertb (nz, nx); None
        leaves an item out (free stream, all Dirichlet, no blowing, all
        wall)."""
        cfg = self.cfg

        def dbl(a, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != shape:
                raise ValueError("expected shape %s, got %s" % (shape, a.shape))
            return _ptr(a), a

        def i32(a, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.int32)
            if a.shape != shape:
                raise ValueError("expected shape %s, got %s" % (shape, a.shape))
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), a

        items = [dbl(inlet, (5, cfg.nz, cfg.ny)), dbl(upper, (5, cfg.nz, cfg.nx)), i32(nonref, (cfg.nz, cfg.nx)),
                 dbl(pertb, (cfg.nz, cfg.nx)), i32(wall, (cfg.nz, cfg.nx))]
        ptrs = [it[0] if it is not None else None for it in items]
        return self._call("set_bc_boundarylayer", ctypes.c_double(aoa), *ptrs)

    def set_bc_time_factor(self, ht):
        return self._call("set_bc_time_factor", ctypes.c_double(ht))

    def filtering(self):
        return self._call("filtering")

    def advance_clock(self):
        return self._call("advance_clock")

    def step(self, nsteps=1):
        return self._call("step", ctypes.c_int32(nsteps))

    def stage(self, krk):
        """One RK stage through the per-routine entry points, in the driver's
        order (OCFD_NS_Solver.f90:33-60)."""
        cfg = self.cfg
        self.exchange_primitives()
        if krk == 1 and cfg.scheme_invis.upper() == "HYBRID":
            self.comput_rhybrid()
        if cfg.if_scheme_character:
            self.du_inviscous_character()
        else:
            self.du_inviscous()
        if cfg.if_viscous:
            self.du_viscous()
        self.time_adv_rk3(krk)
        n = self.comput_duvwt()
        self.bc()
        return n

    def step_by_routines(self, nsteps=1):
        for _ in range(nsteps):
            self.save_fn()
            for krk in (1, 2, 3):
                self.stage(krk)
            self.advance_clock()
            self.filtering()

    # -- diagnostics ----------------------------------------------------------
    def show_flow_msg(self):
        out = (ctypes.c_double * 3)()
        self._call("show_flow_msg", out)
        return np.array(out[:])

    def ana_box(self):
        out = (ctypes.c_double * 20)()
        self._call("ana_box", out)
        return np.array(out[:])


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libocfd_hip.so; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OcfdError("HIP library not built: %s (run __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path)
    lib.ocfd_last_error.restype = ctypes.c_char_p
    lib.ocfd_create.argtypes = [ctypes.POINTER(OcfdConfigC), ctypes.c_int,
                                ctypes.POINTER(ctypes.c_void_p)]
    for name in dir(lib):
        pass
    _lib = lib
    return lib


class Block(BlockBase):
    """One block on one GPU (ocfd_create / ocfd_destroy)."""

    prefix = "ocfd"

    def __init__(self, cfg: CaseConfig, device: int = 0, lib=None):
        super().__init__(lib or load_library(), cfg)
        h = ctypes.c_void_p()
        rc = self.lib.ocfd_create(ctypes.byref(self._c), device, ctypes.byref(h))
        if rc != 0:
            raise OcfdError("ocfd_create failed (%d): %s" % (rc, self._err()))
        self.h = h

    def _err(self):
        msg = self.lib.ocfd_last_error()
        return msg.decode() if msg else ""

    def set_uniform_metrics(self, m10):
        m = np.ascontiguousarray(m10, dtype=np.float64)
        return self._call("set_uniform_metrics", _ptr(m))

    def synchronize(self):
        return self._call("synchronize")

    def show_flow_msg_noalloc(self):
        """ocfd_show_flow_msg_noalloc: E, K, S with the partial sums in the
        library's file-scope device scratch (no allocation)."""
        out = (ctypes.c_double * 3)()
        self._call("show_flow_msg_noalloc", out)
        return np.array(out[:])

    # -- sweep tiles ------------------------------------------------------------
    def cfl_dt(self, cfl):
        """ocfd_cfl_dt: dt = cfl / (largest CFL rate over the cells)."""
        dt = ctypes.c_double()
        self._call("cfl_dt", ctypes.c_double(cfl), ctypes.byref(dt))
        return dt.value

    def set_sweep_tiles(self, shapes):
        a = (ctypes.c_int32 * 3)(*[int(x) for x in shapes])
        return self._call("set_sweep_tiles", a)

    def get_sweep_tiles(self):
        a = (ctypes.c_int32 * 3)()
        self._call("get_sweep_tiles", a)
        return list(a)

    def autotune(self, reps=5):
        """Time every tile shape of each axis' sweep on the current state and
        keep the fastest; returns the chosen shapes.  Overwrites du."""
        a = (ctypes.c_int32 * 3)()
        self._call("autotune", ctypes.c_int32(reps), a)
        return list(a)

    def profile_enable(self, on=True):
        return self._call("profile_enable", ctypes.c_int32(1 if on else 0))

    def profile_reset(self):
        return self._call("profile_reset")

    def profile_read(self, max_entries=64):
        names = (ctypes.c_char * 32 * max_entries)()
        ms = (ctypes.c_double * max_entries)()
        cnt = (ctypes.c_int64 * max_entries)()
        n = self._fn("profile_read")(self.h, max_entries, names, ms, cnt)
        if n < 0:
            raise OcfdError(self._err())
        return {names[i].value.decode(): (ms[i], cnt[i]) for i in range(n)}

    def close(self):
        if self.h is not None:
            self.lib.ocfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class BoxView(Block):
    """A box of a :class:`Batch`: the single-block interface on a handle the
    batch owns (it is not destroyed here)."""

    def __init__(self, lib, cfg: CaseConfig, h):
        BlockBase.__init__(self, lib, cfg)
        self.h = h

    def close(self):
        self.h = None


class Batch:
    """ocfd_batch_*: independent periodic boxes advanced together, one launch
    per kernel phase over all boxes."""

    def __init__(self, cfgs, device: int = 0, lib=None):
        self.lib = lib or load_library()
        self.cfgs = list(cfgs)
        n = len(self.cfgs)
        arr = (OcfdConfigC * n)(*[c.to_c() for c in self.cfgs])
        h = ctypes.c_void_p()
        rc = self.lib.ocfd_batch_create(arr, ctypes.c_int32(n), ctypes.c_int32(device), ctypes.byref(h))
        if rc != 0:
            raise OcfdError("ocfd_batch_create failed (%d): %s" % (rc, self._err()))
        self.h = h
        self.boxes = []
        for i in range(n):
            hb = ctypes.c_void_p()
            self._call("box", ctypes.c_int32(i), ctypes.byref(hb))
            self.boxes.append(BoxView(self.lib, self.cfgs[i], hb))

    def _err(self):
        msg = self.lib.ocfd_last_error()
        return msg.decode() if msg else ""

    def _call(self, name, *args):
        rc = getattr(self.lib, "ocfd_batch_" + name)(self.h, *args)
        if rc < 0:
            raise OcfdError("ocfd_batch_%s failed (%d): %s" % (name, rc, self._err()))
        return rc

    def box(self, i) -> BoxView:
        return self.boxes[i]

    def step(self, nsteps=1):
        return self._call("step", ctypes.c_int32(nsteps))

    def autotune(self, reps=5):
        """Tile shapes per scheme group and the marching segment length."""
        self._call("autotune", ctypes.c_int32(reps))
        info = self.info()
        return {"sweep_tiles": info[1], "kseg": info[3]}

    def set_sweep_tiles(self, group, shapes):
        a = (ctypes.c_int32 * 3)(*[int(x) for x in shapes])
        return self._call("set_sweep_tiles", ctypes.c_int32(group), a)

    def set_launch_mode(self, mode):
        """0: one launch per kernel phase; 1: one persistent launch per step
        (device work queue of (box, phase, tile) items)."""
        return self._call("set_launch_mode", ctypes.c_int32(mode))

    def info(self):
        """(number of scheme groups, their tile shapes, total cells, marching
        segment length)"""
        ng = ctypes.c_int32()
        cells = ctypes.c_int64()
        kseg = ctypes.c_int32()
        self._call("info", ctypes.byref(ng), None, ctypes.byref(cells), None)
        sh = (ctypes.c_int32 * (3 * max(ng.value, 1)))()
        self._call("info", ctypes.byref(ng), sh, ctypes.byref(cells), ctypes.byref(kseg))
        return ng.value, [list(sh[3 * g:3 * g + 3]) for g in range(ng.value)], cells.value, kseg.value

    def synchronize(self):
        for b in self.boxes:
            b.synchronize()

    def close(self):
        if self.h is not None:
            for b in self.boxes:
                b.close()
            self.lib.ocfd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def setup_block(blk: BlockBase, case, use_uniform=None):
    """Grid + initial state as init3d + the first OCFD_bc (OCFD_NS_Solver.f90:14-15)."""
    cfg = case.cfg
    if use_uniform is None:
        use_uniform = cfg.metric_mode == 1
    if use_uniform:
        blk.set_uniform_metrics(case.uniform_metrics)
    else:
        blk.set_grid(case.x, case.y, case.z, case.spans)
    blk.upload_primitives(case.halo_prims(), 0, 0.0)
    blk.bc()
    return blk
