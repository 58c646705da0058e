"""Readers and writers for the solver's on-disk formats.

All files are Fortran unformatted *sequential* files: every record is framed by
a 4-byte little-endian length marker on both sides.

* ``opencfd2.in``  namelist ``&control_opencfd`` (OCFD_readpara.f90:14-22)
* ``opencfd.dat`` / ``OCFDnnnnnnnn.dat``  record (int32 Istep, float64 tt),
  then for each of d, u, v, w, T: nz records of nx*ny float64, i fastest
  (OCFD_IO.f90:24-35 via read_3d, OCFD_IO_mpi.f90:166-202)
* ``OCFD-grid.dat``  GRID1D: x(nx), y(ny), z(nz) (OCFD_mesh.f90:228-230);
  GRID2D_PLANE: x2(nx,ny), y2(nx,ny), z1(nz) (OCFD_mesh.f90:262-264);
  GRID3D: three fields in the read_3d layout (OCFD_mesh.f90:195-197).

Arrays are numpy float64 with shape (nz, ny, nx) in C order, which is the same
memory order as the Fortran (nx, ny, nz) arrays.
"""
from __future__ import annotations

import struct

import numpy as np


def _write_record(fh, payload: bytes) -> None:
    n = len(payload)
    fh.write(struct.pack("<i", n))
    fh.write(payload)
    fh.write(struct.pack("<i", n))


def _read_record(fh) -> bytes:
    head = fh.read(4)
    if len(head) != 4:
        raise EOFError("unexpected end of Fortran sequential file")
    (n,) = struct.unpack("<i", head)
    payload = fh.read(n)
    (m,) = struct.unpack("<i", fh.read(4))
    if m != n or len(payload) != n:
        raise ValueError("corrupt Fortran record framing")
    return payload


def write_flow_data(path, istep: int, tt: float, fields) -> None:
    """Write d, u, v, w, T (each (nz, ny, nx)) in the restart-file layout."""
    with open(path, "wb") as fh:
        _write_record(fh, struct.pack("<id", int(istep), float(tt)))
        for a in fields:
            a = np.ascontiguousarray(a, dtype="<f8")
            for k in range(a.shape[0]):
                _write_record(fh, a[k].tobytes())


def read_flow_data(path, shape):
    """Return (istep, tt, [d, u, v, w, T]) from a restart/snapshot file."""
    nz, ny, nx = shape
    with open(path, "rb") as fh:
        istep, tt = struct.unpack("<id", _read_record(fh))
        out = []
        for _ in range(5):
            a = np.empty((nz, ny, nx), dtype=np.float64)
            for k in range(nz):
                a[k] = np.frombuffer(_read_record(fh), dtype="<f8").reshape(ny, nx)
            out.append(a)
    return istep, tt, out


def write_grid_1d(path, x, y, z) -> None:
    with open(path, "wb") as fh:
        for a in (x, y, z):
            _write_record(fh, np.ascontiguousarray(a, dtype="<f8").tobytes())


def write_grid_2d_plane(path, x2, y2, z1) -> None:
    """x2, y2 have shape (ny, nx) (Fortran x2(nx,ny)); z1 has shape (nz,)."""
    with open(path, "wb") as fh:
        for a in (x2, y2, z1):
            _write_record(fh, np.ascontiguousarray(a, dtype="<f8").tobytes())


def write_grid_3d(path, *fields) -> None:
    """OCFD-grid.dat of Iflag_Gridtype 30 (x, y, z) or 31 (x, y, z and the ten
    metric fields Akx .. Asz, Ajac): each field as nz records of nx*ny values
    (read_mesh3d, OCFD_mesh.f90:188-215)."""
    with open(path, "wb") as fh:
        for a in fields:
            a = np.ascontiguousarray(a, dtype="<f8")
            for k in range(a.shape[0]):
                _write_record(fh, a[k].tobytes())


def _fmt(v) -> str:
    if isinstance(v, str):
        return '"%s"' % v
    if isinstance(v, (bool, np.bool_)):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return repr(float(v)).replace("e", "d")
    if isinstance(v, (list, tuple, np.ndarray)):
        return ", ".join(_fmt(x) for x in v)
    raise TypeError(type(v))


def write_namelist(path, params: dict) -> None:
    """Write ``&control_opencfd``.  Keys whose name contains '(' are written
    verbatim (e.g. ``"ANA_Para(:,1)"``, ``"Filter_Para(:,1)"``)."""
    lines = ["&control_opencfd"]
    for key, val in params.items():
        lines.append(" %s=%s" % (key, _fmt(val)))
    lines.append("/")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
