"""Run configuration: the namelist ``&control_opencfd`` of the reference
(OCFD_readpara.f90:14-65) as a Python object, with conversions to

* the C ABI struct ``ocfd_config`` (include/ocfd_hip.h), and
* the ``opencfd2.in`` namelist the reference solver reads.

Defaults are the reference's defaults (OCFD_readpara.f90:28-65).
"""
from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass, field

SCHEMES = {"WENO5": 1, "WENO7": 2, "OMP6": 3, "UD7L": 4, "CD6": 5, "CD8": 6, "HYBRID": 10}
BCS = {"BC_NONE": 0, "BC_BLUNT2D": 1, "BC_BOUNDARYLAYER": 2, "BC_SWEPTCORNER": 3, "BC_USER_DEF": 99}
GRID_BOX, GRID1D, GRID2D_PLANE, GRID3D, GRID_AND_JACOBIAN3D = 0, 10, 20, 30, 31
PREC_FP64, PREC_MIXED = 0, 1
METRIC_FIELD, METRIC_UNIFORM = 0, 1


class OcfdConfigC(ctypes.Structure):
    """Mirror of ``struct ocfd_config`` in include/ocfd_hip.h."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lap", ctypes.c_int32),
        ("iperiodic", ctypes.c_int32 * 3),
        ("scheme_invis", ctypes.c_int32), ("scheme_vis", ctypes.c_int32),
        ("flux_splitting", ctypes.c_int32), ("if_scheme_character", ctypes.c_int32),
        ("if_viscous", ctypes.c_int32), ("ibc", ctypes.c_int32),
        ("if_mass_force", ctypes.c_int32),
        ("scheme_boundary", ctypes.c_int32 * 6), ("ghost_cell", ctypes.c_int32 * 6),
        ("nfiltering", ctypes.c_int32), ("istep_show", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("metric_mode", ctypes.c_int32),
        ("reserved_i", ctypes.c_int32 * 8),
        ("ma", ctypes.c_double), ("re", ctypes.c_double), ("gamma", ctypes.c_double),
        ("pr", ctypes.c_double), ("ref_amu_t0", ctypes.c_double), ("dt", ctypes.c_double),
        ("ud7l_diss", ctypes.c_double), ("mass_force", ctypes.c_double * 3),
        ("bc_para", ctypes.c_double * 100), ("hybrid_para", ctypes.c_double * 100),
        ("filter_para", (ctypes.c_double * 100) * 10),
        ("reserved_d", ctypes.c_double * 8),
    ]


@dataclass
class CaseConfig:
    nx: int
    ny: int
    nz: int
    lap: int = 4
    iperiodic: tuple = (0, 0, 0)
    scheme_invis: str = "WENO5"
    scheme_vis: str = "CD6"
    flux_splitting: int = 1
    if_scheme_character: int = 0
    if_viscous: int = 1
    boundary_condition: str = "BC_NONE"
    if_mass_force: int = 0
    scheme_boundary: tuple = (0, 0, 0, 0, 0, 0)
    ghost_cell: tuple = (0, 0, 0, 0, 0, 0)
    istep_show: int = 1
    ma: float = 1.0
    re: float = 100.0
    gamma: float = 1.4
    pr: float = 0.7
    ref_amu_t0: float = 288.15
    dt: float = 1.0e-4
    ud7l_diss: float = 1.0
    mass_force: tuple = (0.0, 0.0, 0.0)
    bc_para: dict = field(default_factory=dict)          # 1-based index -> value
    hybrid_para: dict = field(default_factory=dict)      # 1-based index -> value
    filters: list = field(default_factory=list)          # list of 13-value lists
    precision: int = PREC_FP64
    metric_mode: int = METRIC_FIELD
    grid_type: int = GRID_BOX
    periodic_span: tuple = (0.0,) * 9                    # I(3), J(3), K(3)

    def replace(self, **kw) -> "CaseConfig":
        return dataclasses.replace(self, **kw)

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def halo_shape(self):
        L = self.lap
        return (self.nz + 2 * L, self.ny + 2 * L, self.nx + 2 * L)

    def to_c(self) -> OcfdConfigC:
        c = OcfdConfigC()
        c.nx, c.ny, c.nz, c.lap = self.nx, self.ny, self.nz, self.lap
        for a in range(3):
            c.iperiodic[a] = int(self.iperiodic[a])
        c.scheme_invis = SCHEMES[self.scheme_invis.upper()]
        c.scheme_vis = SCHEMES[self.scheme_vis.upper()]
        c.flux_splitting = int(self.flux_splitting)
        c.if_scheme_character = int(self.if_scheme_character)
        c.if_viscous = int(self.if_viscous)
        c.ibc = BCS[self.boundary_condition.upper()]
        c.if_mass_force = int(self.if_mass_force)
        for q in range(6):
            c.scheme_boundary[q] = int(self.scheme_boundary[q])
            c.ghost_cell[q] = int(self.ghost_cell[q])
        c.nfiltering = len(self.filters)
        c.istep_show = int(self.istep_show)
        c.precision = int(self.precision)
        c.metric_mode = int(self.metric_mode)
        c.ma, c.re, c.gamma, c.pr = self.ma, self.re, self.gamma, self.pr
        c.ref_amu_t0, c.dt, c.ud7l_diss = self.ref_amu_t0, self.dt, self.ud7l_diss
        for a in range(3):
            c.mass_force[a] = self.mass_force[a]
        for q in range(100):
            c.hybrid_para[q] = 1.0                    # reference default Hybrid_para(:)=1
        for k, v in self.hybrid_para.items():
            c.hybrid_para[k - 1] = float(v)
        for k, v in self.bc_para.items():
            c.bc_para[k - 1] = float(v)
        if len(self.filters) > 10:
            raise ValueError("NFiltering <= 10")
        for m, fl in enumerate(self.filters):
            for q, v in enumerate(fl):
                c.filter_para[m][q] = float(v)
        return c

    def namelist(self, end_time: float, istep_save: int = 1000000, ana_box_every: int = 0) -> dict:
        nl = {
            "nx_global": self.nx, "ny_global": self.ny, "nz_global": self.nz,
            "npx0": 1, "npy0": 1, "npz0": 1, "LAP": self.lap,
            "Iperiodic_X": self.iperiodic[0], "Iperiodic_Y": self.iperiodic[1],
            "Iperiodic_Z": self.iperiodic[2],
            "Scheme_Invis": self.scheme_invis, "Scheme_Vis": self.scheme_vis,
            "Flux_Splitting": self.flux_splitting,
            "IF_Scheme_Character": self.if_scheme_character, "IF_Viscous": self.if_viscous,
            "BoundaryCondition": self.boundary_condition,
            "IF_Mass_Force": self.if_mass_force, "Mass_Force": list(self.mass_force),
            "Scheme_boundary": list(self.scheme_boundary), "Ghost_Cell": list(self.ghost_cell),
            "Istep_Show": self.istep_show, "Istep_Save": istep_save,
            "Ma": float(self.ma), "Re": float(self.re), "gamma": float(self.gamma),
            "Pr": float(self.pr), "Ref_Amu_T0": float(self.ref_amu_t0), "dt": float(self.dt),
            "End_time": float(end_time), "UD7L_Diss": float(self.ud7l_diss),
            "Iflag_Gridtype": self.grid_type,
            "Periodic_ISpan": list(self.periodic_span[0:3]),
            "Periodic_JSpan": list(self.periodic_span[3:6]),
            "Periodic_KSpan": list(self.periodic_span[6:9]),
            "NFiltering": len(self.filters),
        }
        for k, v in self.hybrid_para.items():
            nl["Hybrid_para(%d)" % k] = float(v)
        for k, v in self.bc_para.items():
            nl["BC_Para(%d)" % k] = float(v)
        for m, fl in enumerate(self.filters):
            nl["Filter_Para(:,%d)" % (m + 1)] = [float(x) for x in fl]
        if ana_box_every:
            nl["ANA_Number"] = 1
            nl["ANA_Para(:,1)"] = [102.0, float(ana_box_every)]
        return nl
