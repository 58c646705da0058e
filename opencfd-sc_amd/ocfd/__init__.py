"""Host side of the MI355X right-hand-side + RK3 path of OpenCFD-SC.

* :mod:`ocfd.config`     run parameters (the reference namelist) and the C ABI struct
* :mod:`ocfd.cases`      initial conditions / grids of the benchmark and test cases
* :mod:`ocfd.fortran_io` the reference's restart, grid and namelist file formats
* :mod:`ocfd.native`     ctypes binding of libocfd_hip.so (the HIP library)
"""
from .config import CaseConfig  # noqa: F401
