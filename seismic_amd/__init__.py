"""seismic_amd — MI355X-native acoustic finite-difference engine.

A drop-in for the libwave propagator of esiwgnahz/seismic and the modeling,
Born, RTM/LSRTM and FWI-gradient drivers above it, behind a C-ABI
(include/seisfd.h, libseisfd.so).  Python here is test/bench plumbing over
that ABI (ctypes); Fortran callers use seismic_amd/fortran/seisfd_mod.f90.
"""
from .abi import (SFD_SNAP_FP32, SFD_SNAP_FP16S, SFD_SNAP_FP16G, SFD_EXTRACT_SINC_V2, SFD_EXTRACT_NEAREST,  # noqa: F401
                  SFD_EXTRACT_SINC_LSRTM, SFD_EXTRACT_SINC_UNIT, SFD_NORM_L2, SFD_NORM_L1, SFD_NORM_HUBER,
                  SFD_NORM_CAUCHY, SFD_NORM_HYPERBOLIC)
from .engine import (Engine, Shot, LBFGS, LBFGSOC, CNN, lbfgs_run, lbfgsoc_step_files, lbfgsoc_rotate_files,  # noqa: F401
                     make_grid, device_info, SeisFDError, Engine3D, Shot3D, make_grid3d)
