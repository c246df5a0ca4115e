"""cosmosphere_amd -- SPH density / smoothing-length and force solver for the AMD MI355X.

The solver itself is native: gfx950 HIP kernels and C++ host code behind a C ABI
(include/cosmosphere.h), built into cosmosphere_amd/lib/libcosmoSPHere{,_quintic}.{so,a}.
This package holds the sources (csrc/), a ctypes mirror of the C ABI (api.py) and the
synthetic workload generators of the benchmark configurations (workloads.py).
"""
from . import workloads  # noqa: F401
from .api import Library, library_path, POS, HSML, VEL, ACCEL, DENS_OUT, GRAD_OUT, THERMO, FORCE_OUT  # noqa: F401
