"""ether_amd - MI355X-native Monte Carlo sweep engine with ETHER's inputs.

The compute path is the gfx950 library ether_amd/lib/libether_gpu.so (HIP
kernels behind the C-ABI of include/ether_gpu.h).  This package is its Python
host side: input readers (inputs), the ctypes binding (engine), the batched
driver (driver), the learned tempering ladder (ladder) and the benchmark
workloads (workloads).  The Fortran driver in ether_amd/fortran binds the same
C-ABI through iso_c_binding.
"""
from .engine import Engine, EtherError, build_neighbours, build_spiral_partners, load_library  # noqa: F401
