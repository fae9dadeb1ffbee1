"""Python host binding of libhpccg_hip (MI355X CG path of HPCCG-F90) over its C ABI.

The Fortran modules under hpccg-f90_amd/fortran are the drop-in for the reference's own
driver; this module is the same boundary seen from Python (tests, benchmark, tooling).
Every call goes through the C ABI declared in include/hpccg_hip.h; there is no CPU fallback:
if the library is missing, importing succeeds but creating a Context raises.
"""
from ._lib import (HpccgError, Context, DeviceArray, lib, lib_path, DOT_SEQUENTIAL, DOT_TREE,  # noqa: F401
                   exported_symbols, xstar_sines, xstar_uniform, batch_dims)
