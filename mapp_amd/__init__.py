"""MI355X-native DMD force field (MAPP's ForceField_DMD / _I / _II) behind a C ABI.

The product is the HIP library mapp_amd/lib/libmapp_dmd.so (sources in
mapp_amd/csrc, ABI in include/mapp_dmd.h).  This package holds its build
recipe, a ctypes view of the ABI and the synthetic workload generator.
"""
from .dmd import DMDError, ForceFieldDMD, load_library, setup  # noqa: F401
