"""lpm_amd: MI355X (gfx950) direct sums of LPM's sphere BVE/SWE solvers.

The compute path is the HIP library ``liblpm_mi355x.so`` behind the C ABI in
``include/lpm_mi355x.h``; this package holds its Python host mirror (mesh
inputs, galleries, BVESphere/BVERK4 and SWE classes).  Importing the package
does not touch the GPU; calling any device entry point without a built
library or without a HIP device raises.
"""
from . import _abi
from ._abi import LpmError

__all__ = ["LpmError", "lib", "version"]


def lib():
    return _abi.lib()


def version() -> int:
    return _abi.lib().lpm_abi_version()
