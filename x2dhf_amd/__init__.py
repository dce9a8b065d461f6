"""x2dhf_amd: MI355X-native relaxation engine for finite-difference diatomic
Hartree-Fock (the x2dhf method).

Layers:
  include/x2gpu.h, x2dhf_amd/csrc/*.hip   HIP kernels + C ABI (libx2gpu.so)
  x2dhf_amd/fortran/*.f90                  Fortran 90 SCF driver (libx2scf.so)
  x2dhf_amd.gpu                            ctypes binding of the C ABI
  x2dhf_amd.scf                            ctypes binding of the SCF driver
"""
from . import gpu  # noqa: F401

__all__ = ["gpu"]
