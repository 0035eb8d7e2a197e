"""gp5_amd: MI355X (gfx950) Gross-Pitaevskii pseudo-spectral core.

Layers:
  include/gp5.h, gp5_amd/csrc/*.hip   HIP kernels + C ABI (libgp5.so)
  gp5_amd/csrc/gp5pp.hpp              C++ facade: gp::functional, gp::operators::laplacian,
                                      gp::stepper, gp::constraint, ... over the C ABI
  gp5_amd/csrc/gpCpp.cpp              pybind11 module gpCpp (reference module surface)
  gp5_amd/capi.py                     ctypes view of the C ABI
  gp5_amd/gp.py                       Python API with the reference's names (gp.geometry, ...)
  gp5_amd/driver.py                   YAML-configured runs (the reference's gp executable)
  gp5_amd/ground_state.py             batched ground states with coarse-grid starts
  gp5_amd/nn.py, gp5_amd/surrogate.py learned warm start (bf16 MFMA U-Net)
"""
from . import capi  # noqa: F401

__all__ = ["capi", "gp"]
