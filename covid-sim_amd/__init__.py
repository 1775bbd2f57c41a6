"""MI355X infection step for CovidSim: host-side package.

  abi         ctypes mirror of include/covidsim_hip.h
  population  population / state containers for the C ABI views
  engine      the HIP library (libcovidsim_hip.so) behind the C ABI
"""
from . import abi, population  # noqa: F401
