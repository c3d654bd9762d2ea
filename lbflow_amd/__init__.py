"""MI355X-native lattice Boltzmann engine with LBFlow's file and process contract.

``lbflow_amd.engine`` binds the C ABI of the HIP engine (include/lbflow.h),
``lbflow_amd.cases`` describes lattices and writes input.prm + geometry.raw,
``lbflow_amd.vti`` reads VTI output,
``lbflow_amd.geometry`` generates the deterministic porous media of the
benchmarks.
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIB_DIR = PACKAGE_DIR / 'lib'
BIN_DIR = PACKAGE_DIR / 'bin'

__all__ = ['PACKAGE_DIR', 'LIB_DIR', 'BIN_DIR']
