"""cupf_amd -- MI355X-native active-parameter-tracking (APT) multi-grain
phase-field step, a drop-in for the step of YigongQin/cuPF's 3Dgrain /
3DgrainMPI on one GPU.

Layers:
  include/pf_capi.h, cupf_amd/csrc/   C ABI + hand-written gfx950 kernels
  cupf_amd/host/                      C++ PhaseField/APTPhaseField drop-in and CLI
  cupf_amd/capi.py                    ctypes binding of the C ABI
  cupf_amd/cases.py                   seeded synthetic inputs (grid, substrate, orientations)
"""
from . import capi, cases  # noqa: F401

__all__ = ["capi", "cases"]
