"""fasteig_amd: MI355X-native (gfx950) block-eigensolver hot path of fasteig.

Layers:
  include/fe_hip.h          C ABI of the kernel library (libfe_hip.so)
  include/fasteig_hip.h     C ABI of the solver layer (libfasteig_hip.so)
  fasteig_amd._fe           ctypes binding of the kernel library
  fasteig_amd.api           ctypes binding of the solver layer (fasteig_problem /
                            fasteig_options / solve, serving loop)
  fasteig_amd.generators    seeded benchmark matrices
"""
__version__ = "0.1.0"
