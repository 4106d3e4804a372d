"""dpmi: MI355X-native se_e2_a energy/force/virial path.

Modules:
  lib      ctypes binding of libdpmi.so (C ABI of include/dpmi.h)
  ops      torch-tensor front-end of the kernel-level entry points
           (deepmd::*_gpu of the reference source/lib)
  engine   whole-path engine handle (cell list .. force/virial)
  model    se_e2_a model data and the compressed-table builder
  systems  seeded water / copper generators
"""
__all__ = ["lib", "ops", "engine", "model", "systems"]
