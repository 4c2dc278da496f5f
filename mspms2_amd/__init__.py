"""mspms2_amd -- MI355X (gfx950) force engine for the mspms2 simulator.

Drop-in replacement for mspms2's frclong / frcshort force routines:
hand-written HIP kernels behind the C ABI of include/mspms_gpu.h
(libmspms_gpu.so), with this package as the Python host side:

  runio    mspms2 run-directory formats (in.mspms / cfg.mspms / xyz.mspms)
  system   the reference's unit reduction -> C ABI structs
  configs  seeded synthetic workloads
  engine   ctypes binding: Engine(system).frclong() / .frcshort()
  build    in-tree hipcc build of the library
"""
__all__ = ["abi", "runio", "system", "configs", "engine", "build"]
