"""MI355X LC2 similarity engine: hand-written HIP kernels behind a C ABI
(include/lc2e.h), with the reference's registration interfaces layered on top.

Submodules:
  engine    ctypes binding of the C ABI (requires the in-tree liblc2e.so)
  geometry  pose -> VirtualToMoving matrices on the host
  synth     seeded synthetic workloads
  build     hipcc build recipe
"""
__all__ = ["engine", "geometry", "synth", "build"]
