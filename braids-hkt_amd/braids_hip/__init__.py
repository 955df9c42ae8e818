"""braids_hip: MI355X-native hydraulic time step of BRAIDS-HKT.

Hand-written HIP kernels for gfx950 (libbraids_hip.so, C ABI) behind a thin
ctypes host layer.  Modules:
  mesh    unstructured triangle meshes, connectivity contract, renumbering
  solver  resident SSPRK(3,3) solver handle (FP64 / FP32+FP64-fixups)
  engine  reference-named operator kernels (Engine.launch)
  build   hipcc build of the library
"""
from .mesh import Mesh, plain_channel, braided_reach, lake_at_rest, structured_triangles  # noqa: F401

__all__ = ['Mesh', 'plain_channel', 'braided_reach', 'lake_at_rest', 'structured_triangles']
