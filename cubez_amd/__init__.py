"""cubez_amd — the cubez simulation step (Integrate, the narrowphase,
ResolveContacts) as hand-written HIP kernels for MI355X (gfx950), behind the
C ABI of include/cubez_amd.h.

  World        batched worlds in HBM, one launch per phase per step
  scenes       seeded scene generators for the benchmark configurations
"""
from . import scenes
from ._native import CubezError, LIB_PATH
from .world import World

__all__ = ["World", "scenes", "CubezError", "LIB_PATH"]
