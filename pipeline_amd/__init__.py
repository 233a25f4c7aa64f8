"""MI355X (gfx950) backend of nvpro-pipeline's scene-update path: transform-tree
world-matrix propagation and bounding-box frustum culling.

The product is the C ABI in include/dpcull_hip.h (libdpcull_hip.so) and the C++
mirrors of dp::culling::Manager and dp::transform::Tree in include/dpcull/.
This package holds their sources (csrc/), the build recipe, a ctypes binding
used by the tests and the benchmark, and the seeded workload generators.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
