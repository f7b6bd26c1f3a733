"""kittlespt_amd — MI355X-native path tracer with the KittlesPT renderer API.

Hot path: hand-written gfx950 HIP kernels in csrc/hip behind the C-ABI of
include/kpt_abi.h (libkpt_hip.so).  This package holds the Python mirror of
the reference's Renderer API (renderer.py), the ABI binding (abi.py, context.py),
seeded synthetic scenes (scenes.py) and the build step (build.py).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
