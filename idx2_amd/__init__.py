"""MI355X (gfx950) implementation of the idx2 encode/decode path.

The compute lives in libidx2g.so (hand-written HIP kernels, C ABI in
include/idx2g.h); this package is the Python-side binding and the synthetic
input generator.  Importing it does not need a GPU; calls that reach the
library raise if the shared library has not been built.
"""
from . import synth  # noqa: F401

__all__ = ["synth", "api"]
