"""otfem: MI355X-native corotated tetrahedral FEM step (OpenTissue::fem path).

Native code lives in ``csrc/`` (HIP kernels for gfx950 and the C ABI declared
in ``include/otfem/otfem.h``); ``fem`` is the Python mirror of the reference
API over that ABI and ``mesh`` holds the caller-side mesh builders.
"""
from . import mesh  # noqa: F401
from . import fem  # noqa: F401

__all__ = ["mesh", "fem"]
