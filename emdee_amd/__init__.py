"""MI355X-native EmDee force engine.

The engine is the HIP/C++ shared library `emdee_amd/libemdee.so`, which
exports the reference C ABI (include/emdee.h).  This package only locates and
binds it; nothing here computes forces.  Importing `load()` raises if the
library has not been built: there is no CPU fallback.
"""
import os

from .emdee import EmDeeLibrary, tEmDee  # noqa: F401

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(PACKAGE_DIR, "libemdee.so")

_lib = None


def load():
    """Return the bound product library, failing loudly if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(hipcc --offload-arch=gfx950)")
        _lib = EmDeeLibrary(LIBRARY_PATH)
    return _lib
