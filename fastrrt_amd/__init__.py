"""fastrrt_amd — FastRRT's parallel planning loop on MI355X (gfx950).

The planner lives in the in-tree shared library fastrrt_amd/lib/libfastrrt.so
(HIP kernels behind the reference's C ABI, include/fastrrt.h).  This package
holds its ctypes binding (`_lib`), the pyfastrrt-compatible interface
(`pyfastrrt`), the seeded scene generator (`scenes`) and the build script.
"""
from . import _lib  # noqa: F401

__all__ = ["load_library"]


def load_library():
    """Load libfastrrt.so; raises ImportError when it has not been built."""
    return _lib.load()
