"""Locate and load the native libraries of the package.

The product path is native: `libscp_amd.so` (HIP kernels + C-ABI) and
`libscp_io.so` (host C: config, SFA container, seeds). Nothing here falls back
to Python or CPU arithmetic: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
_cache = {}


class NativeLibraryMissing(RuntimeError):
    pass


def lib_path(name):
    return os.path.join(LIBDIR, name)


def load(name):
    """Load lib/<name> once; raise NativeLibraryMissing if it was never built."""
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `make -C scp_amd` (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    _cache[name] = lib
    return lib
