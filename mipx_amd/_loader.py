"""Loads the in-tree native libraries of the package.

The libraries are built in-tree by mipx_amd/build.py (hipcc for the HIP
library, g++ for the host-only generator).  A missing library is an error:
there is no fallback path.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
_cache = {}


def library_path(name):
    return os.path.join(PKG_DIR, name)


def load_library(name, local=False):
    """local=True keeps the library's symbols out of the process-wide scope:
    the C++ host library defines mipx:: classes that must not interpose on
    another library of the same namespace loaded later in one process."""
    if name in _cache:
        return _cache[name]
    path = library_path(name)
    if not os.path.exists(path):
        raise ImportError(
            f"{name} is not built ({path} missing); run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m mipx_amd.build` first")
    lib = C.CDLL(path, mode=C.RTLD_LOCAL if local else C.RTLD_GLOBAL)
    _cache[name] = lib
    return lib
