"""Locate and load the package's native libraries (ctypes).

The libraries are built in-tree by ``build.py`` (``__graft_entry__.build()``
drives it).  Loading fails loudly: there is no Python or CPU stand-in for the
solver library.
"""
import ctypes
import os
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_SOLVER = PKG_DIR / "libtfs_solver.so"
LIB_MESH = PKG_DIR / "libtfs_mesh.so"

_cache = {}


class NativeLibraryMissing(RuntimeError):
    pass


def load(path):
    path = pathlib.Path(path)
    key = str(path)
    if key in _cache:
        return _cache[key]
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path.name} is not built ({path}); run build() in __graft_entry__.py")
    lib = ctypes.CDLL(str(path), mode=getattr(os, "RTLD_GLOBAL", 0))
    _cache[key] = lib
    return lib


def mesh_lib():
    return load(LIB_MESH)


def solver_lib():
    return load(LIB_SOLVER)
