"""Loader for the native libraries built in-tree (``ppc_amd/lib``).

The HIP library (``libppc_hip.so``) is the only compute path; nothing here
falls back to Python or to the CPU when it is missing: the loaders raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")
BINDIR = os.path.join(_HERE, "bin")

_libs = {}


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError("native library %s is not built (run __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def hip_lib():
    return _load("libppc_hip.so")


def xppc_lib():
    return _load("libxppc.so")


def ppc_exe():
    path = os.path.join(BINDIR, "ppc")
    if not os.path.exists(path):
        raise RuntimeError("ppc executable is not built")
    return path
