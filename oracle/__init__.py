"""ORACLE — test infrastructure only.

CPU restatement of the reference OPS algorithms on the hot path (see the header of each
oracle/*.c file for the reference file:line it follows). Loaded by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker; never by the
product (ops_amd/), which fails loudly when its HIP libraries are missing.
"""
import ctypes
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _declare(_lib)
    return _lib


def _declare(L):
    d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    L.oracle_laplace.restype = d
    L.oracle_laplace.argtypes = [i, i, i, i, p, p, i]
    for name, (res, args) in _EXTRA.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


_EXTRA = {}


def register(name, res, args):
    _EXTRA[name] = (res, args)
