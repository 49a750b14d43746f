"""Load an application's GPU build or its oracle host build (ctypes)."""
import ctypes
from pathlib import Path

from ops_amd.apps import SIGNATURES

ROOT = Path(__file__).resolve().parents[1]
_cpu = {}


def lib(name, kind):
    if kind == "gpu":
        import ops_amd
        return ops_amd.app(name)
    if name not in _cpu:
        L = ctypes.CDLL(str(ROOT / "oracle" / "_build" / f"lib{name}_cpu.so"), mode=ctypes.RTLD_LOCAL)
        for fn, (res, args) in SIGNATURES[name].items():
            f = getattr(L, fn)
            f.restype = res
            f.argtypes = args
        L.ops_cpu_set_threads.argtypes = [ctypes.c_int]
        _cpu[name] = L
    return _cpu[name]


def argv(*tokens):
    import ops_amd
    return ops_amd.argv(*tokens)


def p(a):
    """ctypes pointer of a contiguous numpy array"""
    return a.ctypes.data_as(ctypes.c_void_p)
