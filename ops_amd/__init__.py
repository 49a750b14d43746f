"""OPS for MI355X: host-side loader for libops_hip.so and the translated applications.

The product is native: libops_hip.so (C ABI declared in include/*.h) and one
lib<app>_hip.so per translated application, all built in tree by `ops_amd.build`.
This module only finds and loads them; it raises if they are missing instead of
falling back to any other implementation.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = Path(__file__).resolve().parent / "lib"
INCLUDE_DIR = ROOT / "include"

_runtime = None
_apps = {}


class NotBuiltError(RuntimeError):
    pass


def _load(path: Path, mode=ctypes.RTLD_GLOBAL):
    if not path.exists():
        raise NotBuiltError(f"{path} is missing: run `python -m ops_amd.build` first")
    return ctypes.CDLL(str(path), mode=mode)


def runtime():
    """libops_hip.so (loaded once, RTLD_GLOBAL so applications bind to it)."""
    global _runtime
    if _runtime is None:
        _runtime = _load(LIB_DIR / "libops_hip.so")
        _declare_runtime(_runtime)
    return _runtime


def app(name: str):
    """lib<name>_hip.so with its extern "C" entry points declared."""
    if name not in _apps:
        runtime()
        # local: two builds of one application (AoS and SoA) export the same stub names
        lib = _load(LIB_DIR / f"lib{name}_hip.so", mode=ctypes.RTLD_LOCAL)
        from . import apps as _a
        _a.declare(name, lib)
        _apps[name] = lib
    return _apps[name]


def _declare_runtime(L):
    i, d, p, c = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
    L.ops_get_last_error.restype = c
    L.ops_hip_kernel_count.restype = i
    L.ops_hip_kernel_stats.restype = i
    L.ops_hip_kernel_stats.argtypes = [i, c, i, ctypes.POINTER(ctypes.c_long),
                                       ctypes.POINTER(d), ctypes.POINTER(d)]
    L.ops_hip_set_timing.argtypes = [i]
    L.ops_hip_get_stream.restype = p
    L.ops_hip_set_block_shape.argtypes = [i, i, i, i]
    L.ops_hip_copy_probe.restype = d
    L.ops_hip_copy_probe.argtypes = [ctypes.c_size_t, i]


def kernel_stats():
    """Per-kernel table: name -> dict(calls, ms, bytes) from the runtime's HIP events."""
    L = runtime()
    out = {}
    name = ctypes.create_string_buffer(256)
    calls, ms, by = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
    for k in range(L.ops_hip_kernel_count()):
        if L.ops_hip_kernel_stats(k, name, 256, ctypes.byref(calls), ctypes.byref(ms),
                                  ctypes.byref(by)):
            n = name.value.decode()
            e = out.setdefault(n, {"calls": 0, "ms": 0.0, "bytes": 0.0, "id": k})
            e["calls"] += calls.value
            e["ms"] += ms.value
            e["bytes"] += by.value
    return out


def argv(*tokens):
    """A C argv array (argv[0] = program name) for ops_init-style entry points."""
    toks = [b"ops_amd"] + [t.encode() if isinstance(t, str) else t for t in tokens]
    arr = (ctypes.c_char_p * (len(toks) + 1))(*toks, None)
    return len(toks), arr
