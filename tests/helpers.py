"""Shared test helpers (no GPU work at import time)."""
import ctypes
import hashlib

import numpy as np


def sha256(arr: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


def first_diff(a, b):
    """Index and values of the first element that differs bitwise (for messages)."""
    av = np.ascontiguousarray(a).view(np.uint64).ravel()
    bv = np.ascontiguousarray(b).view(np.uint64).ravel()
    idx = np.flatnonzero(av != bv)
    if idx.size == 0:
        return None
    i = int(idx[0])
    return i, float(np.asarray(a).ravel()[i]), float(np.asarray(b).ravel()[i]), int(idx.size)
