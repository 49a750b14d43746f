"""Drive the TeaLeaf application (GPU build or oracle host build)."""
import ctypes
import tempfile
from pathlib import Path

import numpy as np

from ops_amd.apps import SIGNATURES

ROOT = Path(__file__).resolve().parents[1]
_cpu = None


def lib(kind):
    global _cpu
    if kind == "gpu":
        import ops_amd
        return ops_amd.app("tealeaf")
    if _cpu is None:
        _cpu = ctypes.CDLL(str(ROOT / "oracle" / "_build" / "libtealeaf_cpu.so"),
                           mode=ctypes.RTLD_LOCAL)
        for fn, (res, args) in SIGNATURES["tealeaf"].items():
            if fn.startswith("tea_unet_"):  # the network lives in the device build only
                continue
            f = getattr(_cpu, fn)
            f.restype = res
            f.argtypes = args
        _cpu.ops_cpu_set_threads.argtypes = [ctypes.c_int]
    return _cpu


FIELDS = ("density", "energy1", "u", "u0", "vector_Kx", "vector_Ky", "vector_r")


def run(kind, deck_text, steps=None, solver=0, dump_dir=None, fetch=FIELDS, threads=1):
    """Returns dict(per_step=[state...], summary, fields)."""
    import ops_amd
    L = lib(kind)
    if kind == "cpu":
        L.ops_cpu_set_threads(threads)
    with tempfile.TemporaryDirectory() as t:
        deck = Path(t) / "tea.in"
        deck.write_text(deck_text)
        toks = [f"OPS_DUMP_DIR={dump_dir}"] if dump_dir else []
        argc, argv = ops_amd.argv(*toks)
        assert L.tealeaf_init(argc, argv, str(deck).encode()) == 0
        assert L.tealeaf_set_solver(solver) == 0
        per = []
        st = np.zeros(7)
        n = 0
        while True:
            assert L.tealeaf_step(1) == 0
            n += 1
            L.tealeaf_state(st.ctypes.data_as(ctypes.c_void_p))
            per.append(st.copy())
            end = int(deck_end_step(deck_text))
            if (steps is not None and n >= steps) or (steps is None and n >= end):
                break
        s = np.zeros(4)
        assert L.tealeaf_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
        got = {}
        for name in fetch:
            cnt = L.tealeaf_fetch(name.encode(), None)
            buf = np.zeros(cnt)
            assert L.tealeaf_fetch(name.encode(), buf.ctypes.data_as(ctypes.c_void_p)) == cnt
            got[name] = buf
        assert L.tealeaf_finalize() == 0
    return {"per_step": per, "summary": s, "fields": got}


def deck_end_step(deck):
    for line in deck.splitlines():
        t = line.replace("=", " ").split()
        if t and t[0] == "end_step":
            return float(t[1])
    return 1e9


def deck_cells(deck):
    nx = ny = 10
    for line in deck.splitlines():
        t = line.replace("=", " ").split()
        if t and t[0] == "x_cells":
            nx = int(float(t[1]))
        if t and t[0] == "y_cells":
            ny = int(float(t[1]))
    return nx, ny


def operator_residual(f, nx, ny, rx, ry):
    """r = u0 - A u on the interior, A the 5-point conduction operator of
    tea_leaf_common_residual_kernel, evaluated independently in numpy."""
    sx = nx + 2
    u = f["u"].reshape(-1, sx)
    u0 = f["u0"].reshape(-1, sx)
    Kx = f["vector_Kx"].reshape(-1, sx)
    Ky = f["vector_Ky"].reshape(-1, sx)
    c = (slice(1, ny + 1), slice(1, nx + 1))
    kx0, kx1 = Kx[1:ny + 1, 1:nx + 1], Kx[1:ny + 1, 2:nx + 2]
    ky0, ky1 = Ky[1:ny + 1, 1:nx + 1], Ky[2:ny + 2, 1:nx + 1]
    Au = ((1.0 + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * u[c]
          - ry * (ky1 * u[2:ny + 2, 1:nx + 1] + ky0 * u[0:ny, 1:nx + 1])
          - rx * (kx1 * u[1:ny + 1, 2:nx + 2] + kx0 * u[1:ny + 1, 0:nx]))
    return u0[c] - Au
