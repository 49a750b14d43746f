"""Drive the CloverLeaf application (GPU build or oracle host build) from a deck."""
import ctypes
import os
import tempfile
from pathlib import Path

import numpy as np

from ops_amd.apps import SIGNATURES

ROOT = Path(__file__).resolve().parents[1]
CPU_LIB = ROOT / "oracle" / "_build" / "libcloverleaf_cpu.so"


def _lib_cpu():
    lib = ctypes.CDLL(str(CPU_LIB), mode=ctypes.RTLD_LOCAL)
    for fn, (res, args) in SIGNATURES["cloverleaf"].items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


_cpu = None


def lib(kind):
    global _cpu
    if kind == "gpu":
        import ops_amd
        return ops_amd.app("cloverleaf")
    if _cpu is None:
        _cpu = _lib_cpu()
    return _cpu


def run(kind, deck_text, steps=None, dump_dir=None, fetch=(), chain=0):
    """Run to the deck's end (or `steps` steps); returns (steps, summary[5], fetched dats).
    chain > 1: the loops between halo exchanges run as loop chains of that many bands."""
    import ops_amd
    L = lib(kind)
    with tempfile.TemporaryDirectory() as t:
        deck = Path(t) / "clover.in"
        deck.write_text(deck_text)
        toks = [f"OPS_DUMP_DIR={dump_dir}"] if dump_dir else []
        argc, argv = ops_amd.argv(*toks)
        assert L.cloverleaf_init(argc, argv, str(deck).encode()) == 0
        L.cloverleaf_set_chain(chain)
        n = L.cloverleaf_step(steps) if steps is not None else L.cloverleaf_run(1 << 30)
        if steps is not None:
            assert n == 0
            n = steps
        assert n >= 0
        s = np.zeros(5)
        assert L.cloverleaf_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
        got = {}
        for name in fetch:
            cnt = L.cloverleaf_fetch(name.encode(), None)
            buf = np.zeros(cnt)
            assert L.cloverleaf_fetch(name.encode(), buf.ctypes.data_as(ctypes.c_void_p)) == cnt
            got[name] = buf
        assert L.cloverleaf_finalize() == 0
    return n, s, got


# apps/c/CloverLeaf/clover_bm16_short.in (the benchmark configuration)
BM16_DECK = """*clover
 state 1 density=0.2 energy 1.0
 state 2 density=1.0 energy=2.5 geometry=rectangle xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0
 x_cells=3840
 y_cells=3840
 xmin=0.0
 ymin=0.0
 xmax=10.0
 ymax=10.0
 initial_timestep=0.04
 timestep_rise=1.5
 max_timestep=0.04
 end_time=0.55
 end_step=87
 test_problem 4
*endclover
"""


SUMMARY_FIELDS = ("volume", "density0", "energy0", "pressure", "xvel0", "yvel0")


def exact_summary(got, nx, ny):
    """field_summary (field_summary_kernel.h) with the per-cell terms evaluated in the
    kernel's operation order and summed exactly (math.fsum), from full-size arrays
    (2-cell halos) as returned by cloverleaf_fetch."""
    import math
    cw, nw = nx + 9, nx + 10      # row widths of cell and node dats, halos included
    c = {k: got[k].reshape(-1, cw) for k in ("volume", "density0", "energy0", "pressure")}
    u = got["xvel0"].reshape(-1, nw)
    v = got["yvel0"].reshape(-1, nw)
    sl = (slice(4, ny + 4), slice(4, nx + 4))
    u00, u10 = u[4:ny + 4, 4:nx + 4], u[4:ny + 4, 5:nx + 5]
    u01, u11 = u[5:ny + 5, 4:nx + 4], u[5:ny + 5, 5:nx + 5]
    v00, v10 = v[4:ny + 4, 4:nx + 4], v[4:ny + 4, 5:nx + 5]
    v01, v11 = v[5:ny + 5, 4:nx + 4], v[5:ny + 5, 5:nx + 5]
    vs = 0.0 + 0.25 * (u00 * u00 + v00 * v00)
    vs = vs + 0.25 * (u10 * u10 + v10 * v10)
    vs = vs + 0.25 * (u01 * u01 + v01 * v01)
    vs = vs + 0.25 * (u11 * u11 + v11 * v11)
    vol = c["volume"][sl]
    mass = vol * c["density0"][sl]
    terms = [vol, mass, mass * c["energy0"][sl], mass * 0.5 * vs, vol * c["pressure"][sl]]
    return np.array([math.fsum(t.ravel()) for t in terms])
