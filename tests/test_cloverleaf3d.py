"""CloverLeaf 3D (an API v1 application: pointer kernels with OPS_ACC<n> accessors): the
oracle (host backend running the same application) pinned bit for bit to the reference's
own dev_seq build of apps/c/CloverLeaf_3D, and the gfx950 path checked against the same
fixtures: every dat, halos and work arrays included, bitwise; the field summary against
the exactly rounded sum of its per-cell terms."""
import ctypes
import hashlib
import math
import tempfile
from pathlib import Path

import numpy as np
import pytest

import ops_amd
from dumps import read_dump
from ops_amd.apps import SIGNATURES
from ops_amd.apps.cloverleaf3d.decks import GOLD_KE_BM, bm_deck

ROOT = Path(__file__).resolve().parents[1]
CPU_LIB = ROOT / "oracle" / "_build" / "libcloverleaf3d_cpu.so"
_cpu = None


def lib(kind):
    global _cpu
    if kind == "gpu":
        return ops_amd.app("cloverleaf3d")
    if _cpu is None:
        _cpu = ctypes.CDLL(str(CPU_LIB), mode=ctypes.RTLD_LOCAL)
        for fn, (res, args) in SIGNATURES["cloverleaf3d"].items():
            f = getattr(_cpu, fn)
            f.restype = res
            f.argtypes = args
        _cpu.ops_cpu_set_threads.argtypes = [ctypes.c_int]
    return _cpu


def run(kind, deck_text, dump_dir=None, fetch=(), threads=1, tokens=()):
    L = lib(kind)
    if kind == "cpu":
        L.ops_cpu_set_threads(threads)
    with tempfile.TemporaryDirectory() as t:
        deck = Path(t) / "clover.in"
        deck.write_text(deck_text)
        argc, argv = ops_amd.argv(*([f"OPS_DUMP_DIR={dump_dir}"] if dump_dir else []), *tokens)
        assert L.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
        n = L.cloverleaf3d_run(1 << 30)
        assert n >= 0
        s = np.zeros(5)
        assert L.cloverleaf3d_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
        got = {}
        for name in fetch:
            cnt = L.cloverleaf3d_fetch(name.encode(), None)
            buf = np.zeros(cnt)
            assert L.cloverleaf3d_fetch(name.encode(), buf.ctypes.data_as(ctypes.c_void_p)) == cnt
            got[name] = buf
        assert L.cloverleaf3d_finalize() == 0
    return n, s, got


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_dump(d, fx):
    got = read_dump(d)
    bad = []
    for name, meta in fx["dats"].items():
        if name not in got:
            bad.append((name, "missing"))
            continue
        arr = got[name]["data"]
        if list(arr.shape) != meta["shape"] or sha(arr) != meta["sha256"]:
            bad.append((name, float(np.asarray(arr, dtype=np.float64).sum()), meta["sum"]))
    return bad


def summary_ref(fx):
    r = fx["reductions"]
    return np.array([r[k]["values"][0] for k in ("vol", "mass", "ie", "ke", "press")])


def mesh(deck):
    v = {}
    for line in deck.splitlines():
        for key in ("x_cells", "y_cells", "z_cells"):
            if line.strip().startswith(key):
                v[key] = int(line.split("=")[1])
    return v["x_cells"], v["y_cells"], v["z_cells"]


SUMMARY_FIELDS = ("volume", "density0", "energy0", "pressure", "xvel0", "yvel0", "zvel0")


def exact_summary(got, nx, ny, nz):
    """field_summary (field_summary_kernel.h) with the per-cell terms evaluated in the
    kernel's operation order and summed exactly (math.fsum), from full-size arrays
    (2-cell halos) as returned by cloverleaf3d_fetch."""
    cell = lambda k: got[k].reshape(nz + 9, ny + 9, nx + 9)[4:nz + 4, 4:ny + 4, 4:nx + 4]
    node = {k: got[k].reshape(nz + 10, ny + 10, nx + 10) for k in ("xvel0", "yvel0", "zvel0")}
    vs = np.zeros((nz, ny, nx))
    for k in range(2):
        for j in range(2):
            for i in range(2):
                sl = (slice(4 + k, nz + 4 + k), slice(4 + j, ny + 4 + j), slice(4 + i, nx + 4 + i))
                u, v, w = node["xvel0"][sl], node["yvel0"][sl], node["zvel0"][sl]
                vs = vs + 0.125 * (u * u + v * v + w * w)
    vol = cell("volume")
    mass = vol * cell("density0")
    terms = [vol, mass, mass * cell("energy0"), mass * 0.5 * vs, vol * cell("pressure")]
    return np.array([math.fsum(t.ravel()) for t in terms])


def check_summary(s, got, fx, nx, ny, nz, tol=1e-13):
    """Sums are order dependent. Against the exactly rounded sum of the same per-cell
    terms a tree sum is within 1e-13; the reference's sequential sum is within n*eps."""
    exact = exact_summary(got, nx, ny, nz)
    assert np.max(np.abs(s - exact) / np.abs(exact)) < tol
    ref = summary_ref(fx)
    assert np.max(np.abs(ref - exact) / np.abs(exact)) < nx * ny * nz * np.finfo(np.float64).eps


# ---------------------------------------------------------------- oracle pins (CPU)

@pytest.mark.parametrize("name", ["clover3d_small", "clover3d_states"])
def test_oracle_bitwise_equals_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    n, s, _ = run("cpu", fx["deck"], dump_dir=str(tmp_path))
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    # one thread accumulates in the reference's order: the summary is bitwise equal too
    assert np.array_equal(s, summary_ref(fx))


def test_bm_deck_is_the_reference_deck(golden):
    assert bm_deck() == golden("clover3d_bm.json")["deck"]


def test_oracle_threaded_benchmark_deck(golden, tmp_path):
    # 96^3, 87 steps on 8 threads: dt comes from a MIN reduction (order independent), so
    # the state is bitwise the reference's; the summary differs only by rounding
    fx = golden("clover3d_bm.json")
    n, s, got = run("cpu", fx["deck"], dump_dir=str(tmp_path), fetch=SUMMARY_FIELDS, threads=8)
    assert n == 87
    assert check_dump(tmp_path, fx) == []
    # each thread adds its share of the cells in sequence: error bound (96^3 / 8) * eps
    check_summary(s, got, fx, 96, 96, 96, tol=96 ** 3 / 8 * np.finfo(np.float64).eps)
    assert abs(100.0 * (s[3] / GOLD_KE_BM) - 100.0) < 0.001


# ---------------------------------------------------------------- gfx950 path

@pytest.mark.gpu
@pytest.mark.parametrize("name", ["clover3d_small", "clover3d_states", "clover3d_bm"])
def test_gpu_bitwise_equals_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    nx, ny, nz = mesh(fx["deck"])
    n, s, got = run("gpu", fx["deck"], dump_dir=str(tmp_path), fetch=SUMMARY_FIELDS)
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    check_summary(s, got, fx, nx, ny, nz)
    if name == "clover3d_bm":
        assert abs(100.0 * (s[3] / GOLD_KE_BM) - 100.0) < 0.001


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["clover3d_small", "clover3d_states"])
def test_gpu_captured_steps_bitwise_equal_reference(golden, tmp_path, name):
    """Steps replayed from the two captured hipGraphs (dt computed on the GPU), in chunks
    separated by eager work (field summaries, state reads, synchronisations): every dat
    bitwise equal to the reference after the deck's steps."""
    fx = golden(f"{name}.json")
    L = lib("gpu")
    L.cloverleaf3d_set_graph(1)
    try:
        deck = tmp_path / "clover.in"
        deck.write_text(fx["deck"])
        argc, argv = ops_amd.argv(f"OPS_DUMP_DIR={tmp_path}")
        assert L.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
        left = fx["steps"]
        for chunk in (2, 1, 3):
            assert L.cloverleaf3d_step(chunk) == 0
            left -= chunk
            s = np.zeros(5)
            assert L.cloverleaf3d_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
        assert L.cloverleaf3d_step(left) == 0
        st = np.zeros(4)
        L.cloverleaf3d_state(st.ctypes.data_as(ctypes.c_void_p))
        assert st[0] == fx["steps"]
        s = np.zeros(5)
        assert L.cloverleaf3d_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
        assert L.cloverleaf3d_finalize() == 0
    finally:
        L.cloverleaf3d_set_graph(0)
    assert check_dump(tmp_path, fx) == []


@pytest.mark.gpu
def test_gpu_autotuned_bitwise_equals_reference(golden, tmp_path):
    """The benchmark deck with the autotuner on: the loops try their candidate workgroup
    shapes on their first launches, so the steps mix shapes; every dat is still bitwise
    the reference's."""
    fx = golden("clover3d_bm.json")
    n, s, got = run("gpu", fx["deck"], dump_dir=str(tmp_path), fetch=SUMMARY_FIELDS,
                    tokens=("OPS_AUTOTUNE=1",))
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    check_summary(s, got, fx, 96, 96, 96)


SPHERE_DECK = """*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 xvel=0.25 geometry=sphere xmin=3.7 ymin=5.2 zmin=4.4 radius=2.3
 x_cells=40
 y_cells=36
 z_cells=44
 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax=10.0
 zmax=10.0
 initial_timestep=0.04
 max_timestep=0.04
 end_step=1
 end_time=100.0
 test_problem 0
 checkpoint_frequency=10
*endclover
"""


@pytest.mark.parametrize("kind", ["cpu", pytest.param("gpu", marks=pytest.mark.gpu)])
def test_sphere_state_covers_the_cells_whose_centre_is_inside(kind):
    """The sphere branch of the generate kernel, which the reference's deck reader cannot
    reach (it never reads a state's radius): after set-up, a cell holds the sphere state's
    density and energy exactly when its centre lies within the radius of the (off-centre)
    sphere centre, the node at its lower corner carries the state's velocity, and the
    flagged cells' volume matches 4/3 pi r^3 to the discretisation error."""
    L = lib(kind)
    if kind == "cpu":
        L.ops_cpu_set_threads(4)
    with tempfile.TemporaryDirectory() as t:
        deck = Path(t) / "clover.in"
        deck.write_text(SPHERE_DECK)
        argc, argv = ops_amd.argv()
        assert L.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
        got = {}
        for name in ("density0", "energy0", "xvel0", "cellx", "celly", "cellz"):
            cnt = L.cloverleaf3d_fetch(name.encode(), None)
            buf = np.zeros(cnt)
            assert L.cloverleaf3d_fetch(name.encode(), buf.ctypes.data_as(ctypes.c_void_p)) == cnt
            got[name] = buf
        assert L.cloverleaf3d_finalize() == 0
    cx, cy, cz = got["cellx"], got["celly"], got["cellz"]
    nx, ny, nz = cx.size, cy.size, cz.size            # cell allocation extents
    rho = got["density0"].reshape(nz, ny, nx)
    e = got["energy0"].reshape(nz, ny, nx)
    vel = got["xvel0"].reshape(nz + 1, ny + 1, nx + 1)[:nz, :ny, :nx]
    # the deck reader moves every state's min corner up by 1/100 of a cell, computed in
    # float (apps/c/CloverLeaf_3D/read_input.cpp:304-315): for a sphere that is its centre
    cen = [c + float(np.float32(10.0 / n)) / 100.0 for c, n in ((3.7, 40), (5.2, 36), (4.4, 44))]
    d2 = ((cx[None, None, :] - cen[0]) ** 2 + (cy[None, :, None] - cen[1]) ** 2 +
          (cz[:, None, None] - cen[2]) ** 2)
    inside = np.sqrt(d2) <= 2.3
    written = rho != 0.0                               # the generate loop's range
    assert written.sum() >= 40 * 36 * 44
    assert np.array_equal(rho[written], np.where(inside, 1.0, 0.2)[written])
    assert np.array_equal(e[written], np.where(inside, 2.5, 1.0)[written])
    assert np.array_equal(vel[written], np.where(inside, 0.25, 0.0)[written])
    # interior cells flagged x cell volume against the sphere's volume (the sphere lies
    # inside the domain: 3.7 - 2.3 > 0 and 4.4 + 2.3 < 10)
    interior = (cx[None, None, :] > 0) & (cx[None, None, :] < 10) & (cy[None, :, None] > 0) & \
        (cy[None, :, None] < 10) & (cz[:, None, None] > 0) & (cz[:, None, None] < 10)
    vol = (inside & interior).sum() * (10 / 40) * (10 / 36) * (10 / 44)
    assert abs(vol / (4.0 / 3.0 * math.pi * 2.3 ** 3) - 1.0) < 0.03
