"""CloverLeaf 2D: the oracle (host backend running the same application) pinned to the
reference's own dev_seq build; the GPU path checked against the reference fixtures and
the oracle."""
import ctypes
import hashlib
from pathlib import Path

import numpy as np
import pytest

import clover_util as cu
from dumps import read_dump

GOLD_KE = {"clover_default": 1.82280367310258, "clover_bm_short": 1.19316898756307,
           "clover_bm16_short": 0.307475452287895}   # field_summary.cpp:91-95


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


def set_cpu_threads(n):
    L = cu.lib("cpu")
    L.ops_cpu_set_threads.argtypes = [ctypes.c_int]
    L.ops_cpu_set_threads(n)


# ---------------------------------------------------------------- oracle pins (CPU)

@pytest.mark.parametrize("name", ["clover_small", "clover_default"])
def test_oracle_bitwise_equals_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    set_cpu_threads(1)
    n, s, _ = cu.run("cpu", fx["deck"], dump_dir=str(tmp_path))
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    # one thread accumulates in the reference's order: the summary is bitwise equal too
    assert np.array_equal(s, summary_ref(fx))


def test_oracle_known_answer_default(golden):
    fx = golden("clover_default.json")
    set_cpu_threads(1)
    _, s, _ = cu.run("cpu", fx["deck"])
    assert abs(100.0 * (s[3] / GOLD_KE["clover_default"]) - 100.0) < 0.001


def check_summary(s, got, fx, nx, ny, tol=1e-13):
    """Sums are order dependent. Against the exactly rounded sum of the same per-cell
    terms the GPU's tree sum is within 1e-13 (a few hundred sequential additions per
    lane at most); the reference's own sequential sum is within its n*eps bound."""
    exact = cu.exact_summary(got, nx, ny)
    assert np.max(np.abs(s - exact) / np.abs(exact)) < tol
    ref = summary_ref(fx)
    n_eps = nx * ny * np.finfo(np.float64).eps
    assert np.max(np.abs(ref - exact) / np.abs(exact)) < n_eps


def test_oracle_threaded_state_is_bitwise_and_summary_close(golden, tmp_path):
    # dt comes from a MIN reduction (order independent), so the state is bitwise equal
    # whatever the thread count; the summary sums differ only by rounding
    fx = golden("clover_bm_short.json")
    set_cpu_threads(8)
    n, s, got = cu.run("cpu", fx["deck"], dump_dir=str(tmp_path), fetch=cu.SUMMARY_FIELDS)
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    # 8 threads each add n/8 cells in sequence: bounded by (n/8)*eps
    check_summary(s, got, fx, 960, 960, tol=960 * 960 / 8 * np.finfo(np.float64).eps)
    assert abs(100.0 * (s[3] / GOLD_KE["clover_bm_short"]) - 100.0) < 0.001


def test_oracle_conserves_mass_and_volume(golden):
    fx = golden("clover_small.json")
    set_cpu_threads(4)
    _, s0, _ = cu.run("cpu", fx["deck"], steps=1)
    _, s1, _ = cu.run("cpu", fx["deck"], steps=10)
    assert abs(s1[0] - s0[0]) < 1e-12 * s0[0]      # volume
    assert abs(s1[1] - s0[1]) < 1e-11 * s0[1]      # mass (reflective boundaries)


# ---------------------------------------------------------------- GPU path

DECK_SIZE = {"clover_small": (48, 40), "clover_default": (10, 2), "clover_bm_short": (960, 960)}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["clover_small", "clover_default", "clover_bm_short"])
def test_gpu_bitwise_equals_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    n, s, got = cu.run("gpu", fx["deck"], dump_dir=str(tmp_path), fetch=cu.SUMMARY_FIELDS)
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    check_summary(s, got, fx, *DECK_SIZE[name])


@pytest.mark.gpu
@pytest.mark.parametrize("name,chain", [("clover_small", 3), ("clover_small", 16),
                                        ("clover_default", 4), ("clover_bm_short", 8)])
def test_gpu_loop_chains_bitwise_equal_reference(golden, tmp_path, name, chain):
    # the loops between halo exchanges run as loop chains (band by band, skewed bands):
    # every dat stays bitwise the reference's
    fx = golden(f"{name}.json")
    n, s, got = cu.run("gpu", fx["deck"], dump_dir=str(tmp_path), fetch=cu.SUMMARY_FIELDS, chain=chain)
    assert n == fx["steps"]
    assert check_dump(tmp_path, fx) == []
    check_summary(s, got, fx, *DECK_SIZE[name])


@pytest.mark.gpu
def test_gpu_matches_oracle_midrun(golden):
    # stop mid-run (not at the deck's end) and compare fields and summary with the oracle
    fx = golden("clover_small.json")
    names = ("density0", "energy1", "xvel1", "yvel0", "mass_flux_x", "vol_flux_y", "viscosity")
    _, sg, fg = cu.run("gpu", fx["deck"], steps=7, fetch=names)
    set_cpu_threads(2)
    _, sc, fc = cu.run("cpu", fx["deck"], steps=7, fetch=names)
    for k in names:
        assert np.array_equal(fg[k].view(np.uint64), fc[k].view(np.uint64)), k
    assert np.max(np.abs(sg - sc) / np.abs(sc)) < 1e-12


@pytest.mark.gpu
def test_gpu_config2_known_answer(golden):
    # clover_bm16_short.in: 3840 x 3840 cells, 87 steps (the benchmark workload)
    p = Path(__file__).parent / "golden" / "clover_bm16_short.json"
    fx = golden("clover_bm16_short.json") if p.exists() else None
    deck = fx["deck"] if fx else cu.BM16_DECK
    names = cu.SUMMARY_FIELDS + ("viscosity", "energy1", "xvel1")
    n, s, got = cu.run("gpu", deck, fetch=names)
    assert n == 87
    assert abs(100.0 * (s[3] / GOLD_KE["clover_bm16_short"]) - 100.0) < 0.001
    exact = cu.exact_summary(got, 3840, 3840)
    assert np.max(np.abs(s - exact) / np.abs(exact)) < 1e-13
    if fx:
        check_summary(s, got, fx, 3840, 3840)
        for k in names:
            meta = fx["dats"][k]
            arr = got[k].reshape(meta["shape"])
            assert sha(arr) == meta["sha256"], k


@pytest.mark.gpu
def test_gpu_captured_steps_bitwise_equal_reference(golden, tmp_path):
    """Steps replayed from the two captured hipGraphs (dt computed on the GPU), in chunks
    separated by eager work (field summaries, state reads, host synchronisations): every
    dat bitwise equal to the reference after the deck's steps, time and dt equal to the
    eager run's."""
    import ops_amd
    fx = golden("clover_small.json")
    L = cu.lib("gpu")
    L.cloverleaf_set_graph(1)
    try:
        deck = tmp_path / "clover.in"
        deck.write_text(fx["deck"])
        argc, argv = ops_amd.argv(f"OPS_DUMP_DIR={tmp_path}")
        assert L.cloverleaf_init(argc, argv, str(deck).encode()) == 0
        left, states = fx["steps"], []
        for chunk in (3, 1, 5):
            assert L.cloverleaf_step(chunk) == 0
            left -= chunk
            s = np.zeros(5)
            assert L.cloverleaf_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
            st = np.zeros(4)
            L.cloverleaf_state(st.ctypes.data_as(ctypes.c_void_p))
            states.append(st)
        assert L.cloverleaf_step(left) == 0
        s = np.zeros(5)
        assert L.cloverleaf_summary(s.ctypes.data_as(ctypes.c_void_p)) == 0
        assert L.cloverleaf_finalize() == 0
    finally:
        L.cloverleaf_set_graph(0)
    assert check_dump(tmp_path, fx) == []
    # the eager run's time and dt after the same step counts
    L.cloverleaf_set_graph(0)
    argc, argv = ops_amd.argv()
    assert L.cloverleaf_init(argc, argv, str(deck).encode()) == 0
    for chunk, st in zip((3, 1, 5), states):
        assert L.cloverleaf_step(chunk) == 0
        ref = np.zeros(4)
        L.cloverleaf_state(ref.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(st, ref)
    assert L.cloverleaf_finalize() == 0
