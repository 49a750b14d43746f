"""The reference's own applications, sources unchanged, as drop-ins on this backend.

ops_amd/dropin.py copies each application's sources from the reference tree, translates
them with the hip_gfx950 target and links them with libops_hip.so. Here:
  - (CPU, where the reference tree is present) all five translate, compile and link;
  - (GPU) the executables run the reference's decks, and every dat they dump at ops_exit
    equals the reference's own sequential build bit for bit (the committed dev_seq
    fixtures), with the reduction handles' last values byte for byte. TeaLeaf's solves sum
    in another order on the GPU: its set-up dats are exact, iteration counts within one,
    and u, energy1 and the temperature within 1e-12 of the reference's sequential build
    (oracle/_ref/tealeaf_dev_seq, run on the same deck) — the reordering alone moves them
    far less than that (test_tealeaf.py::test_oracle_reordered_sums_stay_within_field_bar).
ADI (apps/c/adi) builds, but ends by writing its field with HDF5 (adi.cpp:441-442), which
this image lacks; it is not run.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dumps import read_dump, read_reductions
from helpers import sha256

ROOT = Path(__file__).resolve().parents[1]
DROP = ROOT / "ops_amd" / "_dropin"
REF_BIN = ROOT / "oracle" / "_ref"
REF_TREE = Path("/root/reference/apps/c")


def exe(app):
    p = DROP / app / app
    if not p.exists():
        pytest.skip(f"{p} not built (python -m ops_amd.dropin)")
    return p


def run(binary, cwd, deck=None, deck_name=None, env=None, timeout=900):
    print(f"[dropin] {binary} in {cwd}", flush=True)
    if deck is not None:
        (Path(cwd) / deck_name).write_text(deck)
    r = subprocess.run([str(binary)], cwd=cwd, env=dict(os.environ, **(env or {})),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def run_gpu(app, tmp_path, deck=None, deck_name=None):
    dump = tmp_path / "dump"
    dump.mkdir()
    out = run(exe(app), tmp_path, deck, deck_name, env={"OPS_DUMP_DIR": str(dump)})
    return out, dump


def dump_mismatches(dump, fx):
    got = read_dump(dump)
    bad = []
    for name, meta in fx["dats"].items():
        if name not in got:
            bad.append((name, "missing"))
            continue
        arr = got[name]["data"]
        if list(arr.shape) != meta.get("shape", list(arr.shape)) or sha256(arr) != meta["sha256"]:
            bad.append((name, float(np.asarray(arr, dtype=np.float64).sum()), meta["sum"]))
    return bad


def reduction_mismatches(dump, fx):
    got = read_reductions(dump)
    return [k for k, v in fx["reductions"].items()
            if k not in got or got[k].tobytes().hex() != v["hex"]]


# ---------------------------------------------------------------- build (CPU)

@pytest.mark.skipif(not REF_TREE.exists(), reason="reference tree absent")
def test_reference_apps_translate_compile_and_link(tmp_path):
    if not (ROOT / "ops_amd" / "lib" / "libops_hip.so").exists():
        pytest.skip("libops_hip.so not built")
    from ops_amd import dropin
    exes = dropin.build_all(out=tmp_path)
    assert [e.name for e in exes] == list(dropin.APPS)
    for e in exes:
        assert e.is_file() and os.access(e, os.X_OK), e


def test_makefile_lists_join_continuation_lines(tmp_path):
    from ops_amd.dropin import makefile_lists
    (tmp_path / "Makefile").write_text("HEADERS=a.h\nOPS_FILES=a.cpp b.cpp \\\n  c.cpp\n"
                                       "OPS_GENERATED_SUFFIX=_ops\nOTHER_FILES = d.cpp\n")
    assert makefile_lists(tmp_path) == (["a.cpp", "b.cpp", "c.cpp"], ["d.cpp"])


# ---------------------------------------------------------------- GPU runs

@pytest.mark.gpu
def test_gpu_laplace_step7_equals_reference(golden, tmp_path):
    fx = golden("laplace_step7.json")
    out, dump = run_gpu("laplace2d_step7", tmp_path)
    assert "PASSED" in out, out[-500:]
    assert dump_mismatches(dump, fx) == []
    assert reduction_mismatches(dump, fx) == []


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["clover_small", "clover_default", "clover_bm_short"])
def test_gpu_cloverleaf_equals_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    out, dump = run_gpu("cloverleaf", tmp_path, fx["deck"], "clover.in")
    assert (tmp_path / "clover.out").read_text().count(" Step ") == fx["steps"]
    assert dump_mismatches(dump, fx) == []
    assert reduction_mismatches(dump, fx) == []


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["clover3d_small", "clover3d_states", "clover3d_bm"])
def test_gpu_cloverleaf3d_equals_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    out, dump = run_gpu("cloverleaf3d", tmp_path, fx["deck"], "clover.in")
    assert (tmp_path / "clover.out").read_text().count(" Step ") == fx["steps"]
    if name == "clover3d_bm":
        assert "This test is considered PASSED" in out, out[-500:]
    assert dump_mismatches(dump, fx) == []
    assert reduction_mismatches(dump, fx) == []


TEA_SETUP = ("density", "volume", "xarea", "yarea", "cellx", "celly", "celldx", "celldy",
             "vertexx", "vertexy")


def tea_iterations(text):
    return [int(l.split()[-1]) for l in text.splitlines()
            if "Iteration count" in l and "PPCG" not in l]


def maxrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tea_cg_64", "tea_cg_200", "tea_jacobi_ref", "tea_cheby_64",
                                  "tea_ppcg_64", "tea_cg_jacdiag_64"])
def test_gpu_tealeaf_matches_reference(golden, tmp_path, name):
    fx = golden(f"{name}.json")
    gdir = tmp_path / "gpu"
    gdir.mkdir()
    _, dump = run_gpu("tealeaf", gdir, fx["deck"], "tea.in")
    got = read_dump(dump)
    for k in TEA_SETUP:
        if k in fx["dats"]:
            assert sha256(got[k]["data"]) == fx["dats"][k]["sha256"], k
    iters = tea_iterations((gdir / "tea.out").read_text())
    ref_iters = fx["iteration_counts_printed"]
    assert len(iters) == len(ref_iters)
    assert all(abs(a - b) <= 1 for a, b in zip(iters, ref_iters)), (iters, ref_iters)
    ref_exe = REF_BIN / "tealeaf_dev_seq"
    if not ref_exe.exists():
        pytest.skip("reference dev_seq build absent: fields not compared")
    rdir = tmp_path / "ref"
    rdir.mkdir()
    run(ref_exe, rdir, fx["deck"], "tea.in", env={"OMP_NUM_THREADS": "1",
                                                   "OPS_REF_DUMP_DIR": str(rdir)})
    ref = read_dump(rdir)
    for k in TEA_SETUP:  # the reference build run here is the one the fixture came from
        if k in fx["dats"]:
            assert sha256(ref[k]["data"]) == fx["dats"][k]["sha256"], k
    for k in ("u", "energy1"):
        assert maxrel(got[k]["data"], ref[k]["data"]) < 1e-12, k
    rg, rr = read_reductions(dump), read_reductions(rdir)
    assert abs(rg["temp"][0] / rr["temp"][0] - 1.0) < 1e-12
