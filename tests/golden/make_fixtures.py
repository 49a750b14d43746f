"""Writes the golden fixtures under tests/golden/ from the reference's own dev_seq
builds (oracle/_ref, made by `make -f oracle/Makefile.ref`). Run in the build container,
where /root/reference exists; the fixtures are committed, the binaries are not.

    python tests/golden/make_fixtures.py [laplace] [clover] [tea] [clover3d]

Each fixture holds inputs (decks / parameters) and outputs: scalar results printed by the
reference and, for every dat the run declared, its shape and the SHA-256 of its exact
bytes plus a few float64 statistics, so that a bitwise match can be checked without
storing full fields.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
REF = ROOT / "oracle" / "_ref"
OUT = Path(__file__).resolve().parent


def read_dump(d):
    dats = {}
    for line in (Path(d) / "index.txt").read_text().splitlines():
        f = line.split()
        name = f[0]
        dim, tsize, nd = int(f[1]), int(f[2]), int(f[3])
        size = [int(x) for x in f[4:9]]
        d_m = [int(x) for x in f[9:14]]
        d_p = [int(x) for x in f[14:19]]
        base = [int(x) for x in f[19:24]]
        raw = (Path(d) / f"{name}.bin").read_bytes()
        arr = np.frombuffer(raw, dtype=np.float64 if tsize == 8 else np.float32)
        dats[name] = {"dim": dim, "type_size": tsize, "ndims": nd, "size": size, "d_m": d_m,
                      "d_p": d_p, "base": base, "x_pad": int(f[24]),
                      "base_offset": int(f[25]),
                      "sha256": hashlib.sha256(raw).hexdigest(),
                      "sum": float(arr.astype(np.float64).sum()),
                      "sumsq": float((arr.astype(np.float64) ** 2).sum()),
                      "min": float(arr.min()), "max": float(arr.max())}
    return dats


def read_reductions(d):
    out = {}
    p = Path(d) / "reductions.txt"
    if not p.exists():
        return out
    for line in p.read_text().splitlines():
        f = line.split()
        name, typ, size = f[0], f[1], int(f[2])
        raw = bytes(int(x, 16) for x in f[3:3 + size])
        vals = np.frombuffer(raw, dtype=np.float64 if typ == "double" else np.int32).tolist()
        out[name] = {"type": typ, "hex": raw.hex(), "values": vals}
    return out


def run_ref(binary, cwd, args=(), env_extra=None, deck=None, deck_name=None):
    env = dict(os.environ, OMP_NUM_THREADS="1", OPS_REF_DUMP_DIR=str(cwd))
    env.update(env_extra or {})
    if deck is not None:
        (Path(cwd) / deck_name).write_text(deck)
    r = subprocess.run([str(REF / binary), *args], cwd=cwd, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, check=True)
    return r.stdout


def laplace():
    with tempfile.TemporaryDirectory() as t:
        out = run_ref("laplace2d_step7_dev_seq", t)
        dats = read_dump(t)
        reds = read_reductions(t)
    err_line = [l for l in out.splitlines() if l.strip().startswith("100,")]
    fx = {"source": "apps/c/laplace2d_tutorial/step7 (reference dev_seq)",
          "nx": 4094, "ny": 4094, "iters": 100, "mode": 1,
          "golden_error": 2.421354960840227e-03, "printed": err_line,
          "reductions": reds, "dats": dats}
    (OUT / "laplace_step7.json").write_text(json.dumps(fx, indent=1))




CLOVER_DECKS = {
    # small rectangle problem stopped by end_time after 14 steps
    "clover_small": """*clover
 state 1 density=0.2 energy 1.0
 state 2 density=1.0 energy=2.5 geometry=rectangle xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0
 x_cells=48
 y_cells=40
 xmin=0.0
 ymin=0.0
 xmax=10.0
 ymax=10.0
 initial_timestep=0.04
 timestep_rise=1.5
 max_timestep=0.04
 end_time=0.55
 end_step=25
 summary_frequency=0
 test_problem 0
*endclover
""",
    # apps/c/CloverLeaf/clover.in_default verbatim (known answer ke 1.82280367310258)
    "clover_default": None,
    # apps/c/CloverLeaf/clover_bm_short.in verbatim (known answer ke 1.19316898756307)
    "clover_bm_short": None,
    # apps/c/CloverLeaf/clover_bm16_short.in verbatim (3840^2, 87 steps, ke 0.307475452287895)
    "clover_bm16_short": None,
}
CLOVER_FILES = {"clover_default": "clover.in_default", "clover_bm_short": "clover_bm_short.in",
                "clover_bm16_short": "clover_bm16_short.in"}


def _dump_fixture(dats, reds):
    out = {}
    for k, v in dats.items():
        arr = v["data"]
        out[k] = {"shape": list(arr.shape), "dtype": str(arr.dtype),
                  "sha256": hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest(),
                  "sum": float(np.asarray(arr, dtype=np.float64).sum())}
    return {"dats": out, "reductions": {k: {"hex": v.tobytes().hex(), "values": v.tolist()}
                                        for k, v in reds.items()}}


def clover(which=None):
    sys.path.insert(0, str(ROOT / "tests"))
    from dumps import read_dump, read_reductions as rr
    names = [which] if which else ["clover_small", "clover_default", "clover_bm_short"]
    for name in names:
        deck = CLOVER_DECKS[name]
        if deck is None:
            deck = Path("/root/reference/apps/c/CloverLeaf", CLOVER_FILES[name]).read_text()
        with tempfile.TemporaryDirectory() as t:
            out = run_ref("cloverleaf_dev_seq", t, deck=deck, deck_name="clover.in")
            fx = _dump_fixture(read_dump(t), rr(t))
            steps = (Path(t) / "clover.out").read_text().count(" Step ")
        fx.update({"source": "apps/c/CloverLeaf (reference dev_seq)", "deck": deck,
                   "steps": steps, "tail": out.splitlines()[-6:]})
        (OUT / f"{name}.json").write_text(json.dumps(fx, indent=1))
        print("wrote", name, steps, "steps")


def clover_big():
    clover("clover_bm16_short")


def tea(*which):
    sys.path.insert(0, str(ROOT / "tests"))
    sys.path.insert(0, str(ROOT))
    from dumps import read_dump, read_reductions as rr
    from ops_amd.apps.tealeaf.decks import tea_deck
    decks = {
        "tea_cg_64": tea_deck(64, 3),
        "tea_cg_200": tea_deck(200, 2),
        # apps/c/TeaLeaf/tea.in verbatim: Jacobi, 10x10, 10 steps (temp 157.55084183279294)
        "tea_jacobi_ref": Path("/root/reference/apps/c/TeaLeaf/tea.in").read_text(),
        # the reference's other solvers (apps/c/TeaLeaf/tea_solve.cpp:200-300): Chebyshev
        # after CG presteps, PPCG, and CG with the diagonal Jacobi preconditioner (its block
        # Jacobi preconditioner stops with "block solvers are not supported")
        "tea_cheby_64": tea_deck(64, 3, solver="tl_use_chebyshev"),
        "tea_ppcg_64": tea_deck(64, 3, solver="tl_use_ppcg"),
        "tea_cg_jacdiag_64": tea_deck(64, 3, extra=("tl_preconditioner_type=jac_diag",)),
    }
    only = set(which or decks)
    for name, deck in decks.items():
        if name not in only:
            continue
        with tempfile.TemporaryDirectory() as t:
            out = run_ref("tealeaf_dev_seq", t, deck=deck, deck_name="tea.in")
            fx = _dump_fixture(read_dump(t), rr(t))
            text = (Path(t) / "tea.out").read_text()
        # CG, Jacobi and Chebyshev print "Iteration count <n>"; PPCG prints its CG presteps
        # that way and then "PPCG Iteration count <n> (Total <m>)"
        iters = [int(l.split()[-1].rstrip(")")) for l in text.splitlines()
                 if "Iteration count" in l and "PPCG" not in l]
        ppcg = [int(l.split()[3]) for l in text.splitlines() if "PPCG Iteration count" in l]
        fx.update({"source": "apps/c/TeaLeaf (reference dev_seq)", "deck": deck,
                   "iteration_counts_printed": iters, "tail": out.splitlines()[-4:]})
        if ppcg:
            fx["ppcg_iteration_counts_printed"] = ppcg
        (OUT / f"{name}.json").write_text(json.dumps(fx, indent=1))
        print("wrote", name, iters)


CLOVER3D_DECKS = {
    # apps/c/CloverLeaf_3D/clover.in with a 20x18x16 mesh and 10 steps
    "clover3d_small": ("clover.in", {"x_cells=96": "x_cells=20", "y_cells=96": "y_cells=18",
                                     "z_cells=96": "z_cells=16", "end_step=87": "end_step=10"}),
    # four states: two cuboids (one moving), a point state; 24x20x22, 9 steps, dt rising
    "clover3d_states": ("""*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0 zmin=0.0 zmax=4.0
 state 3 density=0.5 energy=1.5 xvel=0.3 yvel=-0.2 zvel=0.1 geometry=cuboid xmin=6.0 xmax=9.0 ymin=1.0 ymax=4.0 zmin=2.0 zmax=8.0
 state 4 density=2.0 energy=3.0 geometry=point xmin=5.0 ymin=5.0 zmin=5.0
 x_cells=24
 y_cells=20
 z_cells=22
 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax=10.0
 zmax=10.0
 initial_timestep=0.04
 max_timestep=0.04
 timestep_rise=1.5
 end_step=9
 end_time=100.0
 test_problem 0
 checkpoint_frequency=10
*endclover
""", None),
    # apps/c/CloverLeaf_3D/clover.in verbatim: 96^3, 87 steps (known answer ke 20.0546870878964)
    "clover3d_bm": ("clover.in", {}),
}


def clover3d(which=None):
    sys.path.insert(0, str(ROOT / "tests"))
    from dumps import read_dump, read_reductions as rr
    names = [which] if which else list(CLOVER3D_DECKS)
    for name in names:
        src, edits = CLOVER3D_DECKS[name]
        if edits is None:
            deck = src
        else:
            deck = Path("/root/reference/apps/c/CloverLeaf_3D", src).read_text()
            for a, b in edits.items():
                assert a in deck
                deck = deck.replace(a, b)
        with tempfile.TemporaryDirectory() as t:
            out = run_ref("cloverleaf3d_dev_seq", t, deck=deck, deck_name="clover.in")
            fx = _dump_fixture(read_dump(t), rr(t))
            steps = (Path(t) / "clover.out").read_text().count(" Step ")
        fx.update({"source": "apps/c/CloverLeaf_3D (reference dev_seq)", "deck": deck,
                   "steps": steps, "tail": out.splitlines()[-6:]})
        (OUT / f"{name}.json").write_text(json.dumps(fx, indent=1))
        print("wrote", name, steps, "steps")


if __name__ == "__main__":
    which = sys.argv[1:] or ["laplace", "clover"]
    for w in which:
        # "tea:tea_cheby_64,tea_ppcg_64" regenerates only the named fixtures of a group
        group, _, names = w.partition(":")
        globals()[group](*[n for n in names.split(",") if n])
        print("done", w)
