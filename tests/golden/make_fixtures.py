"""Writes the golden fixtures under tests/golden/ from the reference's own dev_seq
builds (oracle/_ref, made by `make -f oracle/Makefile.ref`). Run in the build container,
where /root/reference exists; the fixtures are committed, the binaries are not.

    python tests/golden/make_fixtures.py [laplace] [clover] [tea]

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


if __name__ == "__main__":
    which = sys.argv[1:] or ["laplace"]
    for w in which:
        globals()[w]()
        print("wrote", w)
