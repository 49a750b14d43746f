"""Readers for full-precision dat dumps (index.txt + <name>.bin) written at ops_exit by
the reference dev_seq builds (oracle/ref_dump.cpp), the oracle host backend and
libops_hip.so (OPS_DUMP_DIR=...)."""
from pathlib import Path

import numpy as np


def read_dump(d):
    d = Path(d)
    out = {}
    for line in (d / "index.txt").read_text().splitlines():
        f = line.split()
        name = f[0]
        dim, tsize, nd = int(f[1]), int(f[2]), int(f[3])
        size = [int(x) for x in f[4:9]]
        xpad = int(f[24])
        dt = {8: np.float64, 4: np.float32}.get(tsize, np.float64)
        if name in ("xx", "yy", "zz"):
            dt = np.int32
        raw = np.fromfile(d / f"{name}.bin", dtype=dt)
        shape = tuple(reversed(size[:nd]))
        if dim > 1:
            shape = shape + (dim,)
        arr = raw.reshape(shape)
        if xpad:
            arr = arr[..., : size[0] - xpad] if dim == 1 else arr[..., : size[0] - xpad, :]
        out[name] = {"dim": dim, "size": size, "d_m": [int(x) for x in f[9:14]],
                     "d_p": [int(x) for x in f[14:19]], "x_pad": xpad, "data": arr}
    return out


def read_reductions(d):
    p = Path(d) / "reductions.txt"
    out = {}
    if not p.exists():
        return out
    for line in p.read_text().splitlines():
        f = line.split()
        size = int(f[2])
        raw = bytes(int(x, 16) for x in f[3:3 + size])
        out[f[0]] = np.frombuffer(raw, dtype=np.float64 if f[1] == "double" else np.int32)
    return out
