"""Builds the reference's own OPS applications, sources unchanged, against this backend.

    python -m ops_amd.dropin [--ref DIR] [--out DIR] [app ...]

For each application below, the sources named by the application's own Makefile
(OPS_FILES and OTHER_FILES, e.g. apps/c/CloverLeaf/Makefile:19-27) are copied from the
reference tree into a scratch directory outside this repository, the OPS_FILES are
translated with the hip_gfx950 target (the step the reference's translator performs,
ops_translator/ops-translator/cpp/translator/program.py:20-40), the kernels are compiled
for gfx950 and the host files with g++ against include/, and the objects are linked with
libops_hip.so. Only the executable is kept: <out>/<app>/<app> (default out:
ops_amd/_dropin/), found at run time next to ops_amd/lib/ through its rpath.

No source file is edited. The build needs the reference tree (present where the package
is built, absent on machines that only run what was built).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import shutil
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from ops_amd import build as B  # noqa: E402

REF = Path(os.environ.get("OPS_REFERENCE_DIR", "/root/reference"))
OUT = B.PKG / "_dropin"

# application -> its directory under the reference's apps/c
APPS = {
    "laplace2d_step7": "laplace2d_tutorial/step7",
    "cloverleaf": "CloverLeaf",
    "cloverleaf3d": "CloverLeaf_3D",
    "tealeaf": "TeaLeaf",
    "adi": "adi",
}


def makefile_lists(app_dir: Path):
    """OPS_FILES and OTHER_FILES of the application's Makefile (continuation lines
    joined)."""
    text = (app_dir / "Makefile").read_text().replace("\\\n", " ")
    out = {}
    for var in ("OPS_FILES", "OTHER_FILES"):
        m = re.search(rf"(?m)^\s*{var}\s*=([^\n]*)$", text)
        out[var] = m.group(1).split() if m else []
    return out["OPS_FILES"], out["OTHER_FILES"]


def build_one(app: str, ref: Path = REF, out: Path = OUT, jobs: int = 8) -> Path:
    """Translate, compile and link one reference application; returns the executable."""
    from ops_amd.translator import translate_app

    src_dir = ref / "apps" / "c" / APPS[app]
    if not src_dir.is_dir():
        raise FileNotFoundError(f"{src_dir}: reference application not found")
    ops_files, other_files = makefile_lists(src_dir)
    dest = Path(out) / app
    dest.mkdir(parents=True, exist_ok=True)
    with tempfile.TemporaryDirectory(prefix=f"dropin_{app}_") as t:
        work = Path(t) / "src"
        shutil.copytree(src_dir, work)
        gen = Path(t) / "gen"
        res = translate_app(app, [str(work / f) for f in ops_files], str(gen), str(work))
        objs, futs = [], []
        with cf.ThreadPoolExecutor(jobs) as ex:
            k = Path(res["kernels"])
            ko = Path(t) / (k.stem + ".o")
            objs.append(ko)
            futs.append(ex.submit(B._compile_hip, k, ko, [f"-I{work}"]))
            for h in [*map(Path, res["host"]), *(work / f for f in other_files)]:
                ho = Path(t) / (h.stem + ".o")
                objs.append(ho)
                futs.append(ex.submit(B._compile_cxx, h, ho, [f"-I{work}"]))
            for f in futs:
                f.result()
        exe = dest / app
        tmp = dest / (app + ".tmp")
        rpath = os.path.relpath(B.LIB, dest)
        B._run([B.HIPCC, "--offload-arch=" + B.ARCH, *map(str, objs), f"-L{B.LIB}", "-lops_hip",
                f"-Wl,-rpath,$ORIGIN/{rpath}", "-o", str(tmp)])
        os.replace(tmp, exe)
    return exe


def build_all(apps=None, ref: Path = REF, out: Path = OUT, jobs: int = 8):
    return [build_one(a, ref, out, jobs) for a in (apps or APPS)]


def main(argv=None) -> None:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--ref", default=str(REF), help="reference tree (default %(default)s)")
    p.add_argument("--out", default=str(OUT), help="output directory (default %(default)s)")
    p.add_argument("apps", nargs="*", help=f"applications among {', '.join(APPS)} (default: all)")
    a = p.parse_args(argv)
    for name in a.apps:
        if name not in APPS:
            p.error(f"unknown application {name!r}")
    for exe in build_all(a.apps or None, Path(a.ref), Path(a.out)):
        print(exe)


if __name__ == "__main__":
    main()
