"""Build libops_hip.so and the translated applications, in tree, for gfx950.

    python -m ops_amd.build            # runtime + every application
    python -m ops_amd.build laplace2d  # runtime + one application

Outputs go to ops_amd/lib/. Generated sources go to ops_amd/apps/<app>/_gen/.
Compilation runs hipcc directly (no CMake): device code with --offload-arch=gfx950,
host-only application code with g++.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "ops_amd"
LIB = PKG / "lib"
INC = ROOT / "include"
RT = PKG / "csrc" / "runtime"
APPS = PKG / "apps"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OPS_OFFLOAD_ARCH", "gfx950")
# -ffp-contract=off keeps results bit-comparable with the reference's IEEE CPU build;
# packed FP32 instructions are disabled on purpose.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value"]
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off"]

# application name -> (sources relative to the app dir, kernel-extra flags)
APP_SOURCES = {
    "laplace2d": ["laplace2d.cpp"],
    "apitest": ["apitest.cpp"],
    "cloverleaf": ["clover.cpp"],
    "tealeaf": ["tealeaf.cpp"],
    "adi": ["adi.cpp"],
    "mgrid": ["mgrid.cpp"],
    "lowdim": ["lowdim.cpp"],
    "multicomp": ["multicomp.cpp"],
    "cloverleaf3d": ["clover3d.cpp"],
}
# native sources one application shares with another: application -> [(app dir, file)]
APP_SHARED_NATIVE = {
    "cloverleaf3d": [("cloverleaf", "native/clover_dt.hip")],
}
# translator options per application. ADI's preproc writes ten coefficient and right-hand-
# side arrays of one layout: with x tiles on 128-byte lines of their rows and non-temporal
# stores its 48-block step measured 11.0 ms, against 11.8 with neither and 11.6 with aligned
# tiles alone (plain stores into split lines had measured slower than plain stores, 18.1
# against 17.4 ms with the round-one solver); CloverLeaf keeps unaligned tiles (7 % slower
# aligned, its dats have several row pitches)
APP_TRANSLATE = {
    "adi": {"ntstore": True, "align": True},
}
# builds of an application's sources with extra defines: name -> (application, defines)
APP_VARIANTS = {
    "multicomp_soa": ("multicomp", ["-DOPS_SOA"]),
}


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{out}")
    return out


def _compile_hip(src: Path, obj: Path, extra=()):
    _run([HIPCC, *HIPFLAGS, f"-I{INC}", f"-I{RT}", *extra, "-c", str(src), "-o", str(obj)])


def _compile_cxx(src: Path, obj: Path, extra=()):
    _run(["g++", *CXXFLAGS, f"-I{INC}", *extra, "-c", str(src), "-o", str(obj)])


def build_runtime(jobs: int = 8) -> Path:
    obj = PKG / "_build" / "runtime"
    obj.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    srcs = [RT / "ops_core.cpp", RT / "ops_device.hip", RT / "ops_tridiag.hip"]
    extra_src = sorted((PKG / "csrc" / "solvers").glob("*.hip")) if (PKG / "csrc" / "solvers").exists() else []
    srcs += extra_src
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = []
        for s in srcs:
            o = obj / (s.stem + ".o")
            objs.append(o)
            if s.suffix == ".cpp":
                futs.append(ex.submit(_run, [HIPCC, *HIPFLAGS, "-x", "hip", f"-I{INC}", f"-I{RT}",
                                             "-c", str(s), "-o", str(o)]))
            else:
                futs.append(ex.submit(_compile_hip, s, o))
        for f in futs:
            f.result()
    out = LIB / "libops_hip.so"
    tmp = LIB / "libops_hip.so.tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_app(name: str, jobs: int = 8) -> Path:
    sys.path.insert(0, str(ROOT))
    from ops_amd.translator import translate_app
    base, defines = APP_VARIANTS.get(name, (name, []))
    app_dir = APPS / base
    srcs = [str(app_dir / s) for s in APP_SOURCES[base]]
    gen = app_dir / ("_gen" if base == name else f"_gen_{name}")
    res = translate_app(name, srcs, str(gen), str(app_dir), **APP_TRANSLATE.get(base, {}))
    obj = PKG / "_build" / name
    obj.mkdir(parents=True, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = []
        k = Path(res["kernels"])
        ko = obj / (k.stem + ".o")
        objs.append(ko)
        futs.append(ex.submit(_compile_hip, k, ko, [f"-I{app_dir}", *defines]))
        for h in res["host"]:
            h = Path(h)
            ho = obj / (h.stem + ".o")
            objs.append(ho)
            futs.append(ex.submit(_compile_cxx, h, ho, [f"-I{app_dir}", "-DOPS_APP_LIB", *defines]))
        extra = app_dir / "native"
        natives = sorted(extra.glob("*.hip")) if extra.exists() else []
        natives += [APPS / a / f for a, f in APP_SHARED_NATIVE.get(base, [])]
        if natives:
            for s in natives:
                so = obj / (s.stem + ".o")
                objs.append(so)
                futs.append(ex.submit(_compile_hip, s, so, [f"-I{app_dir}", *defines]))
        for f in futs:
            f.result()
    out = LIB / f"lib{name}_hip.so"
    tmp = LIB / f"lib{name}_hip.so.tmp"
    # -Bsymbolic: an application's calls to its own stubs bind inside the library, whatever
    # else is loaded (builds of one application's sources share stub names)
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *map(str, objs), f"-L{LIB}",
          "-lops_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-Bsymbolic", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def available_apps():
    apps = [a for a in APP_SOURCES if all((APPS / a / s).exists() for s in APP_SOURCES[a])]
    return apps + [v for v, (a, _) in APP_VARIANTS.items() if a in apps]


def build_all(apps=None, jobs: int = 8):
    build_runtime(jobs)
    built = []
    for a in (apps or available_apps()):
        built.append(build_app(a, jobs))
    return built


if __name__ == "__main__":
    names = sys.argv[1:] or None
    for p in build_all(names):
        print(p)
