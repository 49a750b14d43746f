"""The hip_gfx950 translator on small applications (CPU only): call-site rewriting and
prototypes, API v2 (ACC<T>) and API v1 (pointer + OPS_ACC<n>) kernels, the errors it
reports, and that its output compiles for gfx950."""
import shutil
import subprocess
from pathlib import Path

import pytest

from ops_amd.translator import TranslateError, translate_app

ROOT = Path(__file__).resolve().parents[1]

KERNELS = """
#ifndef K_H
#define K_H
void k_v2(const ACC<double> &a, ACC<double> &b) { b(0, 0) = a(1, 0) + a(-1, 0); }
void k_v1(const double *a, double *b, const int *n) {
  b[OPS_ACC1(0, 0)] = a[OPS_ACC0(0, 1)] * n[0];
}
void k_v1_md(const double *v, double *out) {
  out[OPS_ACC1(0, 0)] = v[OPS_ACC_MD0(1, 0, -1)];
}
void k_red(const ACC<double> &a, double *m) { *m = *m > a(0, 0) ? *m : a(0, 0); }
#endif
"""

MAIN = """
#define OPS_2D
#include "ops_seq.h"
#include "k.h"

void run(ops_block blk, ops_dat a, ops_dat b, ops_dat v, ops_stencil s3, ops_stencil s1,
         ops_reduction r) {
  int range[] = {0, 8, 0, 8};
  int n = 3;
  ops_par_loop(k_v2, "k_v2", blk, 2, range,
               ops_arg_dat(a, 1, s3, "double", OPS_READ),
               ops_arg_dat(b, 1, s1, "double", OPS_WRITE));
  ops_par_loop(k_v1, "k_v1", blk, 2, range,
               ops_arg_dat(a, 1, s3, "double", OPS_READ),
               ops_arg_dat(b, 1, s1, "double", OPS_WRITE),
               ops_arg_gbl(&n, 1, "int", OPS_READ));
  ops_par_loop(k_v1_md, "k_v1_md", blk, 2, range,
               ops_arg_dat(v, 2, s3, "double", OPS_READ),
               ops_arg_dat(b, 1, s1, "double", OPS_WRITE));
  ops_par_loop(k_red, "k_red", blk, 2, range,
               ops_arg_dat(a, 1, s1, "double", OPS_READ),
               ops_arg_reduce(r, 1, "double", OPS_MAX));
}
"""


def make_app(tmp_path, kernels=KERNELS, main=MAIN):
    (tmp_path / "k.h").write_text(kernels)
    (tmp_path / "main.cpp").write_text(main)
    return str(tmp_path / "main.cpp")


def test_call_sites_and_prototypes(tmp_path):
    src = make_app(tmp_path)
    r = translate_app("tapp", [src], str(tmp_path / "out"))
    host = Path(r["host"][0]).read_text()
    for k in ("k_v2", "k_v1", "k_v1_md", "k_red"):
        assert f"ops_par_loop_{k}( \"{k}\"" in host         # kernel argument dropped
        assert f"void ops_par_loop_{k}(char const *, ops_block, int, int *" in host
    assert "ops_par_loop(k_" not in host
    # prototypes follow the ops_seq.h include
    assert host.index('#include "ops_seq.h"') < host.index("void ops_par_loop_k_v2(")


def test_v1_kernels_get_the_size_parameter(tmp_path):
    src = make_app(tmp_path)
    r = translate_app("tapp", [src], str(tmp_path / "out"))
    hip = Path(r["kernels"]).read_text()
    # pointer-only kernels take the accessor sizes by value; ACC kernels are unchanged
    assert "void k_v1(const double *a, double *b, const int *n, ops_v1_shape ops_v1)" in hip
    assert "void k_v1_md(const double *v, double *out, ops_v1_shape ops_v1)" in hip
    assert "void k_v2(const ACC<double> &a, ACC<double> &b) {" in hip
    assert "void k_red(const ACC<double> &a, double *m) {" in hip
    # the generated kernel hands over pointers and the sizes of each dat argument
    assert "v1s.xd[0] = p.xd0;" in hip and "v1s.md[0] = p.md0;" in hip
    assert "k_v1((double *)(q0 + " in hip and ", v1s);" in hip
    assert '#include "ops_macros.h"' in hip


def test_errors_name_the_problem(tmp_path):
    bad_count = MAIN.replace('ops_arg_gbl(&n, 1, "int", OPS_READ)',
                             'ops_arg_gbl(&n, 1, "int", OPS_READ), ops_arg_idx()')
    with pytest.raises(TranslateError, match="takes 3 parameters but the loop passes 4"):
        translate_app("tapp", [make_app(tmp_path, main=bad_count)], str(tmp_path / "o1"))
    missing = MAIN.replace("ops_par_loop(k_v2,", "ops_par_loop(k_nope,")
    with pytest.raises(TranslateError, match="kernel k_nope not found"):
        translate_app("tapp", [make_app(tmp_path, main=missing)], str(tmp_path / "o2"))
    gbl_write = MAIN.replace('ops_arg_gbl(&n, 1, "int", OPS_READ)',
                             'ops_arg_gbl(&n, 1, "int", OPS_WRITE)')
    with pytest.raises(TranslateError, match="ops_arg_gbl supports OPS_READ"):
        translate_app("tapp", [make_app(tmp_path, main=gbl_write)], str(tmp_path / "o3"))
    # an ACC kernel cannot take a dat through a pointer parameter
    mixed = KERNELS.replace("void k_v2(const ACC<double> &a, ACC<double> &b)",
                            "void k_v2(const ACC<double> &a, double *b)")
    with pytest.raises(TranslateError, match="must be ACC<T>&"):
        translate_app("tapp", [make_app(tmp_path, kernels=mixed)], str(tmp_path / "o4"))


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc not installed")
def test_output_compiles_for_gfx950(tmp_path):
    src = make_app(tmp_path)
    r = translate_app("tapp", [src], str(tmp_path / "out"))
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    cmd = [hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           f"-I{ROOT / 'include'}", f"-I{ROOT / 'ops_amd' / 'csrc' / 'runtime'}",
           f"-I{tmp_path}", "-c", r["kernels"], "-o", str(tmp_path / "k.o")]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout[-3000:]
    hcmd = ["g++", "-std=c++17", "-fPIC", f"-I{ROOT / 'include'}", f"-I{tmp_path}", "-c",
            r["host"][0], "-o", str(tmp_path / "h.o")]
    p = subprocess.run(hcmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout[-3000:]


def test_pointwise_arguments_and_streaming(tmp_path, monkeypatch):
    from ops_amd.translator import hip_gfx950 as hg
    from ops_amd.translator import cparse
    text = KERNELS + """
void k_pw(const ACC<double> &a, ACC<double> &b, const ACC<double> &c, ACC<double> &d) {
  b(0, 0) = 2.0 * a(0, 0) + c(1, 0);
  d(0, 0) = d(0, 0) + a(0, 0);
}
"""
    f = {fd.name: fd for fd in cparse.find_functions(text)}
    assert hg.pointwise_params(text, f["k_pw"]) == {0, 1, 3}     # c reads (1, 0)
    assert hg.pointwise_params(text, f["k_v1"]) == {1}           # a reads (0, 1); n is n[0]
    assert hg.pointwise_params(text, f["k_v1_md"]) == {1}        # v is a component access
    main = MAIN.replace("}\n", """  ops_par_loop(k_pw, "k_pw", blk, 2, range,
               ops_arg_dat(a, 1, s1, "double", OPS_READ),
               ops_arg_dat(b, 1, s1, "double", OPS_WRITE),
               ops_arg_dat(v, 1, s3, "double", OPS_READ),
               ops_arg_dat(a, 1, s1, "double", OPS_RW));
}
""", 1)
    src = make_app(tmp_path, kernels=text, main=main)
    monkeypatch.setattr(hg, "STREAM", True)
    hip = Path(translate_app("tapp", [src], str(tmp_path / "out"))["kernels"]).read_text()
    body = hip[hip.index("void ops_kern_k_pw("):hip.index("void ops_kern_k_pw_alias(")]
    # READ and RW point-wise arguments are loaded non-temporally and the RW one is stored
    # back; the WRITE argument and the stencil argument are left on the dat's memory
    assert "double v0 = __builtin_nontemporal_load(e0);" in body
    assert "double v3 = __builtin_nontemporal_load(e3);" in body and "*e3 = v3;" in body
    assert "v1 = " not in body and "v2 = " not in body
    alias = hip[hip.index("void ops_kern_k_pw_alias("):]
    assert "nontemporal" not in alias.split("__global__")[0]
    # the batch kernel (its launches never alias) streams like the no-alias one
    batch = hip[hip.index("void ops_kern_k_pw_batch("):]
    assert "double v0 = __builtin_nontemporal_load(e0);" in batch.split("__global__")[0]


def test_write_once_arguments_get_non_temporal_stores(tmp_path):
    from ops_amd.translator import hip_gfx950 as hg
    from ops_amd.translator import cparse
    text = KERNELS + """
void k_w(const ACC<double> &a, ACC<double> &b, ACC<double> &c, ACC<double> &d, ACC<double> &e) {
  double t = a(0, 0);
  b(0, 0) = 2.0 * t;
  if (t > 0.0) c(0, 0) = t;
  d(0, 0) = t;
  e(0, 0) = d(0, 0) + 1.0;
}
void k_ret(const ACC<double> &a, ACC<double> &b) {
  if (a(0, 0) < 0.0) return;
  b(0, 0) = a(0, 0);
}
"""
    f = {fd.name: fd for fd in cparse.find_functions(text)}
    # b: one top-level write; c: conditional; d: read after the write; e: stencil-free
    # single write whose right side reads another argument
    assert hg.written_once_params(text, f["k_w"]) == {1, 4}
    assert hg.written_once_params(text, f["k_ret"]) == set()
    assert hg.written_once_params(text, f["k_v1"]) == {1}        # b[OPS_ACC1(0, 0)] = ...
    main = MAIN.replace("}\n", """  ops_par_loop(k_w, "k_w", blk, 2, range,
               ops_arg_dat(a, 1, s1, "double", OPS_READ),
               ops_arg_dat(b, 1, s1, "double", OPS_WRITE),
               ops_arg_dat(v, 1, s1, "double", OPS_WRITE),
               ops_arg_dat(a, 1, s1, "double", OPS_RW),
               ops_arg_dat(v, 1, s1, "double", OPS_RW));
}
""", 1)
    src = make_app(tmp_path, kernels=text, main=main)
    hip = Path(translate_app("tapp", [src], str(tmp_path / "out"), ntstore=True)["kernels"]).read_text()
    body = hip[hip.index("void ops_kern_k_w("):hip.index("void ops_kern_k_w_alias(")]
    # only b (WRITE, written once) goes through a register and a non-temporal store; e is
    # written once but its access is RW, so it stays on the dat's memory
    assert "double v1;" in body and "__builtin_nontemporal_store(v1, e1);" in body
    assert body.count("__builtin_nontemporal_store") == 1
    alias = hip[hip.index("void ops_kern_k_w_alias("):].split("__global__")[0]
    assert "nontemporal" not in alias
    batch = hip[hip.index("void ops_kern_k_w_batch("):].split("__global__")[0]
    assert batch.count("__builtin_nontemporal_store(v1, e1);") == 1
    off = Path(translate_app("tapp", [src], str(tmp_path / "o2"), ntstore=False)["kernels"]).read_text()
    assert "nontemporal_store" not in off
