"""The C-ABI boundary (CPU only): every function that a public header declares inside an
extern "C" block is exported by the library that implements it, and the checkpointing
entry points behave as the reference's build without HDF5 does
(ops/c/src/externlib/ops_checkpointing.cpp:1734-1777)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "ops_amd" / "lib"

# header -> library that implements its extern "C" declarations
HEADERS = {
    ROOT / "include" / "ops_lib_core.h": "libops_hip.so",
    ROOT / "include" / "ops_hip_rt.h": "libops_hip.so",
    ROOT / "include" / "ops_tridiag.h": "libops_hip.so",
    ROOT / "include" / "ops_macros.h": "libops_hip.so",
    ROOT / "ops_amd" / "apps" / "tealeaf" / "tealeaf_unet.h": "libtealeaf_hip.so",
}


def c_declarations(text):
    """Names of the functions and variables declared in the extern "C" blocks of a header:
    comments and preprocessor lines are dropped, the blocks are cut out by brace matching,
    and each top-level statement that is not a typedef, a struct or an inline definition
    contributes the identifier before its first '(' (a function) or its last identifier
    (an extern variable)."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    names = set()
    for m in re.finditer(r'extern\s+"C"\s*\{', text):
        depth, i, start = 1, m.end(), m.end()
        while depth:
            depth += {"{": 1, "}": -1}.get(text[i], 0)
            i += 1
        body = text[start:i - 1]
        # drop nested brace groups (struct/enum bodies, inline function bodies)
        flat, depth = [], 0
        for ch in body:
            if ch == "{":
                depth += 1
                flat.append(" {} ")
            elif ch == "}":
                depth -= 1
            elif depth == 0:
                flat.append(ch)
        for stmt in "".join(flat).split(";"):
            s = " ".join(stmt.split())
            if not s or s.startswith("typedef") or "{}" in s and "(" not in s.split("{}")[0]:
                continue
            if s.startswith(("struct ", "enum ", "union ")) and "(" not in s:
                continue
            if "(" in s:
                if "{}" in s:
                    continue  # inline definition
                m2 = re.search(r"([A-Za-z_]\w*)\s*\($", s[:s.index("(") + 1])
                if m2:
                    names.add(m2.group(1))
            elif s.startswith("extern "):
                names.add(re.findall(r"[A-Za-z_]\w*", s)[-1])
    return names


def exported(libname):
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB / libname)],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if len(l.split()) >= 3}


@pytest.fixture(scope="module")
def built():
    if not (LIB / "libops_hip.so").exists():
        import __graft_entry__
        __graft_entry__.build()
    return True


def test_parser_finds_known_declarations():
    names = c_declarations((ROOT / "include" / "ops_lib_core.h").read_text())
    for n in ("ops_init", "ops_decl_dat_char", "ops_halo_transfer", "ops_reduction_result_char",
              "ops_checkpointing_init", "ops_get_last_error"):
        assert n in names
    assert "ops_block_core" not in names and "OPS_CHECKPOINT_FASTFW" not in names
    assert "ops_tridMultiDimBatch" in c_declarations((ROOT / "include" / "ops_tridiag.h").read_text())


@pytest.mark.parametrize("header", sorted(HEADERS), ids=lambda h: h.name)
def test_every_declared_symbol_is_exported(built, header):
    names = c_declarations(header.read_text())
    assert names, f"no extern \"C\" declarations found in {header.name}"
    missing = sorted(names - exported(HEADERS[header]))
    assert not missing, f"{HEADERS[header]} does not export {missing} declared in {header.name}"


def test_checkpointing_is_inert_like_the_reference_without_hdf5(built):
    L = ctypes.CDLL(str(LIB / "libops_hip.so"))
    for f in ("ops_checkpointing_init", "ops_checkpointing_before",
              "ops_checkpointing_name_before", "ops_checkpointing_fastfw",
              "ops_checkpointing_manual_datlist_fastfw",
              "ops_checkpointing_manual_datlist_fastfw_trigger"):
        getattr(L, f).restype = ctypes.c_bool
    # init: not restarting, for every option combination the reference app uses
    assert L.ops_checkpointing_init(b"check.h5", ctypes.c_double(15.0), 1 | 2 | 4) is False
    assert L.ops_checkpointing_init(b"check.h5", ctypes.c_double(15.0), 0) is False
    # every loop runs
    rng = (ctypes.c_int * 4)(0, 4, 0, 4)
    assert L.ops_checkpointing_before(None, 0, rng, 7) is True
    assert L.ops_checkpointing_name_before(None, 0, rng, b"loop") is True
    # no checkpoint is ever taken, so the fast-forward calls report false
    payload = ctypes.create_string_buffer(16)
    assert L.ops_checkpointing_fastfw(16, payload) is False
    assert L.ops_checkpointing_manual_datlist_fastfw(0, None, 16, payload) is False
    assert L.ops_checkpointing_manual_datlist_fastfw_trigger(0, None, 16, payload) is False
    L.ops_checkpointing_initphase_done()
    L.ops_checkpointing_manual_datlist(0, None)
    L.ops_checkpointing_exit(None)
    L.ops_checkpointing_reduction(None)
