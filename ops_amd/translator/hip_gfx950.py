"""`hip_gfx950` code-generation target of the OPS translator.

For an application it emits
  * `<src>_ops.cpp` per source file: the program with every `ops_par_loop(kernel, ...)`
    rewritten to `ops_par_loop_<kernel>(...)` and the prototypes inserted after the OPS
    include (the reference's program rewrite: ops_translator/ops-translator/cpp/
    translator/program.py:10-48);
  * `<app>_kernels.hip`: one translation unit holding every loop of the application as a
    gfx950 kernel plus its host stub, the `ops_decl_const` symbols, and the user kernel
    headers compiled for the device.

Kernel shape (one thread per iteration point): 2D/3D workgroups of 256 lanes with 64
lanes along x so that each wavefront reads whole row segments, an XCD-aware tile map,
ACC<T> views at the thread's point, wavefront butterfly + workgroup reductions finished
on the device, small ops_arg_gbl values passed by value in the kernel arguments.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple

from . import cparse

ACCESS = {"OPS_READ": 0, "OPS_WRITE": 1, "OPS_RW": 2, "OPS_INC": 3, "OPS_MIN": 4, "OPS_MAX": 5}
RED_OP = {"OPS_INC": "ops_red_inc", "OPS_MIN": "ops_red_min", "OPS_MAX": "ops_red_max"}
CTYPES = {"double": "double", "float": "float", "int": "int", "ll": "long long",
          "long long": "long long"}


class TranslateError(Exception):
    pass


@dataclass
class Arg:
    kind: str                 # dat | gbl | reduce | idx
    expr: str = ""            # dat / pointer / handle expression
    dim: str = "1"
    stencil: str = ""
    typ: str = ""
    acc: str = "OPS_READ"
    opt: bool = False
    prolong: bool = False     # stencil declared with ops_decl_prolong_stencil

    def key(self):
        return (self.kind, self.typ, self.acc if self.kind != "dat" else self.acc, self.dim
                if self.kind != "dat" else (self.dim if cparse.int_literal(self.dim) == 1 else "m"),
                self.prolong)


@dataclass
class Loop:
    kernel: str
    name_expr: str
    args: List[Arg]
    file: str
    line: int
    call: cparse.Call
    stub: str = ""


@dataclass
class Const:
    name: str
    dim: str
    typ: str


@dataclass
class Param:
    text: str
    is_acc: bool
    ctype: str
    const: bool


def parse_param(p: str) -> Param:
    s = " ".join(p.split())
    m = re.search(r"ACC\s*<\s*([^>]+?)\s*>", s)
    if m:
        return Param(s, True, m.group(1), s.startswith("const"))
    m = re.match(r"(const\s+)?([A-Za-z_][\w\s]*?)\s*\*", s)
    if m:
        return Param(s, False, m.group(2).strip(), bool(m.group(1)))
    raise TranslateError(f"unsupported kernel parameter '{p}' (expected ACC<T>& or T*)")


def parse_arg(text: str) -> Arg:
    masked = cparse.mask(text)
    calls = cparse.find_calls(text, masked, ["ops_arg_dat_opt", "ops_arg_dat", "ops_arg_gbl",
                                             "ops_arg_reduce", "ops_arg_idx"])
    if not calls or calls[0].start != len(text) - len(text.lstrip()):
        raise TranslateError(f"unrecognised ops_par_loop argument: {text}")
    c = calls[0]
    a = c.args
    if c.name == "ops_arg_idx":
        return Arg("idx")
    if c.name in ("ops_arg_dat", "ops_arg_dat_opt"):
        return Arg("dat", a[0], a[1], a[2], cparse.unquote(a[3]), a[4].strip(),
                   c.name == "ops_arg_dat_opt")
    if c.name == "ops_arg_gbl":
        acc = a[3].strip()
        if acc != "OPS_READ":
            raise TranslateError("ops_arg_gbl supports OPS_READ; use ops_arg_reduce for reductions")
        return Arg("gbl", a[0], a[1], "", cparse.unquote(a[2]), acc)
    return Arg("reduce", a[0], a[1], "", cparse.unquote(a[2]), a[3].strip())


def scan_source(path: str):
    text = Path(path).read_text()
    masked = cparse.mask(text)
    loops, consts = [], []
    for c in cparse.find_calls(text, masked, "ops_par_loop"):
        if len(c.args) < 5:
            raise TranslateError(f"{path}:{c.line}: ops_par_loop needs kernel, name, block, dim, range")
        loops.append(Loop(c.args[0], c.args[1], [parse_arg(x) for x in c.args[5:]], path,
                          c.line, c))
    for c in cparse.find_calls(text, masked, ["ops_decl_const", "ops_update_const"]):
        if c.name == "ops_decl_const":
            consts.append(Const(cparse.unquote(c.args[0]), c.args[1].strip(),
                                cparse.unquote(c.args[2])))
    m = re.search(r"(?m)^\s*#\s*define\s+OPS_([1-5])D\b", masked)
    ndim = int(m.group(1)) if m else None
    return text, loops, consts, ndim


def kernel_headers(app_dir: str, sources: List[str]) -> Dict[str, Tuple[str, cparse.FuncDef]]:
    """Map kernel function name -> (header path, definition) for every header in the
    application directory."""
    out = {}
    for p in sorted(Path(app_dir).glob("*.h")):
        text = p.read_text()
        for f in cparse.find_functions(text):
            out[f.name] = (str(p), f)
    return out


def strip_include_guard(text: str) -> str:
    """The header without its own include guard (`#ifndef G` / `#define G` before any other
    code, and the file's last `#endif`): headers are pasted into one translation unit, and
    two headers of an application may share a guard name (the reference's TeaLeaf
    tea_leaf_jacobi_kernels.h reuses tea_leaf_cheby_kernels.h's), which would drop the
    second one."""
    masked = cparse.mask(text)
    m = re.match(r"\s*#\s*ifndef\s+(\w+)\s*\n\s*#\s*define\s+(\w+)[^\n]*\n", masked)
    if not m or m.group(1) != m.group(2):
        return text
    ends = [e for e in re.finditer(r"(?m)^\s*#\s*endif\b[^\n]*$", masked)]
    if not ends or masked[ends[-1].end():].strip():
        return text
    e = ends[-1]
    # blank the guard lines, keeping every other character (and line) in place
    return (" " * 0 + "\n" * text[:m.end()].count("\n") + text[m.end():e.start()] +
            "\n" * text[e.start():e.end()].count("\n") + text[e.end():])


def is_v1_kernel(text: str, f: cparse.FuncDef) -> bool:
    """API v1 kernel: only pointer parameters, dats addressed with the OPS_ACC<n> macros
    (directly or through macros of the application's own)."""
    return not any(re.search(r"\bACC\s*<", p) for p in f.params)


def add_v1_param(text: str, f: cparse.FuncDef) -> str:
    """Append the by-value `ops_v1_shape ops_v1` parameter that carries the sizes behind
    the OPS_ACC<n> macros (it shadows the host-side global of ops_macros.h)."""
    head = text[:f.params_close]
    sep = ", " if f.params else ""
    if not f.params:
        # `void k()` or `void k(void)`: replace the (possibly void) list
        op = text.rindex("(", f.start, f.params_close)
        head = text[:op + 1]
    return head + sep + "ops_v1_shape ops_v1" + text[f.params_close:]


def pointwise_params(text: str, f: cparse.FuncDef) -> set:
    """Indices of the parameters that the kernel body touches only at offset 0: every
    occurrence of the name is an accessor call with all-zero offsets, `a(0, 0)` (ACC) or
    `a[OPS_ACC<n>(0, 0, 0)]` (API v1). Any other use (another offset, a component index, a
    method, passing the object on, use through a macro of the application) keeps the
    parameter off the list."""
    body = cparse.mask(text[f.body_open:f.end])
    out = set()
    for i, ptxt in enumerate(f.params):
        m = re.search(r"([A-Za-z_]\w*)\s*$", ptxt)
        if not m:
            continue
        name = m.group(1)
        occ = list(re.finditer(r"\b%s\b" % re.escape(name), body))
        ok = bool(occ)
        for mm in occ:
            rest = body[mm.end():].lstrip()
            if not (re.match(r"\(\s*0\s*(?:,\s*0\s*){0,2}\)", rest) or
                    re.match(r"\[\s*OPS_ACC\d+\s*\(\s*0\s*(?:,\s*0\s*){0,2}\)\s*\]", rest)):
                ok = False
                break
        if ok:
            out.add(i)
    return out


def written_once_params(text: str, f: cparse.FuncDef) -> set:
    """Indices of the parameters that the kernel body writes exactly once at the point
    itself and never reads: the name occurs once, as the left side of `a(0, 0) = ...` (ACC)
    or `a[OPS_ACC<n>(0, 0)] = ...` (API v1), and that assignment starts a statement at the
    top level of the body (not under if/else/for/while/switch), in a body without return
    or goto."""
    body = cparse.mask(text[f.body_open:f.end])
    if re.search(r"\b(return|goto)\b", body):
        return set()
    out = set()
    for i, ptxt in enumerate(f.params):
        m = re.search(r"([A-Za-z_]\w*)\s*$", ptxt)
        if not m:
            continue
        name = m.group(1)
        occ = list(re.finditer(r"\b%s\b" % re.escape(name), body))
        if len(occ) != 1:
            continue
        o = occ[0]
        rest = body[o.end():]
        if not (re.match(r"\s*\(\s*0\s*(?:,\s*0\s*){0,2}\)\s*=(?!=)", rest) or
                re.match(r"\s*\[\s*OPS_ACC\d+\s*\(\s*0\s*(?:,\s*0\s*){0,2}\)\s*\]\s*=(?!=)", rest)):
            continue
        before = body[:o.start()]
        if before.count("{") - before.count("}") != 1:
            continue  # inside a nested block
        # the statement begins right here: only whitespace since the last ; { or }
        last = max(before.rfind(";"), before.rfind("{"), before.rfind("}"))
        if before[last + 1:].strip():
            continue
        out.add(i)
    return out


def nt_stored(loop: Loop, i: int, written_once: set) -> bool:
    """A WRITE dat argument that the kernel assigns once at its own point and never reads
    (one component, no prolong addressing, always present): the generated kernel hands the
    kernel a register and writes it with a non-temporal store afterwards."""
    a = loop.args[i]
    return (a.kind == "dat" and i in written_once and a.acc == "OPS_WRITE" and
            not a.opt and not a.prolong and cparse.int_literal(a.dim) == 1)


def streamed(loop: Loop, i: int, pointwise: set) -> bool:
    """A dat argument read once per point and only at the point itself (READ or RW,
    one component, no prolong addressing, always present): the generated kernel loads it
    with a non-temporal load into a register and hands the kernel that register (RW: stored
    back with a plain store after the kernel)."""
    a = loop.args[i]
    return (a.kind == "dat" and i in pointwise and a.acc in ("OPS_READ", "OPS_RW") and
            not a.opt and not a.prolong and cparse.int_literal(a.dim) == 1)


# ------------------------------------------------------------------ code generation

def const_dim(a: Arg, loop: Loop) -> str:
    """Dimension of a gbl/reduce argument as a compile-time expression: an integer
    literal or a macro / constant the kernels translation unit can see."""
    d = a.dim.strip()
    if cparse.int_literal(d) is not None or re.fullmatch(r"[A-Za-z_]\w*", d):
        return d
    raise TranslateError(f"{loop.file}:{loop.line}: argument dim '{d}' must be a literal or a "
                         "compile-time constant name")


# loop points per lane along x (blockDim.x apart)
VX = int(os.environ.get("OPS_GEN_VX", "1"))
# loops whose only reductions are MIN/MAX (order-free, so results are unchanged) take this
# many points per lane: the per-workgroup reduction tree is then paid once per 4 x 256 points
RED_VX = int(os.environ.get("OPS_GEN_RED_VX", "4"))


# non-temporal loads of point-wise arguments, opt-in (OPS_GEN_STREAM=1): on gfx950 they
# lift a plain 5-read/2-write stream from 5.9 to 6.7 TB/s, but in the stencil kernels,
# whose other arguments share L2 with them, the CloverLeaf steps gained under 1 % (2D)
# or lost up to 8 % (3D at 96^3)
STREAM = os.environ.get("OPS_GEN_STREAM", "0") == "1"
# non-temporal stores of the arguments a kernel writes once at its own point and never
# reads (OPS_GEN_NTSTORE=0 turns them off)
NTSTORE = os.environ.get("OPS_GEN_NTSTORE", "1") == "1"


def loop_vx(loop: Loop) -> int:
    reds = [a for a in loop.args if a.kind == "reduce"]
    if reds and all(a.acc in ("OPS_MIN", "OPS_MAX") for a in reds):
        return max(VX, RED_VX)
    return VX


# x tiles that start on 128-byte lines of a written argument's rows (OPS_GEN_ALIGN=1, or the
# per-application option): whole-line stores for kernels that write many arrays of one
# layout (ADI's preproc: 12.2 -> 11.6 ms per step, more with non-temporal stores). Off by
# default: when a loop's dats have different row pitches the other arguments stay split,
# and the CloverLeaf 2D step measured 7 % slower with them
ALIGN = os.environ.get("OPS_GEN_ALIGN", "0") == "1"
_align_on = ALIGN  # the current application's setting (translate_app)


def align_anchor(loop: Loop) -> int:
    """The dat argument whose rows the x tiles of a launch start on 128-byte lines of (the
    first written one-component dat; -1 for loops with reductions, several points per lane
    or no such argument, and with aligned tiles off). The runtime decides per launch
    (ops_hip_loop.align)."""
    if not _align_on or any(a.kind == "reduce" for a in loop.args) or loop_vx(loop) != 1:
        return -1
    for i, a in enumerate(loop.args):
        if (a.kind == "dat" and a.acc in ("OPS_WRITE", "OPS_RW", "OPS_INC") and not a.prolong and
                cparse.int_literal(a.dim) == 1):
            return i
    return -1


def gen_kernel(loop: Loop, params: List[Param], ndim: int, v1: bool = False,
               pointwise: Optional[set] = None, written_once: Optional[set] = None) -> str:
    k = loop.stub
    L = []
    # kernel-argument struct
    L.append(f"struct ops_kp_{k} {{")
    L.append("  int cnt0, cnt1, cnt2, st0, st1, st2;")
    if any(a.kind == "reduce" for a in loop.args):
        L.append("  unsigned ntiles, tg0, tg1;  // tile count and tile grid (grid-stride walk)")
    if align_anchor(loop) >= 0:
        L.append("  int al;  // elements per 128-byte line of the anchor argument (aligned x tiles), or 0")
    if not any(a.kind == "reduce" for a in loop.args):
        L.append("  int rev;  // walk each XCD's run of tiles from its end (every other launch)")
    for i, a in enumerate(loop.args):
        if a.kind == "dat":
            t = params[i].ctype
            L.append(f"  {t} *p{i}; int sx{i}, sy{i}, sz{i}, xd{i}, yd{i}, zd{i}, md{i};")
            if a.prolong:
                L.append(f"  int mx{i}, my{i}, mz{i};")
        elif a.kind == "gbl":
            L.append(f"  ops_gbl<{CTYPES.get(a.typ, a.typ)}, {const_dim(a, loop)}> g{i};")
        elif a.kind == "reduce":
            t = CTYPES.get(a.typ, a.typ)
            L.append(f"  {t} *rp{i};")
    L.append("};")
    L.append("")
    if batchable(loop):
        # a batch of up to 8 launches travels by value as the kernel argument (kept under
        # 4 KB: fewer loops per batch for large argument structs)
        L.append(f"constexpr int ops_nb_{k} = (int)(3000 / sizeof(ops_kp_{k})) < 8 ? "
                 f"(int)(3000 / sizeof(ops_kp_{k})) : 8;")
        L.append(f"struct ops_kb_{k} {{")
        L.append(f"  ops_kp_{k} p[ops_nb_{k} > 0 ? ops_nb_{k} : 1];")
        L.append("  unsigned first[9], nb[8], g0[8], g1[8], b0[8], b1[8], b2[8];")
        L.append("  int n;")
        L.append("};")
        L.append("")
    if noalias_variant(loop):
        # the same body twice: with every dat as a no-alias pointer parameter (the usual
        # launch: loads need not wait behind the stores of other arguments), and reading
        # the pointers from the struct (a dat passed to two arguments, one of them writing)
        # the no-alias variant also streams the point-wise arguments (non-temporal loads);
        # the aliasing variant keeps every access on the dat's memory
        L.append(kernel_body(loop, params, ndim, k, True, v1, pointwise or set(),
                             written_once or set()))
        L.append("")
        L.append(kernel_body(loop, params, ndim, k + "_alias", False, v1))
    else:
        L.append(kernel_body(loop, params, ndim, k, False, v1))
    if batchable(loop):
        L.append("")
        L.append(batch_kernel(loop, params, ndim, v1, pointwise or set(), written_once or set()))
    return "\n".join(L)


def batchable(loop: Loop) -> bool:
    """Loops that may join a launch batch: no reductions (their per-workgroup partials and
    finisher belong to one launch), no prolong addressing, one point per lane."""
    return (not any(a.kind == "reduce" for a in loop.args) and
            not any(a.kind == "dat" and a.prolong for a in loop.args) and loop_vx(loop) == 1)


def batch_kernel(loop: Loop, params: List[Param], ndim: int, v1: bool, pointwise: set = frozenset(),
                 written_once: set = frozenset()) -> str:
    """The batch variant: the aliasing-safe body (pointers from the argument struct) behind
    a prologue that finds the workgroup's loop in the batch (prefix sums of workgroups, each
    loop's share rounded up to a multiple of the 8 XCDs, so workgroup b of a loop runs on
    XCD b % 8), then its tile (the loop's XCD remap over its own tiles; the padding
    workgroups return) and the lane's position in that loop's workgroup shape (threads
    beyond it return; the body has no barriers)."""
    k = loop.stub
    # pointers from the argument struct; the streamed loads and non-temporal stores of the
    # no-alias kernel still apply (a launch whose dats alias does not join a batch)
    body = kernel_body(loop, params, ndim, k + "_batch", False, v1, pointwise, written_once,
                       stream_from_struct=True)
    sig = f"void ops_kern_{k}_batch(const ops_kp_{k} p) {{"
    assert sig in body
    body = body.replace(sig, f"void ops_kern_{k}_batch(const ops_kb_{k} B) {{")
    pro = ("  unsigned bx, by, bz;\n  ops_block_coords_dir(bx, by, bz, p.rev);\n"
           "  const unsigned tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;")
    assert pro in body
    body = body.replace(pro, "\n".join([
        "  int j = 0;",
        "  while (j + 1 < B.n && blockIdx.x >= B.first[j + 1]) j++;",
        f"  const ops_kp_{k} &p = B.p[j];",
        "  const unsigned bdx = B.b0[j], bdy = B.b1[j], bdz = B.b2[j];",
        "  const unsigned lb = blockIdx.x - B.first[j];",
        "  if (lb >= B.nb[j]) return;",
        "  const unsigned lt = ops_xcd_remap_dir(lb, B.nb[j], p.rev);",
        "  const unsigned bx = lt % B.g0[j], by = (lt / B.g0[j]) % B.g1[j], bz = lt / (B.g0[j] * B.g1[j]);",
        "  const unsigned tx = threadIdx.x % bdx, ty = (threadIdx.x / bdx) % bdy, tz = threadIdx.x / (bdx * bdy);",
        "  if (tz >= bdz) return;"]))
    for d in "xyz":
        body = body.replace(f"blockDim.{d}", f"bd{d}")
    assert "blockDim" not in body and "threadIdx.y" not in body
    return body


def noalias_variant(loop: Loop) -> bool:
    """a no-alias kernel variant pays when some dat is written and another dat is accessed"""
    dats = [a for a in loop.args if a.kind == "dat"]
    return len(dats) >= 2 and any(a.acc != "OPS_READ" for a in dats)


def dat_offset(i: int, ndim: int, aligned: bool = False) -> str:
    """Element offset of dat argument i at point (x, y, z): the workgroup origin's part in
    64 bits, then the lane's part (unsigned, 24-bit products along x and y). Dimensions
    beyond the application's are 0 inside the kernel's bounds test, so they are left out.
    aligned: the lane's x offset is txm = tx - mis, which can be negative: its product is a
    signed 24-bit multiply and the lane's part a signed sum (the runtime keeps the x step
    below 2^23 and the lane's part below 2^31)."""
    uni = [f"(long long)x0 * p.sx{i}"]
    if aligned:
        lane = [f"__mul24(txm, p.sx{i})"]
    else:
        lane = [f"(unsigned)__umul24(tx, (unsigned)p.sx{i})"]
    if ndim >= 2:
        uni.append(f"(long long)y0 * p.sy{i}")
        lane.append(f"(int)__umul24(ty, (unsigned)p.sy{i})" if aligned else f"(unsigned)__umul24(ty, (unsigned)p.sy{i})")
    if ndim >= 3:
        uni.append(f"(long long)z0 * p.sz{i}")
        lane.append(f"(int)(tz * (unsigned)p.sz{i})" if aligned else f"tz * (unsigned)p.sz{i}")
    return f"({' + '.join(uni)}) + ({' + '.join(lane)})"


def kernel_body(loop: Loop, params: List[Param], ndim: int, name: str, noalias: bool,
                v1: bool = False, pointwise: set = frozenset(),
                written_once: set = frozenset(), stream_from_struct: bool = False) -> str:
    """One generated kernel. noalias: dats are __restrict__ pointer parameters q<i>, and the
    point-wise arguments may be streamed (non-temporal loads / deferred non-temporal
    stores); stream_from_struct: pointers from the argument struct with the same streaming
    (the batch kernel: its launches never alias)."""
    k = loop.stub
    nt_ok = noalias or stream_from_struct
    L = []
    red = [i for i, a in enumerate(loop.args) if a.kind == "reduce"]
    extra = "".join(f", {params[i].ctype} *__restrict__ q{i}" for i, a in enumerate(loop.args)
                    if a.kind == "dat") if noalias else ""
    L.append(f"__global__ __launch_bounds__(256) void ops_kern_{name}(const ops_kp_{k} p{extra}) {{")
    for i in red:
        a = loop.args[i]
        t = CTYPES.get(a.typ, a.typ)
        d = const_dim(a, loop)
        L.append(f"  {t} r{i}[{d}];")
        L.append(f"  for (int d = 0; d < {d}; d++) r{i}[d] = {RED_OP[a.acc]}<{t}>::identity();")
    if red:
        # loops with reductions walk their tiles grid-stride over a capped 1-D grid (the
        # runtime sets p.ntiles and the tile grid p.tg0 x p.tg1 x ...), so the partials to
        # fold stay few; tile vt is placed with the XCD remap of the full tile count (the
        # grid is a multiple of 8 workgroups, so all of a workgroup's tiles sit on its XCD)
        L.append("  for (unsigned vt = blockIdx.x; vt < p.ntiles; vt += gridDim.x) {")
        L.append("  unsigned bx, by, bz;")
        L.append("  ops_tile_coords(vt, p.ntiles, p.tg0, p.tg1, bx, by, bz);")
    else:
        L.append("  unsigned bx, by, bz;")
        L.append("  ops_block_coords_dir(bx, by, bz, p.rev);")
    # the workgroup's origin (x0, y0, z0) is wave-uniform, so its part of every address is
    # scalar arithmetic; the lane's part uses 24-bit multiplies (the runtime checks that the
    # x and y steps fit), where a full 32-bit multiply per argument and dimension would be
    # a quarter-rate vector instruction
    L.append("  const unsigned tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;")
    L.append("  const int y0 = (int)(by * blockDim.y), z0 = (int)(bz * blockDim.z);")
    L.append("  const int y = y0 + (int)ty;")
    L.append("  const int z = z0 + (int)tz;")
    vx = loop_vx(loop)
    anc = align_anchor(loop)
    if anc >= 0:
        # x tiles start on 128-byte lines of the anchor argument's row (y, z): the lane's x
        # moves back by that row's offset into its first line (0 without alignment)
        row = f"p.p{anc}"
        if ndim >= 2:
            row += f" + (long long)y * p.sy{anc}"
        if ndim >= 3:
            row += f" + (long long)z * p.sz{anc}"
        L.append(f"  const int mis = p.al ? (int)(((uintptr_t)({row}) / sizeof(*p.p{anc})) & "
                 "(unsigned)(p.al - 1)) : 0;")
        L.append("  const int txm = (int)tx - mis;")
    if vx > 1:
        # vx points per lane, blockDim.x apart: every pass of the wavefront stays contiguous
        L.append("#pragma unroll")
        L.append(f"  for (int v = 0; v < {vx}; v++) {{")
        L.append(f"  const int x0 = (int)((bx * {vx} + v) * blockDim.x);")
    else:
        L.append("  const int x0 = (int)(bx * blockDim.x);")
    if anc >= 0:
        L.append("  const int x = x0 + txm;")
        L.append("  if (x >= 0 && x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {")
    else:
        L.append("  const int x = x0 + (int)tx;")
        L.append("  if (x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {")
    if v1:
        # sizes behind OPS_ACC<n>(...) for the dat arguments (xdim<n> = size[0], ...)
        L.append("    ops_v1_shape v1s;")
        for i, a in enumerate(loop.args):
            if a.kind == "dat":
                L.append(f"    v1s.xd[{i}] = p.xd{i}; v1s.yd[{i}] = p.yd{i}; "
                         f"v1s.zd[{i}] = p.zd{i}; v1s.md[{i}] = p.md{i};")
    call = []
    after = []  # stores of streamed read-write arguments
    for i, a in enumerate(loop.args):
        pr = params[i]
        if a.kind == "dat":
            if not pr.is_acc and not v1:
                raise TranslateError(f"{loop.file}:{loop.line}: kernel {loop.kernel} parameter {i} "
                                     "must be ACC<T>& (or a pointer addressed with OPS_ACC<n>) "
                                     "for an ops_arg_dat argument")
            t = pr.ctype
            off = dat_offset(i, ndim, anc >= 0)
            if a.prolong:
                # coarse index = global loop index / mgrid stride (C division, as the
                # reference's generated loops: templates/cpp/mpi_openmp/loop_host.cpp.j2:234-237)
                off = (f"((p.st0 + x) / p.mx{i}) * p.sx{i} + ((p.st1 + y) / p.my{i}) * p.sy{i}"
                       f" + ((p.st2 + z) / p.mz{i}) * p.sz{i}")
            # (pointer + uniform part) + lane part: the first sum stays scalar
            ptr = f"{'q' if noalias else 'p.p'}{i} + {off}" if not a.prolong else \
                f"{'q' if noalias else 'p.p'}{i} + ({off})"
            if nt_ok and streamed(loop, i, pointwise):
                L.append(f"    {t} *const e{i} = {ptr};")
                L.append(f"    {t} v{i} = __builtin_nontemporal_load(e{i});")
                if a.acc == "OPS_RW":
                    after.append(f"    *e{i} = v{i};")
                ptr = f"&v{i}"
            elif nt_ok and nt_stored(loop, i, written_once):
                L.append(f"    {t} *const e{i} = {ptr};")
                L.append(f"    {t} v{i};")
                after.append(f"    __builtin_nontemporal_store(v{i}, e{i});")
                ptr = f"&v{i}"
            if not pr.is_acc:
                # API v1: the pointer at this point; OPS_ACC<n> adds the stencil offset
                call.append(f"({t} *)({ptr})")
                continue
            mlit = cparse.int_literal(a.dim)
            if mlit == 1:
                if ndim == 1:
                    ctor = f"{ptr}"
                elif ndim == 2:
                    ctor = f"p.xd{i}, {ptr}"
                else:
                    ctor = f"p.xd{i}, p.yd{i}, {ptr}"
            else:
                if ndim == 1:
                    ctor = f"p.md{i}, p.xd{i}, {ptr}"
                elif ndim == 2:
                    ctor = f"p.md{i}, p.xd{i}, p.yd{i}, {ptr}"
                else:
                    ctor = f"p.md{i}, p.xd{i}, p.yd{i}, p.zd{i}, {ptr}"
            cq = "const " if pr.const else ""
            L.append(f"    {cq}ACC<{t}> a{i}({ctor});")
            call.append(f"a{i}")
        elif a.kind == "gbl":
            L.append(f"    const ops_gbl<{CTYPES.get(a.typ, a.typ)}, {const_dim(a, loop)}> g{i} = p.g{i};")
            call.append(f"g{i}.v")
        elif a.kind == "reduce":
            call.append(f"r{i}")
        else:
            L.append(f"    int idx{i}[3] = {{p.st0 + x, p.st1 + y, p.st2 + z}};")
            call.append(f"idx{i}")
    if v1:
        call.append("v1s")
    L.append(f"    {loop.kernel}({', '.join(call)});")
    L += after
    L.append("  }")
    if vx > 1:
        L.append("  }")
    if red:
        L.append("  }")
    for i in red:
        a = loop.args[i]
        t = CTYPES.get(a.typ, a.typ)
        d = const_dim(a, loop)
        L.append(f"  ops_reduce_partial<{RED_OP[a.acc]}<{t}>, {t}, {d}>(r{i}, p.rp{i});")
    L.append("}")
    return "\n".join(L)


def gen_stub(loop: Loop, params: List[Param], ns: str) -> str:
    k = loop.stub
    n = len(loop.args)
    sig = ", ".join(f"ops_arg arg{i}" for i in range(n))
    L = []
    bat = batchable(loop)
    if bat:
        # the pending batch of this loop's kernel and its launcher
        L += [f"static {ns}::ops_kb_{k} ops_batch_{k};",
              f"static int ops_batch_n_{k} = 0;",
              f"static void *ops_batch_stream_{k} = nullptr;",
              f"static void ops_batch_flush_{k}(void) {{",
              f"  const int n = ops_batch_n_{k};",
              "  if (n == 0) return;",
              f"  ops_batch_n_{k} = 0;",
              f"  ops_batch_{k}.n = n;",
              f"  hipLaunchKernelGGL({ns}::ops_kern_{k}_batch, dim3(ops_batch_{k}.first[n]), dim3(256), 0, "
              f"(hipStream_t)ops_batch_stream_{k}, ops_batch_{k});",
              f"  ops_hip_check_launch(\"{k} (batch)\");",
              "}", ""]
    # the stub again for a sub-range, with the recorded arguments (loop chains)
    proto = (f"void ops_par_loop_{k}(char const *name, ops_block block, int dim, int *range"
             + (f", {sig}" if n else "") + ")")
    L += [proto + ";",
          f"static void ops_relaunch_{k}(const char *name, ops_block block, int dim, int *range, ops_arg *a) {{",
          f"  ops_par_loop_{k}(name, block, dim, range" + "".join(f", a[{i}]" for i in range(n)) + ");",
          "}", ""]
    L += [proto + " {"]
    L.append(f"  static int kid = ops_hip_register_kernel(\"{loop.kernel}\");")
    L.append(f"  ops_arg args[{max(n, 1)}] = {{{', '.join(f'arg{i}' for i in range(n)) or 'ops_arg_idx()'}}};")
    L.append("  ops_hip_loop lp;")
    L.append(f"  lp.vx = {loop_vx(loop)};")
    L.append(f"  lp.align_arg = {align_anchor(loop)};")
    mask = sum(1 << i for i, a in enumerate(loop.args) if a.kind == "dat" and a.prolong)
    L.append(f"  lp.prolong_mask = {mask}u;")
    L.append(f"  lp.batch = {'ops_batch_flush_' + k if bat else 'nullptr'};")
    L.append(f"  lp.relaunch = ops_relaunch_{k};")
    L.append(f"  if (ops_hip_loop_begin(&lp, kid, name, block, dim, range, args, {n})) {{")
    L.append(f"    {ns}::ops_kp_{k} p;")
    L.append("    p.cnt0 = lp.count[0]; p.cnt1 = lp.count[1]; p.cnt2 = lp.count[2];")
    L.append("    p.st0 = lp.start[0]; p.st1 = lp.start[1]; p.st2 = lp.start[2];")
    if align_anchor(loop) >= 0:
        L.append("    p.al = lp.align;")
    if not any(a.kind == "reduce" for a in loop.args):
        L.append("    p.rev = lp.rev;")
    if any(a.kind == "reduce" for a in loop.args):
        L.append("    p.ntiles = lp.tiles; p.tg0 = lp.tgrid[0]; p.tg1 = lp.tgrid[1];")
    for i, a in enumerate(loop.args):
        if a.kind == "dat":
            t = params[i].ctype
            L.append(f"    p.p{i} = ({t} *)lp.ptr[{i}]; p.sx{i} = lp.step[{i}][0]; p.sy{i} = lp.step[{i}][1]; "
                     f"p.sz{i} = lp.step[{i}][2];")
            L.append(f"    p.xd{i} = lp.xdim[{i}]; p.yd{i} = lp.ydim[{i}]; p.zd{i} = lp.zdim[{i}]; p.md{i} = lp.mdim[{i}];")
            if a.prolong:
                L.append(f"    p.mx{i} = lp.mg[{i}][0]; p.my{i} = lp.mg[{i}][1]; p.mz{i} = lp.mg[{i}][2];")
        elif a.kind == "gbl":
            L.append(f"    memcpy(p.g{i}.v, arg{i}.data, sizeof(p.g{i}.v));")
        elif a.kind == "reduce":
            t = CTYPES.get(a.typ, a.typ)
            L.append(f"    p.rp{i} = ({t} *)lp.red_part[{i}];")
    # timed launches go between the runtime's two events (they bracket the kernel itself)
    launch = ("if (lp.timed) hipExtLaunchKernelGGL({0}, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), "
              "dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, "
              "(hipEvent_t)lp.ev_start, (hipEvent_t)lp.ev_stop, 0, p{1}); "
              "else hipLaunchKernelGGL({0}, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), "
              "dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, p{1});")
    if bat:
        # join the pending batch (an untimed launch while batching is on), else launch alone
        # after whatever this kernel has pending
        L.append(f"    if ({ns}::ops_nb_{k} >= 2 && !lp.timed && !lp.alias && "
                 f"ops_hip_batch_join(kid, ops_batch_flush_{k})) {{")
        L.append(f"      if (ops_batch_n_{k} == {ns}::ops_nb_{k}) ops_batch_flush_{k}();")
        L.append(f"      const int j = ops_batch_n_{k}++;")
        L.append(f"      auto &B = ops_batch_{k};")
        L.append("      if (j == 0) B.first[0] = 0;")
        L.append("      B.p[j] = p;")
        L.append("      B.g0[j] = lp.grid[0]; B.g1[j] = lp.grid[1];")
        L.append("      B.b0[j] = lp.block[0]; B.b1[j] = lp.block[1]; B.b2[j] = lp.block[2];")
        L.append("      B.nb[j] = lp.nblocks;")
        L.append("      B.first[j + 1] = B.first[j] + ((lp.nblocks + 7u) & ~7u);")
        L.append(f"      ops_batch_stream_{k} = lp.stream;")
        L.append("      ops_hip_loop_end(&lp, args, " + str(n) + ");")
        L.append("      return;")
        L.append("    }")
        L.append(f"    ops_batch_flush_{k}();")
    if noalias_variant(loop):
        ptrs = "".join(f", p.p{i}" for i, a in enumerate(loop.args) if a.kind == "dat")
        L.append("    if (lp.alias)")
        L.append("    {")
        L.append("      " + launch.format(f"{ns}::ops_kern_{k}_alias", ""))
        L.append("    } else {")
        L.append("      " + launch.format(f"{ns}::ops_kern_{k}", ptrs))
        L.append("    }")
    else:
        L.append("    " + launch.format(f"{ns}::ops_kern_{k}", ""))
    L.append("    ops_hip_check_launch(name);")
    L.append("  }")
    L.append(f"  ops_hip_loop_end(&lp, args, {n});")
    L.append("}")
    return "\n".join(L)


def const_decl(c: Const) -> Tuple[str, str]:
    t = CTYPES.get(c.typ, c.typ)
    d = cparse.int_literal(c.dim)
    if d is None:
        return f"__device__ {t} *{c.name};", "-1"
    if d == 1:
        return f"__constant__ {t} {c.name};", f"(long)sizeof({t})"
    return f"__constant__ {t} {c.name}[{d}];", f"(long)sizeof({t}) * {d}"


def translate_app(app_name: str, sources: List[str], out_dir: str,
                  app_dir: Optional[str] = None, ntstore: Optional[bool] = None,
                  align: Optional[bool] = None) -> Dict[str, object]:
    """Translate the application made of `sources`; returns the generated file list.
    `ntstore` overrides OPS_GEN_NTSTORE (non-temporal stores of write-once arguments),
    `align` OPS_GEN_ALIGN (x tiles on 128-byte lines of a written argument's rows)."""
    global _align_on
    _align_on = ALIGN if align is None else align
    use_nt = NTSTORE if ntstore is None else ntstore
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    app_dir = app_dir or str(Path(sources[0]).parent)
    headers = kernel_headers(app_dir, sources)
    scanned = {s: scan_source(s) for s in sources}
    ndims = {nd for (_, _, _, nd) in scanned.values() if nd is not None}
    if len(ndims) > 1:
        raise TranslateError(f"{app_name}: sources mix OPS_nD dimensionalities {ndims}")
    ndim = ndims.pop() if ndims else 2
    consts: Dict[str, Const] = {}
    for (_, _, cs, _) in scanned.values():
        for c in cs:
            consts.setdefault(c.name, c)
    # stencils declared as prolong stencils (their loops divide the loop index by the mgrid
    # stride; the runtime checks every launch against the stencil it is given)
    prolong = set()
    for (text, _, _, _) in scanned.values():
        masked = cparse.mask(text)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*=\s*ops_decl_prolong_stencil\s*\(", masked):
            prolong.add(m.group(1))
    for (_, lps, _, _) in scanned.values():
        for lp in lps:
            for a in lp.args:
                if a.kind == "dat" and a.stencil.strip() in prolong:
                    a.prolong = True
    # stubs: one per (kernel, argument signature)
    variants: Dict[str, List[tuple]] = {}
    loops_all: List[Loop] = []
    for s in sources:
        for lp in scanned[s][1]:
            key = tuple(a.key() for a in lp.args)
            vs = variants.setdefault(lp.kernel, [])
            if key not in vs:
                vs.append(key)
            idx = vs.index(key)
            lp.stub = lp.kernel if idx == 0 else f"{lp.kernel}__v{idx}"
            loops_all.append(lp)
    ns = f"ops_gen_{re.sub(r'[^A-Za-z0-9_]', '_', app_name)}"
    # ---- kernels translation unit
    used_headers = []
    for lp in loops_all:
        if lp.kernel not in headers:
            raise TranslateError(f"{lp.file}:{lp.line}: kernel {lp.kernel} not found in any header of {app_dir}")
        h = headers[lp.kernel][0]
        if h not in used_headers:
            used_headers.append(h)
    used_kernels = {lp.kernel for lp in loops_all}
    v1_kernels = set()
    hoisted, bodies = [], []
    for h in used_headers:
        text = strip_include_guard(Path(h).read_text())
        # API v1 kernels take the accessor sizes as one more (by-value) parameter; so do the
        # header's other pointer-only functions that address through OPS_ACC<n> (kernels no
        # loop of this application uses still compile for the device)
        fdefs = [f for f in cparse.find_functions(text)
                 if is_v1_kernel(text, f) and
                 (f.name in used_kernels or
                  re.search(r"\bOPS_ACC(?:_MD)?\d+\s*\(", cparse.mask(text[f.body_open:f.end])))]
        for f in sorted(fdefs, key=lambda f: f.start, reverse=True):
            text = add_v1_param(text, f)
            v1_kernels.add(f.name)
        keep = []
        for line in text.splitlines():
            if re.match(r"\s*#\s*include\b", line):
                if line.strip() not in hoisted:
                    hoisted.append(line.strip())
                keep.append("")
            else:
                keep.append(line)
        bodies.append((h, "\n".join(keep)))
    H = [f"// Generated by ops_amd.translator (target hip_gfx950) for application '{app_name}'.",
         "// Do not edit: regenerate from the application sources.",
         f"#define OPS_{ndim}D",
         "#include <hip/hip_runtime.h>",
         "#include <hip/hip_ext.h>",
         "#include <string.h>",
         "#include <math.h>",
         "#include \"ops_lib_core.h\"",
         "#include \"ops_hip_rt.h\"",
         "#include \"ops_hip_kernel.h\"",
         "#include \"ops_macros.h\""]
    if (Path(app_dir) / "user_types.h").exists():
        H.append("#include \"user_types.h\"")
    H += hoisted
    H.append("")
    H.append(f"namespace {ns} {{")
    H.append("")
    H.append("// ---- ops_decl_const symbols")
    regs = []
    for c in consts.values():
        decl, size = const_decl(c)
        H.append(decl)
        regs.append(f"    {{\"{c.name}\", (const void *)&{ns}::{c.name}, {size}}},")
    H.append("")
    H.append("// ---- user kernels, compiled for host and gfx950")
    H.append("#pragma clang force_cuda_host_device begin")
    for h, body in bodies:
        H.append(f"// from {os.path.basename(h)}")
        H.append(body)
    H.append("#pragma clang force_cuda_host_device end")
    H.append("")
    done = set()
    stub_defs = []
    for lp in loops_all:
        if lp.stub in done:
            continue
        done.add(lp.stub)
        fdef = headers[lp.kernel][1]
        params = [parse_param(p) for p in fdef.params]
        if len(params) != len(lp.args):
            raise TranslateError(f"{lp.file}:{lp.line}: kernel {lp.kernel} takes {len(params)} parameters "
                                 f"but the loop passes {len(lp.args)} arguments")
        htext = Path(headers[lp.kernel][0]).read_text()
        pw = pointwise_params(htext, fdef) if STREAM else set()
        wo = written_once_params(htext, fdef) if use_nt else set()
        H.append(gen_kernel(lp, params, ndim, lp.kernel in v1_kernels, pw, wo))
        H.append("")
        stub_defs.append(gen_stub(lp, params, ns))
    H.append(f"}}  // namespace {ns}")
    H.append("")
    H += stub_defs
    H.append("")
    # this application's ops_decl_const homes: the application defines ops_decl_const_char
    # itself, as the reference's generated master file does (cuda/master_kernel.cpp.j2:
    # 40-67), so two applications in one process that declare constants of one name (of
    # different types) each write their own symbols (the library links -Bsymbolic)
    H.append("namespace {")
    H.append("struct ops_const_home { const char *name; const void *symbol; long nbytes; };")
    H.append("const ops_const_home ops_const_homes[] = {")
    H += regs or ["    {nullptr, nullptr, 0},"]
    H.append("};")
    H.append("const ops_const_home *ops_const_find(const char *name) {")
    H.append("  for (const ops_const_home &h : ops_const_homes)")
    H.append("    if (h.name && !strcmp(h.name, name)) return &h;")
    H.append("  return nullptr;")
    H.append("}")
    H.append("}  // namespace")
    H.append("")
    H.append("extern \"C\" void ops_decl_const_char(OPS_instance *, int dim, char const *, int type_size, "
             "char *data, char const *name) {")
    H.append("  const ops_const_home *h = ops_const_find(name);")
    H.append("  if (h) ops_hip_const_set(h->symbol, h->nbytes, h->name, data, (long)dim * type_size);")
    H.append("}")
    H.append("")
    H.append("extern \"C\" void *ops_hip_const_device_ptr(const char *name) {")
    H.append("  const ops_const_home *h = ops_const_find(name);")
    H.append("  return ops_hip_const_symbol_ptr(h ? h->symbol : nullptr, h ? h->nbytes : 0, name);")
    H.append("}")
    kpath = out / f"{app_name}_kernels.hip"
    _write_if_changed(kpath, "\n".join(H) + "\n")
    # ---- rewritten host sources
    host_files = []
    for s in sources:
        text, loops, _, _ = scanned[s]
        new = text
        for lp in sorted(loops, key=lambda l: l.call.start, reverse=True):
            c = lp.call
            # drop the kernel argument and rename the callee
            first_end = c.arg_spans[1][0] if len(c.arg_spans) > 1 else c.close
            new = new[:c.start] + f"ops_par_loop_{lp.stub}(" + new[first_end:]
        protos = []
        seen = set()
        for lp in loops:
            if lp.stub in seen:
                continue
            seen.add(lp.stub)
            protos.append(f"void ops_par_loop_{lp.stub}(char const *, ops_block, int, int *"
                          + ", ops_arg" * len(lp.args) + ");")
        if protos:
            lines = new.split("\n")
            at = None
            for i, line in enumerate(lines):
                if re.match(r'\s*#\s*include\s*[<"]ops_seq(_v2)?\.h[>"]', line):
                    at = i + 1
            if at is None:
                for i, line in enumerate(lines):
                    if re.match(r"\s*#\s*include\b", line):
                        at = i + 1
            at = at or 0
            lines[at:at] = ["/* ops_par_loop declarations (hip_gfx950) */"] + protos
            new = "\n".join(lines)
        hp = out / (Path(s).stem + "_ops.cpp")
        _write_if_changed(hp, new)
        host_files.append(str(hp))
    return {"kernels": str(kpath), "host": host_files, "ndim": ndim,
            "loops": [(lp.kernel, lp.stub, lp.file, lp.line) for lp in loops_all],
            "consts": list(consts)}


def _write_if_changed(path: Path, text: str):
    if path.exists() and path.read_text() == text:
        return
    path.write_text(text)
