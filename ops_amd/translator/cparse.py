"""Small C/C++ source scanner for the OPS translator.

The call grammar the translator needs is narrow (the reference parses the same shapes
with libclang: ops_translator/ops-translator/cpp/parser.py:80-433): calls to
ops_par_loop / ops_arg_* / ops_decl_const and top-level function definitions in kernel
headers. A scanner that blanks comments and literals and then matches balanced
brackets covers it without a compiler front end.
"""
from __future__ import annotations

import re
from dataclasses import dataclass
from typing import List, Optional, Tuple


def mask(text: str) -> str:
    """Return `text` with comments and the insides of string/char literals replaced by
    spaces (newlines kept), so that positions map 1:1 onto the original."""
    out = list(text)
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c == "/" and i + 1 < n and text[i + 1] == "/":
            j = text.find("\n", i)
            j = n if j < 0 else j
            for k in range(i, j):
                out[k] = " "
            i = j
        elif c == "/" and i + 1 < n and text[i + 1] == "*":
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
            for k in range(i, j):
                if out[k] != "\n":
                    out[k] = " "
            i = j
        elif c in "\"'":
            q = c
            j = i + 1
            while j < n and text[j] != q:
                if text[j] == "\\":
                    j += 1
                j += 1
            for k in range(i + 1, min(j, n)):
                if out[k] != "\n":
                    out[k] = " "
            i = j + 1
        else:
            i += 1
    return "".join(out)


_OPEN = {"(": ")", "[": "]", "{": "}"}


def match_bracket(masked: str, open_idx: int) -> int:
    """Index of the bracket closing the one at `open_idx`."""
    stack = []
    for i in range(open_idx, len(masked)):
        c = masked[i]
        if c in _OPEN:
            stack.append(_OPEN[c])
        elif c in ")]}":
            if not stack or stack[-1] != c:
                raise ValueError(f"unbalanced bracket at offset {i}")
            stack.pop()
            if not stack:
                return i
    raise ValueError(f"unterminated bracket at offset {open_idx}")


def split_top(masked: str, lo: int, hi: int) -> List[Tuple[int, int]]:
    """Spans of the comma-separated items in masked[lo:hi] at bracket depth 0."""
    spans, depth, start = [], 0, lo
    for i in range(lo, hi):
        c = masked[i]
        if c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
        elif c == "," and depth == 0:
            spans.append((start, i))
            start = i + 1
    if masked[start:hi].strip():
        spans.append((start, hi))
    return spans


@dataclass
class Call:
    name: str
    start: int   # offset of the callee name
    open: int    # offset of "("
    close: int   # offset of ")"
    args: List[str]
    arg_spans: List[Tuple[int, int]]
    line: int


def find_calls(text: str, masked: str, names) -> List[Call]:
    if isinstance(names, str):
        names = [names]
    pat = re.compile(r"(?<![A-Za-z0-9_])(" + "|".join(re.escape(n) for n in names) + r")\s*\(")
    calls = []
    for m in pat.finditer(masked):
        op = m.end() - 1
        cl = match_bracket(masked, op)
        spans = split_top(masked, op + 1, cl)
        args = [text[s:e].strip() for s, e in spans]
        calls.append(Call(m.group(1), m.start(), op, cl, args, spans,
                          text.count("\n", 0, m.start()) + 1))
    return calls


@dataclass
class FuncDef:
    name: str
    start: int        # start of the declaration (return type)
    end: int          # one past the closing brace
    params: List[str]
    body_open: int
    params_close: int = -1   # index of the ')' closing the parameter list


_FUNC = re.compile(r"(?m)^[ \t]*((?:static\s+|inline\s+)*void)\s+([A-Za-z_]\w*)\s*\(")


def find_functions(text: str, masked: Optional[str] = None) -> List[FuncDef]:
    """Top-level `void name(...) { ... }` definitions (OPS user kernels return void)."""
    if masked is None:
        masked = mask(text)
    out = []
    depth_at = _brace_depths(masked)
    for m in _FUNC.finditer(masked):
        if depth_at[m.start()] != 0:
            continue
        op = m.end() - 1
        cl = match_bracket(masked, op)
        j = cl + 1
        while j < len(masked) and masked[j] in " \t\r\n":
            j += 1
        if j >= len(masked) or masked[j] != "{":
            continue  # a declaration
        be = match_bracket(masked, j)
        # parameter text without its comments (a parameter list may carry `// ...` notes)
        params = [" ".join(masked[s:e].split()) for s, e in split_top(masked, op + 1, cl)]
        if params == ["void"]:
            params = []
        out.append(FuncDef(m.group(2), m.start(), be + 1, params, j, cl))
    return out


def _brace_depths(masked: str) -> List[int]:
    d, out = 0, [0] * (len(masked) + 1)
    for i, c in enumerate(masked):
        out[i] = d
        if c == "{":
            d += 1
        elif c == "}":
            d -= 1
    out[len(masked)] = d
    return out


def unquote(s: str) -> str:
    s = s.strip()
    if len(s) >= 2 and s[0] == '"' and s[-1] == '"':
        return s[1:-1]
    return s


def int_literal(s: str) -> Optional[int]:
    s = s.strip()
    return int(s) if re.fullmatch(r"[0-9]+", s) else None
