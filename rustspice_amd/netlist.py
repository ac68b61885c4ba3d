"""SPICE netlist front end: text to the device records the GPU stamp path reads.

Host-side restatement of RustSpice's netlist reader and circuit builder:

- ``parse_netlist`` / ``parse_netlist_file``: crates/sim-core/src/netlist.rs:135-328
  (continuation lines, comments, control and device statements), with
  ``.include`` expansion (:1214-1264);
- ``elaborate_netlist``: :858-932 (``.param`` tables and expressions
  :1266-1349, :1557-1841; ``.subckt`` expansion :1351-1439, :1861-1952);
- ``build_circuit``: :934-1211 with the NodeTable of crates/sim-core/src/circuit.rs:31-57
  and the options of crates/sim-core/src/options.rs:185-360;
- ``lower``: the built circuit as ``rs_device`` records (include/rs_circuit.h),
  reading every Instance the way the stamp functions do
  (crates/sim-core/src/stamp.rs:128-145, 242-553, 715-787).

The reference's own behaviour is kept where it is surprising, because the
records must produce the same matrices:

- a PULSE/PWL/SIN source's waveform goes into the statement's extras and never
  reaches ``Instance.value``, so such a source is not stamped at all (its nodes
  still exist);
- ``.model X PMOS`` does not make a device PMOS: only a ``type`` parameter
  containing ``pmos`` (or equal to ``p``) or a ``pmos`` key does, and the model
  level defaults to 49 (BSIM3);
- ``.ic`` nodes are entered in the node table before any device node;
- inside a subcircuit every node that is not a port, ``0`` included, becomes
  ``<instance>:<node>``; the body's ``.model`` names are scoped to
  ``<instance>.<model>`` while its devices keep the unscoped name;
- device values pass through ``.param`` evaluation and come back as the text
  of the double (``1k`` becomes ``1000``).

Devices the GPU stamp path does not evaluate (L, E, G, F, H, Verilog-A) make
``lower`` raise :class:`UnsupportedDevice`; nothing falls back to the CPU.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import netlists as NL

# ---------------------------------------------------------------------------
# Rust string and number semantics the reader depends on
# ---------------------------------------------------------------------------
#: char::is_whitespace (Unicode White_Space); Python's str.isspace also takes
#: the four information separators \x1c-\x1f, which Rust does not
_WS = ("\t\n\x0b\x0c\r \x85\xa0\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007"
       "\u2008\u2009\u200a\u2028\u2029\u202f\u205f\u3000")
_WS_SPLIT = re.compile("[" + re.escape(_WS) + "]+")
_ASCII_LOWER = str.maketrans("ABCDEFGHIJKLMNOPQRSTUVWXYZ", "abcdefghijklmnopqrstuvwxyz")
_ASCII_UPPER = str.maketrans("abcdefghijklmnopqrstuvwxyz", "ABCDEFGHIJKLMNOPQRSTUVWXYZ")


def _lower(s: str) -> str:
    """str::to_ascii_lowercase."""
    return s.translate(_ASCII_LOWER)


def _upper(s: str) -> str:
    return s.translate(_ASCII_UPPER)


def _trim(s: str) -> str:
    return s.strip(_WS)


def _split_ws(s: str) -> list[str]:
    """str::split_whitespace."""
    return [t for t in _WS_SPLIT.split(s) if t]


def _lines(text: str) -> list[str]:
    """str::lines: split at \\n, drop one trailing \\r per line, no final empty line."""
    if not text:
        return []
    parts = text.split("\n")
    if text.endswith("\n"):
        parts.pop()
    return [p[:-1] if p.endswith("\r") else p for p in parts]


def _trim_start_matches(s: str, pat: str) -> str:
    while pat and s.startswith(pat):
        s = s[len(pat):]
    return s


def _trim_end_matches(s: str, pat: str) -> str:
    while pat and s.endswith(pat):
        s = s[: -len(pat)]
    return s


_F64 = re.compile(r"[+-]?((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?|[iI][nN][fF]([iI][nN][iI][tT][yY])?|[nN][aA][nN])")
_USIZE = re.compile(r"\+?\d+")
_I64 = re.compile(r"[+-]?\d+")


def rust_f64(s: str):
    """str::parse::<f64>: sign, digits with optional point and exponent, or
    inf/infinity/nan in any case; no whitespace, underscores or hex."""
    return float(s) if _F64.fullmatch(s) else None


def _rust_usize(s: str):
    return int(s) if _USIZE.fullmatch(s) else None


def _rust_i64(s: str):
    if not _I64.fullmatch(s):
        return None
    v = int(s)
    return v if -(1 << 63) <= v < (1 << 63) else None


def parse_number_with_suffix(token: str):
    """netlist.rs:1635-1662 (the stamp's copy, stamp.rs:720-745, is the same):
    ASCII-lower-cased and trimmed; ``meg`` or one letter of f p n u m k g t
    scales the number before it; the product is rounded once."""
    t = _trim(_lower(token))
    if t.endswith("meg"):
        num, mult = t[:-3], 1e6
    else:
        cut = max(len(t) - 1, 0)
        mult = {"f": 1e-15, "p": 1e-12, "n": 1e-9, "u": 1e-6, "m": 1e-3, "k": 1e3, "g": 1e9,
                "t": 1e12}.get(t[cut:])
        num = t[:cut] if mult is not None else t
        mult = 1.0 if mult is None else mult
    v = rust_f64(num)
    return None if v is None else v * mult


def _number_or_plain(s: str):
    """``parse_number_with_suffix(s).or_else(|| s.parse().ok())``."""
    v = parse_number_with_suffix(s)
    return rust_f64(s) if v is None else v


def rust_f64_display(v: float) -> str:
    """``f64::to_string``: the shortest digits that read back to v, written
    out without an exponent; ``inf``, ``-inf``, ``NaN``; ``-0`` for negative zero."""
    if v != v:
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    r = repr(abs(v))
    mant, _, exp = r.partition("e")
    e = int(exp) if exp else 0
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    if not digits:
        return sign + "0"
    # value = 0.<ip+fp> * 10^(len(ip) + e), with leading zeros of ip removed
    point = len(ip) + e - (len(ip + fp) - len((ip + fp).lstrip("0")))
    digits = digits.rstrip("0") or "0"
    if point <= 0:
        out = "0." + "0" * (-point) + digits
    elif point >= len(digits):
        out = digits + "0" * (point - len(digits))
    else:
        out = digits[:point] + "." + digits[point:]
    return sign + out


# ---------------------------------------------------------------------------
# Syntax tree (netlist.rs:1-133)
# ---------------------------------------------------------------------------
@dataclass
class Param:
    key: str
    value: str


@dataclass
class PolySpec:
    degree: int
    coeffs: list


@dataclass
class DeviceStmt:
    name: str
    kind: str  # "R" "C" "L" "V" "I" "D" "M" "E" "G" "F" "H" "X" or "?" (unknown)
    nodes: list
    model: str | None
    control: str | None
    value: str | None
    params: list
    extras: list
    poly: PolySpec | None
    raw: str
    line: int
    ac_mag: float | None = None
    ac_phase: float | None = None

    def copy(self) -> "DeviceStmt":
        return DeviceStmt(self.name, self.kind, list(self.nodes), self.model, self.control, self.value,
                          [Param(p.key, p.value) for p in self.params], list(self.extras), self.poly,
                          self.raw, self.line, self.ac_mag, self.ac_phase)


@dataclass
class ControlStmt:
    command: str
    kind: str  # "param" "model" "subckt" "ends" "include" "op" "dc" "tran" "ac" "ic" "end" "hdl" "osdi" "option" "other"
    args: list
    params: list
    model_name: str | None = None
    model_type: str | None = None
    subckt_name: str | None = None
    subckt_ports: list = field(default_factory=list)
    raw: str = ""
    line: int = 0


@dataclass
class Comment:
    text: str


class Empty:
    pass


@dataclass
class ParseError:
    line: int
    message: str


@dataclass
class NetlistAst:
    title: str | None
    statements: list
    errors: list


_CONTROL_KINDS = {".param": "param", ".model": "model", ".subckt": "subckt", ".ends": "ends",
                  ".include": "include", ".op": "op", ".dc": "dc", ".tran": "tran", ".ac": "ac",
                  ".ic": "ic", ".end": "end", ".hdl": "hdl", ".osdi": "osdi", ".option": "option",
                  ".options": "option"}
_DEVICE_LETTERS = "RCLVIDMEGFHX"


def parse_netlist(text: str) -> NetlistAst:
    """netlist.rs:144-194. A statement is parsed when the next one starts, and
    carries that line's number; comment and empty lines in between come first."""
    ast = NetlistAst(None, [], [])
    pending = ""
    lines = _lines(text)
    for idx, raw in enumerate(lines):
        line_no = idx + 1
        t = _trim(raw)
        if not t:
            ast.statements.append(Empty())
            continue
        if t.startswith("*"):
            ast.statements.append(Comment(t))
            continue
        if t.startswith("+"):
            if not pending:
                ast.errors.append(ParseError(line_no, "continuation line has no preceding line"))
                continue
            pending += " " + _trim(_trim_start_matches(t, "+"))
            continue
        if pending:
            _parse_statement(pending, line_no, ast)
        pending = t
    if pending:
        _parse_statement(pending, len(lines), ast)
    return ast


def _read_with_includes(path: str, visited: set, errors: list) -> str:
    """netlist.rs:1214-1264: ``.include "file"`` lines replaced by the file's
    text (relative to the including file), each file at most once."""
    key = os.path.normpath(path)
    if key in visited:
        errors.append(ParseError(0, f"include circular reference: {path}"))
        return ""
    visited.add(key)
    try:
        with open(path, encoding="utf-8") as f:
            content = f.read()
    except (OSError, UnicodeDecodeError):
        errors.append(ParseError(0, f"cannot read file: {path}"))
        content = ""
    base = os.path.dirname(path) or "."
    out = []
    for line in _lines(content):
        t = _trim(line)
        if _lower(t).startswith(".include"):
            toks = _split_ws(t)
            inc = (toks[1] if len(toks) > 1 else "").strip('"')
            if not inc:
                errors.append(ParseError(0, f"include statement missing path: {path}"))
                continue
            out.append(_read_with_includes(os.path.join(base, inc), visited, errors) + "\n")
        else:
            out.append(line + "\n")
    return "".join(out)


def parse_netlist_file(path: str) -> NetlistAst:
    """netlist.rs:135-142."""
    errors: list = []
    text = _read_with_includes(path, set(), errors)
    ast = parse_netlist(text)
    ast.errors.extend(errors)
    return ast


def _parse_statement(line: str, line_no: int, ast: NetlistAst) -> None:
    """netlist.rs:196-328."""
    toks = _split_ws(line)
    if not toks:
        ast.statements.append(Empty())
        return
    first, rest = toks[0], toks[1:]
    if first.startswith("."):
        command = _lower(first)
        args, params = _split_args_params(rest)
        kind = _CONTROL_KINDS.get(command, "other")
        c = ControlStmt(command, kind, args, params, raw=line, line=line_no)
        if command == ".title" and args:
            ast.title = " ".join(args)
        if kind == "model":
            if len(args) >= 2:
                c.model_name, c.model_type = args[0], args[1]
            else:
                ast.errors.append(ParseError(line_no, "model statement missing name/type"))
        if kind == "subckt":
            if args:
                c.subckt_name, c.subckt_ports = args[0], args[1:]
            else:
                ast.errors.append(ParseError(line_no, "subckt statement missing name"))
        ast.statements.append(c)
        return
    letter = _upper(first[0])
    kind = letter if letter in _DEVICE_LETTERS else "?"
    if kind == "?":
        ast.errors.append(ParseError(line_no, f"unknown device type: {first}"))
    args, params = _split_args_params(rest)
    nodes, model, value, extras, poly = _split_device_fields(kind, args)
    control = _control_name(kind, args)
    _validate(first, kind, nodes, model, control, value, extras, poly, line_no, ast.errors)
    ac_mag, ac_phase = _ac_params(kind, args)
    ast.statements.append(DeviceStmt(first, kind, nodes, model, control, value, params, extras, poly,
                                     line, line_no, ac_mag, ac_phase))


def _split_top_level_commas(s: str) -> list[str]:
    """netlist.rs:374-396."""
    out, depth, start = [], 0, 0
    for i, ch in enumerate(s):
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth = max(depth - 1, 0)
        elif ch == "," and depth == 0:
            out.append(s[start:i])
            start = i + 1
    if start < len(s):
        out.append(s[start:])
    return out


def _split_args_params(tokens: list[str]):
    """netlist.rs:330-372: ``key=value`` tokens (split at top-level commas,
    one enclosing parenthesis dropped) become parameters, the rest arguments."""
    args, params = [], []
    for tok in tokens:
        parts = _split_top_level_commas(tok) if ("=" in tok and "," in tok) else [tok]
        for raw in parts:
            part = _trim(raw)
            if not part:
                continue
            if part.startswith("(") and "=" in part:
                part = part[1:]
            if part.endswith(")") and "=" in part:
                eq = part.find("=")
                if "(" not in part[eq + 1:]:
                    part = part[:-1]
            if not part:
                continue
            if "=" in part:
                k, v = part.split("=", 1)
                params.append(Param(k, v))
            else:
                args.append(part)
    return args, params


def _is_poly(tok: str) -> bool:
    return _upper(tok).startswith("POLY")


def _is_waveform(tok: str) -> bool:
    u = _upper(tok)
    return u in ("AC", "SIN", "PULSE", "EXP", "SFFM", "PWL") or u.startswith(
        ("SIN(", "PULSE(", "EXP(", "SFFM(", "PWL(", "AC("))


def _split_device_fields(kind: str, args: list[str]):
    """netlist.rs:398-541."""
    if not args:
        return [], None, None, [], None
    nodes, model, value, extras, poly = [], None, None, [], None
    n = len(args)
    if kind in "RCLVI":
        if n >= 3:
            nodes = args[0:2]
            if _lower(args[2]) == "dc" and n >= 4:
                value, extras = args[3], args[4:]
            elif kind in "VI" and _is_waveform(args[2]):
                extras = args[2:]
            else:
                value, extras = args[2], args[3:]
        else:
            nodes = list(args)
    elif kind == "D":
        if n >= 3:
            nodes, model, extras = args[0:2], args[2], args[3:]
        else:
            nodes = list(args)
    elif kind == "M":
        if n >= 5:
            nodes, model, extras = args[0:4], args[4], args[5:]
        elif n == 4:
            nodes, model = args[0:3] + ["0"], args[3]
        else:
            nodes = list(args)
    elif kind in "EG":
        pidx = next((i for i, a in enumerate(args) if _is_poly(a)), None)
        if pidx is not None:
            if n >= 2:
                nodes = args[0:2]
            if pidx > 2:
                extras = args[2:pidx]
            extras = extras + args[pidx:]
        elif n >= 5:
            nodes, value, extras = args[0:4], args[4], args[5:]
        else:
            nodes = list(args)
    elif kind in "FH":
        pidx = next((i for i, a in enumerate(args) if _is_poly(a)), None)
        if pidx is not None:
            if n >= 2:
                nodes = args[0:2]
            extras = args[pidx:]
        elif n >= 4:
            nodes, value, extras = args[0:2], args[3], args[4:]
        else:
            nodes = list(args)
    elif kind == "X":
        if n >= 2:
            nodes, model = args[:-1], args[-1]
        else:
            nodes = list(args)
    else:
        nodes = list(args)
    if kind in "EGFH" and kind != "?":
        poly, extras = _parse_poly(extras)
        if poly is not None:
            value = None
    return list(nodes), model, value, list(extras), poly


def _parse_poly(tokens: list[str]):
    """netlist.rs:1525-1555. The degree of ``POLY(n)`` is read from the token as
    written after stripping an upper-case ``POLY(``: ``poly(2)`` gives degree 1."""
    for idx, tok in enumerate(tokens):
        u = _upper(tok)
        if not u.startswith("POLY"):
            continue
        deg_tok, skip = None, 1
        if u.startswith("POLY(") and u.endswith(")"):
            deg_tok = _trim_end_matches(_trim_start_matches(tok, "POLY("), ")")
        elif idx + 1 < len(tokens):
            nxt = _trim(tokens[idx + 1])
            if nxt.startswith("(") and nxt.endswith(")"):
                deg_tok, skip = _trim_end_matches(_trim_start_matches(nxt, "("), ")"), 2
            elif all("0" <= ch <= "9" for ch in nxt):
                deg_tok, skip = nxt, 2
        deg = _rust_usize(deg_tok) if deg_tok is not None else None
        return PolySpec(1 if deg is None else deg, tokens[idx + skip:]), tokens[:idx]
    return None, list(tokens)


def _control_name(kind: str, args: list[str]):
    """netlist.rs:1441-1452: F/H name their controlling source third."""
    if kind in "FH" and kind != "?" and len(args) > 2 and not _is_poly(args[2]):
        return args[2]
    return None


def _ac_params(kind: str, args: list[str]):
    """netlist.rs:1454-1476: the numbers after the first ``AC`` of a V/I source."""
    if kind not in "VI" or kind == "?":
        return None, None
    idx = next((i for i, a in enumerate(args) if _upper(a) == "AC"), None)
    if idx is None:
        return None, None
    mag = _number_or_plain(args[idx + 1]) if idx + 1 < len(args) else None
    ph = _number_or_plain(args[idx + 2]) if idx + 2 < len(args) else None
    return mag, ph


def _fields(nodes, model, control, value, extras, poly) -> str:
    """netlist.rs:1491-1523."""
    parts = []
    if nodes:
        parts.append("nodes=" + ",".join(nodes))
    if model is not None:
        parts.append(f"model={model}")
    if control is not None:
        parts.append(f"ctrl={control}")
    if value is not None:
        parts.append(f"value={value}")
    if extras:
        parts.append("extras=" + ",".join(extras))
    if poly is not None:
        parts.append(f"poly=deg{poly.degree} coeffs={len(poly.coeffs)}")
    return "[" + " ".join(parts) + "]" if parts else ""


def _validate(name, kind, nodes, model, control, value, extras, poly, line_no, errors) -> None:
    """netlist.rs:583-856: the field-count and missing-field checks."""
    if kind == "?":
        return
    ff = _fields(nodes, model, control, value, extras, poly)

    def err(msg):
        errors.append(ParseError(line_no, msg))

    if not nodes:
        err(f"device missing node definitions: {name} {ff}")
        return
    has_wave = any(_is_waveform(t) for t in extras)
    if kind in "RCLVI":
        if len(nodes) != 2:
            err(f"{name} requires 2 nodes, got={len(nodes)} {ff}")
        if value is None and not (kind in "VI" and has_wave):
            err(f"{name} missing value {ff}")
        if kind in "RCL" and extras:
            err(f"{name} has extra fields {ff}")
        if kind in "VI" and extras and not has_wave:
            err(f"{name} has extra fields {ff}")
        if kind in "VI" and extras:
            wave = _upper(extras[0])
            if wave in ("AC", "SIN", "PULSE", "EXP", "SFFM", "PWL") and len(extras) == 1:
                err(f"{name} waveform {wave} missing parameters {ff}")
    elif kind == "D":
        if len(nodes) != 2:
            err(f"{name} requires 2 nodes, got={len(nodes)} {ff}")
        if model is None:
            err(f"{name} missing model name {ff}")
    elif kind == "M":
        if len(nodes) < 4:
            err(f"{name} requires at least 4 nodes, got={len(nodes)} {ff}")
        if model is None:
            err(f"{name} missing model name {ff}")
    elif kind in "EG":
        if poly is not None:
            if len(nodes) != 2:
                err(f"{name} POLY requires 2 output nodes, got={len(nodes)} {ff}")
        elif len(nodes) != 4:
            err(f"{name} requires 4 nodes, got={len(nodes)} {ff}")
        if value is None and poly is None:
            err(f"{name} missing gain value {ff}")
        if poly is None and extras:
            err(f"{name} has extra fields {ff}")
        if poly is not None:
            _check_poly(name, poly, poly.degree * 2, "control nodes", ff, err)
    elif kind in "FH":
        if len(nodes) != 2:
            err(f"{name} requires 2 nodes, got={len(nodes)} {ff}")
        if control is None and poly is None:
            err(f"{name} missing control source {ff}")
        if value is None and poly is None:
            err(f"{name} missing gain value {ff}")
        if poly is None and extras:
            err(f"{name} has extra fields {ff}")
        if poly is not None:
            _check_poly(name, poly, poly.degree, "control sources", ff, err)
    elif kind == "X":
        if model is None:
            err(f"{name} missing subcircuit name {ff}")


def _check_poly(name, poly, expected, what, ff, err) -> None:
    if len(poly.coeffs) < expected:
        err(f"{name} POLY insufficient {what}, expected={expected} got={len(poly.coeffs)} {ff}")
    elif len(poly.coeffs) - expected < poly.degree + 1:
        err(f"{name} POLY insufficient coefficients, expected>={poly.degree + 1} "
            f"got={len(poly.coeffs) - expected} {ff}")


# ---------------------------------------------------------------------------
# Expressions (netlist.rs:1557-1841)
# ---------------------------------------------------------------------------
def _tokenize(expr: str) -> list:
    toks, buf = [], []

    def flush():
        if buf:
            s = "".join(buf)
            v = parse_number_with_suffix(s)
            toks.append(("num", v) if v is not None else ("id", s))
            buf.clear()

    for ch in expr:
        if ch in _WS:
            flush()
        elif ch in "+-*/^":
            flush()
            toks.append(("op", ch))
        elif ch == ",":
            flush()
            toks.append(("comma", None))
        elif ch == "(":
            flush()
            toks.append(("lp", None))
        elif ch == ")":
            flush()
            toks.append(("rp", None))
        else:
            buf.append(ch)
    flush()
    return toks


_PREC = {"+": 1, "-": 1, "*": 2, "/": 2, "^": 3}


def _to_rpn(toks: list):
    """Shunting-yard (netlist.rs:1681-1765): a ``-`` with no value before it
    pushes 0 first; every operator is left-associative; ``f(`` makes a call
    whose argument count is the commas seen plus one."""
    out, ops, argc = [], [], []
    prev_value = False
    for i, (kind, v) in enumerate(toks):
        nxt = toks[i + 1][0] if i + 1 < len(toks) else None
        if kind == "num":
            out.append((kind, v))
        elif kind == "id":
            if nxt == "lp":
                ops.append(("func", v))
            else:
                out.append((kind, v))
        elif kind == "op":
            if not prev_value and v == "-":
                out.append(("num", 0.0))
            while ops and ops[-1][0] == "op" and _PREC[ops[-1][1]] >= _PREC[v]:
                out.append(ops.pop())
            ops.append((kind, v))
        elif kind == "comma":
            while ops and ops[-1][0] != "lp":
                out.append(ops.pop())
            if argc:
                argc[-1] += 1
        elif kind == "lp":
            if ops and ops[-1][0] == "func":
                argc.append(0)
            ops.append((kind, None))
        elif kind == "rp":
            while ops:
                top = ops.pop()
                if top[0] == "lp":
                    break
                out.append(top)
            if ops and ops[-1][0] == "func":
                n = (argc.pop() if argc else 0) + 1
                out.append(("call", (ops.pop()[1], n)))
        prev_value = kind in ("num", "id", "rp")
    while ops:
        top = ops.pop()
        if top[0] in ("lp", "rp"):
            return None
        out.append(top)
    return out


def _fdiv(a: float, b: float) -> float:
    if b == 0.0:
        if a == 0.0 or a != a:
            return math.nan
        return math.copysign(math.inf, a) * math.copysign(1.0, b)
    return a / b


def _powf(a: float, b: float) -> float:
    try:
        return math.pow(a, b)
    except ValueError:
        return math.nan
    except OverflowError:
        neg = a < 0 and float(b).is_integer() and int(b) % 2 == 1
        return -math.inf if neg else math.inf


def _fmax(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return a if a > b else b


def _fmin(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return a if a < b else b


def _call(name: str, args: list):
    """netlist.rs:1827-1841."""
    n = _lower(name)
    if n == "max" and len(args) == 2:
        return _fmax(*args)
    if n == "min" and len(args) == 2:
        return _fmin(*args)
    if n == "abs" and len(args) == 1:
        return abs(args[0])
    if n == "if" and len(args) == 3:
        return args[1] if args[0] != 0.0 else args[2]
    return None


def eval_expression(expr: str, local: dict, parent: dict | None = None, glob: dict | None = None):
    """eval_expression_scoped (netlist.rs:1564-1587): identifiers are looked up
    (lower-cased) in local, parent, then global, and read as numbers."""
    parent = local if parent is None else parent
    glob = local if glob is None else glob
    toks = _tokenize(expr)
    if not toks:
        return None
    rpn = _to_rpn(toks)
    if rpn is None:
        return None
    st = []
    for kind, v in rpn:
        if kind == "num":
            st.append(v)
        elif kind == "id":
            k = _lower(v)
            s = local.get(k)
            if s is None:
                s = parent.get(k)
            if s is None:
                s = glob.get(k)
            x = _number_or_plain(s) if s is not None else None
            if x is None:
                return None
            st.append(x)
        elif kind == "op":
            if len(st) < 2:
                return None
            b, a = st.pop(), st.pop()
            st.append({"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b,
                       "/": lambda: _fdiv(a, b), "^": lambda: _powf(a, b)}[v]())
        elif kind == "call":
            name, n = v
            if len(st) < n:
                return None
            args = st[len(st) - n:]
            del st[len(st) - n:]
            r = _call(name, args)
            if r is None:
                return None
            st.append(r)
        else:
            return None
    return st[0] if len(st) == 1 else None


# ---------------------------------------------------------------------------
# Elaboration: .param and .subckt (netlist.rs:858-932, 1266-1439, 1861-1952)
# ---------------------------------------------------------------------------
@dataclass
class SubcktDef:
    name: str
    ports: list
    params: list
    body: list
    line: int


@dataclass
class ElaboratedNetlist:
    instances: list
    subckt_models: list
    control_count: int
    error_count: int
    errors: list


def _extract_subckts(stmts: list):
    top, defs, errors = [], [], []
    i = 0
    while i < len(stmts):
        s = stmts[i]
        if isinstance(s, ControlStmt) and s.kind == "subckt":
            name = s.subckt_name if s.subckt_name is not None else "unknown"
            i += 1
            body, found = [], False
            while i < len(stmts):
                b = stmts[i]
                i += 1
                if isinstance(b, ControlStmt) and b.kind == "ends":
                    found = True
                    break
                body.append(b)
            if not found:
                errors.append(ParseError(s.line, f"subckt {name} missing .ends"))
            defs.append(SubcktDef(name, list(s.subckt_ports), list(s.params), body, s.line))
        else:
            top.append(s)
            i += 1
    return top, defs, errors


def _eval_to_text(expr: str, local, parent, glob) -> str:
    v = eval_expression(expr, local, parent, glob)
    return rust_f64_display(v) if v is not None else expr


def _param_table(stmts: list) -> dict:
    """netlist.rs:1266-1282: ``.param`` values evaluated in order."""
    params: dict = {}
    for s in stmts:
        if isinstance(s, ControlStmt) and s.kind == "param":
            for p in s.params:
                params[_lower(p.key)] = _eval_to_text(p.value, params, params, params)
    return params


def _body_params(body: list) -> list:
    """netlist.rs:1407-1427: the body's own ``.param`` lines (not nested ones)."""
    out, depth = [], 0
    for s in body:
        if isinstance(s, ControlStmt):
            if s.kind == "subckt":
                depth += 1
            elif s.kind == "ends":
                depth = max(depth - 1, 0)
            elif s.kind == "param" and depth == 0:
                out.extend(s.params)
    return out


def _local_params(d: SubcktDef, inst: DeviceStmt, body_params: list, parent: dict, glob: dict) -> dict:
    """netlist.rs:1284-1314: definition defaults, then body ``.param``, then the
    instance's own parameters, each evaluated against what came before."""
    params: dict = {}
    for p in list(d.params) + list(body_params) + list(inst.params):
        params[_lower(p.key)] = _eval_to_text(p.value, params, parent, glob)
    return params


def _resolve(local: dict, glob: dict, token: str):
    k = _lower(token)
    if k in local:
        return local[k]
    if k in glob:
        return glob[k]
    v = eval_expression(token, local, {}, glob)
    return rust_f64_display(v) if v is not None else None


def _apply_params(glob: dict, local: dict, dev: DeviceStmt) -> None:
    """netlist.rs:1316-1336: value, model name and parameter values resolved."""
    if dev.value is not None:
        r = _resolve(local, glob, dev.value)
        if r is not None:
            dev.value = r
    if dev.model is not None:
        r = _resolve(local, glob, dev.model)
        if r is not None:
            dev.model = r
    for p in dev.params:
        r = _resolve(local, glob, p.value)
        if r is not None:
            p.value = r


def _expand(inst: DeviceStmt, d: SubcktDef, defs: dict, local: dict, glob: dict, errors: list,
            models: list) -> list:
    """netlist.rs:1869-1952."""
    body, nested, nerr = _extract_subckts(d.body)
    errors.extend(nerr)
    nmap = {nd.name: nd for nd in nested}
    for k, v in defs.items():
        nmap.setdefault(k, v)
    ports = {}
    for port, node in zip(d.ports, inst.nodes):
        ports[_lower(port)] = node
    out = []
    for s in body:
        if isinstance(s, DeviceStmt):
            sc = s.copy()
            sc.name = f"{inst.name}.{s.name}"
            sc.nodes = [ports.get(_lower(nd), f"{inst.name}:{nd}") for nd in s.nodes]
            if sc.kind == "X":
                child = nmap.get(sc.model) if sc.model is not None else None
                if child is not None:
                    cp = _local_params(child, sc, _body_params(child.body), local, glob)
                    out.extend(_expand(sc, child, nmap, cp, glob, errors, models))
                    continue
                errors.append(ParseError(sc.line, f"subcircuit not defined: {sc.model!r}"))
            _apply_params(glob, local, sc)
            out.append(sc)
        elif isinstance(s, ControlStmt) and s.kind == "model":
            m = ControlStmt(s.command, s.kind, list(s.args), list(s.params), s.model_name, s.model_type,
                            s.subckt_name, list(s.subckt_ports), s.raw, s.line)
            if m.model_name is not None:
                m.model_name = f"{inst.name}.{m.model_name}"
            models.append(m)
    return out


def elaborate_netlist(ast: NetlistAst) -> ElaboratedNetlist:
    """netlist.rs:858-932."""
    errors = list(ast.errors)
    top, defs, serr = _extract_subckts(ast.statements)
    errors.extend(serr)
    glob = _param_table(top)
    dmap = {d.name: d for d in defs}
    instances, models, controls = [], [], 0
    for s in top:
        if isinstance(s, DeviceStmt):
            if s.kind == "X":
                d = dmap.get(s.model) if s.model is not None else None
                if d is not None:
                    lp = _local_params(d, s, _body_params(d.body), {}, glob)
                    instances.extend(_expand(s, d, dmap, lp, glob, errors, models))
                    continue
                errors.append(ParseError(s.line, f"subcircuit not defined: {s.model!r}"))
            inst = s.copy()
            _apply_params(glob, {}, inst)
            instances.append(inst)
        elif isinstance(s, ControlStmt):
            controls += 1
    return ElaboratedNetlist(instances, models, controls, len(errors), errors)


# ---------------------------------------------------------------------------
# Circuit (netlist.rs:934-1211, circuit.rs, options.rs)
# ---------------------------------------------------------------------------
#: option_defs (options.rs:39-110): name -> (type, default, range)
OPTION_DEFS = {
    "abstol": ("float", 1e-12, (0.0, 1e-3)),
    "reltol": ("float", 1e-3, (0.0, 1.0)),
    "vntol": ("float", 1e-6, (0.0, 1.0)),
    "gmin": ("float", 1e-12, (0.0, 1e-3)),
    "itl1": ("int", 100, (1, 10000)),
    "itl4": ("int", 50, (1, 10000)),
    "temp": ("float", 27.0, (-273.15, 1000.0)),
    "tnom": ("float", 27.0, (-273.15, 1000.0)),
    "solver": ("str", "auto", ("auto", "dense", "sparse", "sparselu", "sparselubtf", "bbd", "faer",
                               "klu", "nativeklu")),
    "solver_parallel": ("int", 0, (0, 256)),
}


class SimOptions:
    """options.rs:112-330: typed options with range checks; a value that does
    not parse or is out of range leaves the option as it was."""

    def __init__(self):
        self.values = {k: d[1] for k, d in OPTION_DEFS.items()}
        self.is_set = set()
        self.warnings: list[str] = []

    def set(self, key: str, raw: str) -> None:
        k = _lower(key)
        d = OPTION_DEFS.get(k)
        if d is None:
            self.warnings.append(f"unknown option '{key}' ignored")
            return
        typ, _, rng = d
        if typ == "float":
            v = _number_or_plain(raw)
            ok = v is not None and rng[0] < v < rng[1]
        elif typ == "int":
            v = _rust_i64(raw)
            ok = v is not None and rng[0] <= v <= rng[1]
        else:
            v = raw
            ok = _lower(v) in rng
        if not ok:
            self.warnings.append(f"option '{k}' value '{raw}' rejected")
            return
        self.values[k] = v
        self.is_set.add(k)


@dataclass
class Model:
    name: str
    model_type: str
    params: dict


@dataclass
class Instance:
    name: str
    kind: str
    nodes: list  # node ids
    model: int | None
    params: dict
    value: str | None
    control: str | None
    ac_mag: float | None
    ac_phase: float | None
    poly: dict | None


@dataclass
class SpiceCircuit:
    """crates/sim-core/src/circuit.rs:120-216: node table, models, instances,
    analyses, options, initial conditions."""
    node_names: list
    node_ids: dict
    models: list
    model_ids: dict
    instances: list
    analyses: list
    options: SimOptions
    initial_conditions: dict
    va_files: list
    title: str | None = None

    def ensure_node(self, name: str) -> int:
        i = self.node_ids.get(name)
        if i is None:
            i = len(self.node_names)
            self.node_ids[name] = i
            self.node_names.append(name)
        return i

    @property
    def node_count(self) -> int:
        return len(self.node_names)


def _model_entry(c: ControlStmt) -> Model:
    return Model(_lower(c.model_name), _lower(c.model_type), {_lower(p.key): p.value for p in c.params})


def _insert_model(circ: SpiceCircuit, m: Model) -> None:
    circ.model_ids[m.name] = len(circ.models)
    circ.models.append(m)


def build_circuit(ast: NetlistAst, elab: ElaboratedNetlist) -> SpiceCircuit:
    """netlist.rs:934-1211. Every control statement of the file (subcircuit
    bodies included) is read first, then the elaborated devices in order."""
    circ = SpiceCircuit([], {}, [], {}, [], [], SimOptions(), {}, [], ast.title)
    circ.ensure_node("0")
    for s in ast.statements:
        if not isinstance(s, ControlStmt):
            continue
        a = s.args
        if s.kind == "model":
            if s.model_name is not None and s.model_type is not None:
                _insert_model(circ, _model_entry(s))
        elif s.kind == "op":
            circ.analyses.append(("op",))
        elif s.kind == "dc":
            if len(a) >= 4:
                vals = [_number_or_plain(t) for t in a[1:4]]
                circ.analyses.append(("dc", a[0]) + tuple(0.0 if v is None else v for v in vals))
        elif s.kind == "tran":
            if len(a) >= 2:
                tstep = _number_or_plain(a[0])
                tstop = _number_or_plain(a[1])
                tstep = 0.0 if tstep is None else tstep
                tstop = 0.0 if tstop is None else tstop
                tstart = _number_or_plain(a[2]) if len(a) > 2 else None
                tmax = _number_or_plain(a[3]) if len(a) > 3 else None
                circ.analyses.append(("tran", tstep, tstop, 0.0 if tstart is None else tstart,
                                      tstop if tmax is None else tmax))
        elif s.kind == "ac":
            if len(a) >= 4:
                sweep = {"dec": "dec", "oct": "oct", "lin": "lin"}.get(_lower(a[0]), "dec")
                pts = _rust_usize(a[1])
                f0 = _number_or_plain(a[2])
                f1 = _number_or_plain(a[3])
                circ.analyses.append(("ac", sweep, 10 if pts is None else pts, 1.0 if f0 is None else f0,
                                      1e6 if f1 is None else f1))
        elif s.kind == "ic":
            for p in s.params:
                k = _trim(_lower(p.key))
                if not (k.startswith("v(") and k.endswith(")")):
                    continue
                node = _trim(k[2:-1])
                v = parse_number_with_suffix(p.value)
                if node and v is not None:
                    circ.initial_conditions[circ.ensure_node(node)] = v
        elif s.kind == "option":
            for p in s.params:
                circ.options.set(p.key, p.value)
            for t in a:
                circ.options.set(t, "true")
        elif s.kind in ("hdl", "osdi"):
            if a:
                circ.va_files.append(a[0].strip('"').strip("'"))
    for m in elab.subckt_models:
        if m.model_name is not None and m.model_type is not None:
            _insert_model(circ, _model_entry(m))
    for d in elab.instances:
        if d.kind == "?":
            continue
        nodes = [circ.ensure_node(nm) for nm in d.nodes]
        mid = circ.model_ids.get(_lower(d.model)) if d.model is not None else None
        params = dict(circ.models[mid].params) if mid is not None else {}
        for p in d.params:
            params[_lower(p.key)] = p.value
        poly = None
        if d.poly is not None:
            if d.kind in "EG":
                nc = d.poly.degree * 2
                ctl, coeffs = [], []
                for i, item in enumerate(d.poly.coeffs):
                    if i < nc:
                        nid = circ.ensure_node(item)
                        if i % 2 == 1:
                            ctl.append((circ.ensure_node(d.poly.coeffs[i - 1]), nid))
                    else:
                        v = _number_or_plain(item)
                        if v is not None:
                            coeffs.append(v)
                poly = dict(degree=d.poly.degree, coeffs=coeffs, control_nodes=ctl, control_sources=[])
            elif d.kind in "FH":
                ns = d.poly.degree
                srcs = [it for i, it in enumerate(d.poly.coeffs) if i < ns]
                coeffs = [v for v in (_number_or_plain(it) for i, it in enumerate(d.poly.coeffs) if i >= ns)
                          if v is not None]
                poly = dict(degree=d.poly.degree, coeffs=coeffs, control_nodes=[], control_sources=srcs)
        circ.instances.append(Instance(d.name, d.kind, nodes, mid, params, d.value, d.control,
                                       d.ac_mag, d.ac_phase, poly))
    if not circ.analyses:
        circ.analyses.append(("op",))
    return circ


def load(text: str | None = None, path: str | None = None) -> SpiceCircuit:
    """Parse, elaborate and build a netlist given as text or as a file path."""
    if (text is None) == (path is None):
        raise ValueError("give the netlist text or its path")
    ast = parse_netlist_file(path) if path is not None else parse_netlist(text)
    circ = build_circuit(ast, elaborate_netlist(ast))
    circ.parse_errors = list(ast.errors)
    return circ


# ---------------------------------------------------------------------------
# Lowering to rs_device records (stamp.rs)
# ---------------------------------------------------------------------------
class UnsupportedDevice(ValueError):
    """A device the GPU stamp path does not evaluate (L, E, G, F, H, VA)."""


def _param_value(params: dict, keys) -> float | None:
    """stamp.rs:749-757."""
    for k in keys:
        s = params.get(_lower(k))
        if s is not None:
            v = _number_or_plain(s)
            if v is not None:
                return v
    return None


def _as_u32(v: float) -> int:
    """``f64 as u32``: truncation, saturating, NaN to 0."""
    if v != v:
        return 0
    return int(min(max(v, 0.0), 4294967295.0))


def _is_pmos(params: dict) -> bool:
    """stamp.rs:403-410."""
    t = params.get("type")
    if t is not None:
        t = _lower(t)
        return "pmos" in t or t == "p"
    return "pmos" in params


@dataclass
class Lowered:
    """The records one netlist stamps as, and what was left out."""
    circuit: NL.Circuit
    node_names: list
    skipped: list  # (instance name, reason): devices the reference stamps nothing for
    source_index: dict  # V/I instance name -> record index
    ac_sources: dict  # record index -> (ac_mag, ac_phase) of V/I sources
    record_of: dict  # instance index -> record index (instances that stamp)
    ac_circuit: NL.Circuit = None  # the records the AC stamps read (stamp.rs:1525-1747)
    ac_mag: np.ndarray = None  # per AC record: ac_mag of a V/I source (0 when none)
    ac_phase: np.ndarray = None  # per AC record: ac_phase in degrees (0 when none)


def lower(circ: SpiceCircuit) -> Lowered:
    """One rs_device per instance the stamp loop would stamp, in instance order.
    A device whose stamp returns an error (missing or unparsable value, wrong
    node count: stamp.rs:242-349) stamps nothing; the engine ignores the error
    (engine.rs:262-269), so the device is listed in ``skipped``."""
    b = NL._Builder()
    ba = NL._Builder()  # AC: a V/I source stamps whether or not its value parses
    amag, aph = [], []
    skipped, src_index, ac, rec_of = [], {}, {}, {}
    models: list = []
    model_key: dict = {}
    seen_named = set()
    ac_v_names: list = []

    def one(inst):
        k = inst.kind
        if k in ("L", "E", "G", "F", "H"):
            raise UnsupportedDevice(f"{inst.name}: device kind {k} is not evaluated by the GPU stamp path")
        if k == "X":
            skipped.append((inst.name, "subcircuit instance left unexpanded"))
            return
        nd = inst.nodes
        if k in "RCVI":
            if len(nd) != 2:
                skipped.append((inst.name, "needs 2 nodes"))
                return
            v = parse_number_with_suffix(inst.value) if inst.value is not None else None
            if k in "VI":
                if k == "V" and any(nm == inst.name for nm in ac_v_names):
                    raise UnsupportedDevice(f"{inst.name}: a repeated V name shares its aux row in the reference")
                if k == "V":
                    ac_v_names.append(inst.name)
                ba.add(NL.RS_DEV_V if k == "V" else NL.RS_DEV_I, nd, 0.0 if v is None else v)
                amag.append(0.0 if inst.ac_mag is None else inst.ac_mag)
                aph.append(0.0 if inst.ac_phase is None else inst.ac_phase)
            if v is None:
                skipped.append((inst.name, "no numeric value (waveforms are not passed on by the parser)"))
                return
            if k in "RC":
                ba.add(NL.RS_DEV_R if k == "R" else NL.RS_DEV_C, nd, v)
                amag.append(0.0)
                aph.append(0.0)
            if k in "CV":
                # state (capacitor history) and aux rows are keyed by name in the reference
                if (k, inst.name) in seen_named:
                    raise UnsupportedDevice(f"{inst.name}: a repeated {k} name shares state in the reference")
                seen_named.add((k, inst.name))
            kind = {"R": NL.RS_DEV_R, "C": NL.RS_DEV_C, "V": NL.RS_DEV_V, "I": NL.RS_DEV_I}[k]
            if k in "VI":
                src_index[inst.name] = len(b.rows)
                if inst.ac_mag is not None:
                    ac[len(b.rows)] = (inst.ac_mag, 0.0 if inst.ac_phase is None else inst.ac_phase)
            b.add(kind, nd, v)
        elif k == "D":
            if len(nd) != 2:
                skipped.append((inst.name, "needs 2 nodes"))
                return
            isat = _param_value(inst.params, ["is"])
            em = _param_value(inst.params, ["n", "nj"])
            b.add(NL.RS_DEV_D, nd, 0.0, [1e-14 if isat is None else isat, 1.0 if em is None else em])
            ba.rows.append(b.rows[-1])
            amag.append(0.0)
            aph.append(0.0)
        elif k == "M":
            if len(nd) < 4:
                skipped.append((inst.name, "needs 4 nodes"))
                return
            lv = _param_value(inst.params, ["level"])
            level = 49 if lv is None else _as_u32(lv)
            pm = _is_pmos(inst.params)
            w = _param_value(inst.params, ["w"])
            ln = _param_value(inst.params, ["l"])
            w = 1e-6 if w is None else w
            ln = 1e-6 if ln is None else ln
            rec, _ = NL.build_bsim_params(inst.params, level, pm)
            flags = NL.RS_F_PMOS if pm else 0
            if level == 1:
                b.add(NL.RS_DEV_M, nd[:4], 0.0,
                      [1.0, float(rec["vth0"]), float(rec["u0"]), float(rec["tox"]), w, ln], flags)
                ba.rows.append(b.rows[-1])
                amag.append(0.0)
                aph.append(0.0)
            else:
                temp = _param_value(inst.params, ["temp"])
                sa = _param_value(inst.params, ["sa"])
                sb = _param_value(inst.params, ["sb"])
                rec = rec.copy()
                rec["sa"] = 0.0 if sa is None else sa
                rec["sb"] = 0.0 if sb is None else sb
                key = rec.tobytes()
                idx = model_key.get(key)
                if idx is None:
                    idx = model_key[key] = len(models)
                    models.append(rec)
                b.add(NL.RS_DEV_M, nd[:4], 0.0,
                      NL.bsim_params(level, idx, w, ln, 300.15 if temp is None else temp), flags)
                ba.rows.append(b.rows[-1])
                amag.append(0.0)
                aph.append(0.0)
        else:
            skipped.append((inst.name, f"unknown device kind {k}"))

    for ii, inst in enumerate(circ.instances):
        before = len(b.rows)
        one(inst)
        if len(b.rows) > before:
            rec_of[ii] = before
    c = b.build(circ.node_count, circ.title or "netlist", models if models else None)
    ca = ba.build(circ.node_count, (circ.title or "netlist") + " (ac)", models if models else None)
    return Lowered(c, list(circ.node_names), skipped, src_index, ac, rec_of, ca,
                   np.array(amag, dtype=np.float64), np.array(aph, dtype=np.float64))


#: the solver kinds of rustspice_amd.solver
_SOLVER_KINDS = {"nativeklu": 0, "sparse": 1, "sparselu": 1, "bbd": 2, "dense": 3, "klu": 1}


def solver_kind(circ: SpiceCircuit) -> int:
    """solver_type_from_options (engine.rs:972-984) and create_solver
    (solver.rs:169-205) for the kinds this library provides. ``klu`` without
    SuiteSparse is SparseLU (solver.rs:185-195); ``auto`` is Dense up to 50
    nodes (solver.rs:222-225); beyond that the reference picks faer, and
    ``faer``/``sparselubtf`` are not provided here."""
    s = _lower(str(circ.options.values["solver"]))
    if s in _SOLVER_KINDS:
        return _SOLVER_KINDS[s]
    if s not in ("faer", "sparselubtf") and circ.node_count <= 50:
        return 3
    raise UnsupportedDevice(f"solver '{s}' (auto beyond 50 nodes picks faer) is not provided by the GPU "
                            "library: set .option solver=nativeklu, sparselu, bbd or dense")


def newton_options(circ: SpiceCircuit) -> dict:
    """newton_config_from_options (engine.rs:987-999)."""
    o = circ.options.values
    return dict(max_iters=int(o["itl1"]), abs_tol=float(o["vntol"]), rel_tol=float(o["reltol"]),
                gmin=float(o["gmin"]))


def to_numpy_ic(circ: SpiceCircuit) -> np.ndarray | None:
    """The ``.ic`` values as a node_count vector (the TRAN's initial guess,
    engine.rs:343-348), or None when the netlist has none."""
    if not circ.initial_conditions:
        return None
    x = np.zeros(circ.node_count)
    for nid, v in circ.initial_conditions.items():
        if nid < circ.node_count:
            x[nid] = v
    return x
