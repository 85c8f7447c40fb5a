"""NMODL -> HIP/C++ generator for Hodgkin-Huxley-style mechanisms (cnexp).

The reference turns a model's .mod files into CUDA device code with its own
parser and emitter (PYS/extractModel.py:1905-2000 reads the mod files'
blocks, :2003- writes the per-mechanism state and current functions into
AllModels.cu).  This module does the same job for the subset of NMODL that
HH-style membrane mechanisms use, and emits host- and device-callable C++:

* blocks NEURON (SUFFIX, USEION ... READ ... WRITE ..., NONSPECIFIC_CURRENT,
  RANGE, GLOBAL), UNITS (named physical constants such as FARADAY), PARAMETER
  (defaults), CONSTANT, STATE, ASSIGNED, BREAKPOINT (SOLVE states METHOD
  cnexp, assignments), DERIVATIVE states, INITIAL, PROCEDURE and FUNCTION;
* statements: assignments, ``x' = expr``, LOCAL, if / else if / else,
  procedure calls; UNITSOFF / UNITSON lines and TABLE statements (with their
  DEPEND and FROM ... TO ... WITH lines) are ignored, units after numbers in
  expressions (``10 (degC)``) dropped; VERBATIM, KINETIC, NET_RECEIVE and
  other solve methods are refused.

Each mechanism becomes a struct whose fields are its variables (v, dt,
PARAMETER defaults, STATE, ASSIGNED, ion variables) and whose member
functions are its blocks, so that PROCEDUREs act on the instance as in
NEURON.  ``x' = expr`` must be linear in x (A + B x); it is integrated with
NEURON's cnexp formula x += (1 - exp(dt B)) (-A/B - x), which is the form of
the reference's generated code (e.g. Code/OneKernel/src/AllModels.cu:302).
exp, log and pow are ngmath.h's ng_exp, ng_log and ng_pow (the same on the
host and the device).

    python -m neurogpu_amd.nmodl -o gen_mechs.h a.mod b.mod ...

also writes a C-ABI evaluation entry per mechanism (``--harness``) that the
tests compile on the host.  ``--set NAME`` adds SetGen_NAME, a mechanism set
the simulation kernel instantiates (emit_set: mask dispatch, state slots,
the parameter-matrix layout, calcium bookkeeping), with the model's
temperature (``--celsius``), constant ion values (``--ion``), constant
parameters (``--constant``, ``--global name_suffix=value``) and specialised
masks (``--masks``).  ``--fast`` emits the fast forms (fast_expr,
fast_cnexp, _inline_functions, _ExpShare) instead of the reference's
operation order; both agree to rounding.
"""
from __future__ import annotations

import argparse
import ast
import copy
import dataclasses
import json
import math
import re
import sys
from typing import Dict, List, Optional, Tuple

# physical constants of the UNITS block, as NEURON's units database defines them
UNIT_CONSTANTS = {"faraday": 96485.309, "k-mole": 8.313424, "R": 8.313424}
MATH = {"exp": "ng_exp", "log": "ng_log", "pow": "ng_pow", "fabs": "fabs", "sqrt": "sqrt", "tanh": "tanh",
        "sin": "sin", "cos": "cos"}


class NmodlError(ValueError):
    pass


@dataclasses.dataclass
class Block:
    kind: str
    name: str
    args: List[str]
    body: str


@dataclasses.dataclass
class Mechanism:
    suffix: str
    params: Dict[str, Optional[float]]          # PARAMETER (None: no default, e.g. celsius)
    constants: Dict[str, float]                 # UNITS and CONSTANT
    states: List[str]
    assigned: List[str]
    ions: Dict[str, Tuple[List[str], List[str]]]   # ion -> (READ, WRITE)
    nonspecific: List[str]
    ranges: List[str]
    blocks: Dict[str, Block]                    # BREAKPOINT, INITIAL, DERIVATIVE <name>, PROCEDURE <name>, FUNCTION <name>
    solve: Optional[str]                        # the DERIVATIVE block BREAKPOINT solves

    @property
    def currents(self) -> List[str]:
        out = list(self.nonspecific)
        for ion, (_, w) in self.ions.items():
            out += [x for x in w if x == "i" + ion]
        return out


def strip_comments(text: str) -> str:
    text = re.sub(r"COMMENT.*?ENDCOMMENT", "", text, flags=re.S)
    return re.sub(r":[^\n]*", "", text)


def _match_brace(text: str, i: int) -> int:
    """Index just past the brace that closes the one before position i."""
    depth = 1
    while depth:
        if i >= len(text):
            raise NmodlError("unbalanced braces")
        depth += {"{": 1, "}": -1}.get(text[i], 0)
        i += 1
    return i


_HEAD = re.compile(r"\b(NEURON|UNITS|PARAMETER|CONSTANT|STATE|ASSIGNED|BREAKPOINT|INITIAL|INDEPENDENT|"
                   r"DERIVATIVE|PROCEDURE|FUNCTION|KINETIC|NET_RECEIVE|LINEAR|NONLINEAR)\b\s*(\w*)\s*")


def _balanced(text: str, i: int) -> int:
    """Index just past the parenthesis that closes the one at position i."""
    depth, j = 0, i
    while True:
        if j >= len(text):
            raise NmodlError("unbalanced parentheses")
        depth += {"(": 1, ")": -1}.get(text[j], 0)
        j += 1
        if depth == 0:
            return j


def blocks(text: str) -> List[Block]:
    """The top-level blocks: keyword, name, argument names (units dropped)
    and body."""
    out, i = [], 0
    while True:
        m = _HEAD.search(text, i)
        if not m:
            return out
        j, args = m.end(), []
        if j < len(text) and text[j] == "(":
            k = _balanced(text, j)
            inner = text[j + 1:k - 1]
            depth, cur, parts = 0, "", []
            for ch in inner:
                depth += {"(": 1, ")": -1}.get(ch, 0)
                if ch == "," and depth == 0:
                    parts.append(cur)
                    cur = ""
                else:
                    cur += ch
            parts.append(cur)
            args = [p.split("(")[0].strip() for p in parts if p.strip()]
            j = k
        while True:                                   # units of a FUNCTION's value, e.g. (mV)
            while j < len(text) and text[j] in " \t\r\n":
                j += 1
            if j < len(text) and text[j] == "(":
                j = _balanced(text, j)
            else:
                break
        if j >= len(text) or text[j] != "{":
            i = m.end()
            continue
        end = _match_brace(text, j + 1)
        out.append(Block(m.group(1), m.group(2), args, text[j + 1:end - 1]))
        i = end


def _names(body: str) -> List[str]:
    """Declared names of a STATE / ASSIGNED / PARAMETER-like block."""
    body = re.sub(r"\([^)]*\)", " ", body)           # units
    body = re.sub(r"<[^>]*>", " ", body)             # limits
    body = re.sub(r"=\s*[-+0-9.eE]+", " ", body)
    return [t for t in re.findall(r"[A-Za-z_]\w*", body) if t not in ("FROM", "TO")]


def parse(text: str) -> Mechanism:
    """A Mechanism from the text of a .mod file."""
    text = strip_comments(text)
    if "VERBATIM" in text:
        raise NmodlError("VERBATIM blocks are not supported")
    bl = blocks(text)
    suffix, ions, nonspec, ranges = None, {}, [], []
    params, consts, states, assigned, named = {}, {}, [], [], {}
    solve = None
    for b in bl:
        if b.kind == "NEURON":
            for line in b.body.split("\n"):
                line = line.strip()
                if line.startswith("SUFFIX"):
                    suffix = line.split()[1]
                elif line.startswith("USEION"):
                    m = re.match(r"USEION\s+(\w+)(?:\s+READ\s+([\w\s,]+?))?(?:\s+WRITE\s+([\w\s,]+?))?(?:\s+VALENCE.*)?$", line)
                    if not m:
                        raise NmodlError(f"cannot read {line!r}")
                    rd = [x.strip() for x in (m.group(2) or "").split(",") if x.strip()]
                    wr = [x.strip() for x in (m.group(3) or "").split(",") if x.strip()]
                    ions[m.group(1)] = (rd, wr)
                elif line.startswith("NONSPECIFIC_CURRENT"):
                    nonspec += [x.strip() for x in line.split(None, 1)[1].split(",")]
                elif line.startswith("RANGE") or line.startswith("GLOBAL"):
                    ranges += [x.strip() for x in line.split(None, 1)[1].split(",")]
        elif b.kind == "UNITS":
            for m in re.finditer(r"(\w+)\s*=\s*\(\s*([\w-]+)\s*\)", b.body):
                if m.group(2) in UNIT_CONSTANTS:
                    consts[m.group(1)] = UNIT_CONSTANTS[m.group(2)]
        elif b.kind in ("PARAMETER", "CONSTANT"):
            # name [= value] [(units)] [<lo, hi>], any number to a line
            body = re.sub(r"<[^>]*>", " ", re.sub(r"\([^)]*\)", " ", b.body))
            for m in re.finditer(r"([A-Za-z_]\w*)\s*(?:=\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?))?", body):
                val = float(m.group(2)) if m.group(2) is not None else None
                (consts if b.kind == "CONSTANT" else params)[m.group(1)] = val
        elif b.kind == "STATE":
            states += _names(b.body)
        elif b.kind == "ASSIGNED":
            assigned += _names(b.body)
        elif b.kind == "BREAKPOINT":
            named["BREAKPOINT"] = b
            m = re.search(r"SOLVE\s+(\w+)\s+METHOD\s+(\w+)", b.body)
            if m:
                if m.group(2) != "cnexp":
                    raise NmodlError(f"METHOD {m.group(2)} is not supported (cnexp only)")
                solve = m.group(1)
        elif b.kind == "INITIAL":
            named["INITIAL"] = b
        elif b.kind in ("DERIVATIVE", "PROCEDURE", "FUNCTION"):
            named[f"{b.kind} {b.name}"] = b
        elif b.kind in ("KINETIC", "NET_RECEIVE", "LINEAR", "NONLINEAR"):
            raise NmodlError(f"{b.kind} blocks are not supported")
    if suffix is None:
        raise NmodlError("no SUFFIX")
    return Mechanism(suffix, params, consts, states, assigned, ions, nonspec, ranges, named, solve)


# ------------------------------------------------------------------ statements
@dataclasses.dataclass
class Stmt:
    kind: str                     # assign, deriv, call, if, local
    target: str = ""
    expr: str = ""
    branches: Optional[list] = None   # if: [(cond or None, [Stmt])]


def _expr(e: str) -> str:
    return e.strip()


def statements(body: str) -> List[Stmt]:
    """The statements of a block body."""
    out, i, n = [], 0, len(body)
    while i < n:
        while i < n and body[i] in " \t\r\n;":
            i += 1
        if i >= n:
            break
        rest = body[i:]
        m = re.match(r"if\s*\(", rest)
        if m:
            branches = []
            j = i + m.end()
            while True:
                depth, k = 1, j
                while depth:
                    depth += {"(": 1, ")": -1}.get(body[k], 0)
                    k += 1
                cond = body[j:k - 1]
                b0 = body.index("{", k)
                b1 = _match_brace(body, b0 + 1)
                branches.append((cond, statements(body[b0 + 1:b1 - 1])))
                i = b1
                m2 = re.match(r"\s*else\s*if\s*\(", body[i:])
                if m2:
                    j = i + m2.end()
                    continue
                m3 = re.match(r"\s*else\s*\{", body[i:])
                if m3:
                    b0 = i + m3.end()
                    b1 = _match_brace(body, b0)
                    branches.append((None, statements(body[b0:b1 - 1])))
                    i = b1
                break
            out.append(Stmt("if", branches=branches))
            continue
        eol = body.find("\n", i)
        line = body[i:] if eol < 0 else body[i:eol]
        i = n if eol < 0 else eol + 1
        line = line.strip()
        if (not line or line in ("UNITSOFF", "UNITSON") or line.startswith("TABLE") or line.startswith("SOLVE")
                or line.startswith("DEPEND") or re.match(r"FROM\b.*\bTO\b.*\bWITH\b", line)):
            # TABLE (with its DEPEND and FROM ... TO ... WITH lines) asks NEURON
            # to interpolate; the equations are evaluated as written instead
            continue
        if line.startswith("LOCAL"):
            for v in line[5:].split(","):
                out.append(Stmt("local", target=v.strip()))
            continue
        m = re.match(r"([A-Za-z_]\w*)\s*'\s*=(.*)$", line)
        if m:
            out.append(Stmt("deriv", m.group(1), _expr(m.group(2))))
            continue
        m = re.match(r"([A-Za-z_]\w*)\s*=(.*)$", line)
        if m:
            out.append(Stmt("assign", m.group(1), _expr(m.group(2))))
            continue
        m = re.match(r"([A-Za-z_]\w*)\s*\((.*)\)\s*$", line)
        if m:
            out.append(Stmt("call", m.group(1), m.group(2)))
            continue
        raise NmodlError(f"cannot read statement {line!r}")
    return out


# ------------------------------------------------------------------ expressions
_NUM_UNITS = re.compile(r"(?<![A-Za-z_\d.])(\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)\s*\(\s*[A-Za-z][\w/ ]*\)")


def _py(e: str) -> ast.expr:
    e = _NUM_UNITS.sub(r"\1", e)                     # a unit after a number: 10 (degC)
    e = e.replace("^", "**")
    try:
        return ast.parse(e, mode="eval").body
    except SyntaxError as err:
        raise NmodlError(f"cannot read expression {e!r}: {err}") from None


_BIN = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}
_CMP = {ast.Eq: "==", ast.NotEq: "!=", ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">="}


def cxx(node: ast.expr) -> str:
    """C++ for an expression tree."""
    if isinstance(node, ast.BinOp):
        if isinstance(node.op, ast.Pow):
            return f"ng_pow({cxx(node.left)}, {cxx(node.right)})"
        return f"({cxx(node.left)} {_BIN[type(node.op)]} {cxx(node.right)})"
    if isinstance(node, ast.UnaryOp):
        if isinstance(node.op, ast.USub):
            return f"(-{cxx(node.operand)})"
        if isinstance(node.op, ast.UAdd):
            return cxx(node.operand)
        if isinstance(node.op, ast.Not):
            return f"(!{cxx(node.operand)})"
    if isinstance(node, ast.Constant):
        return repr(float(node.value))
    if isinstance(node, ast.Name):
        return node.id
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
        fn = MATH.get(node.func.id, node.func.id)
        return f"{fn}({', '.join(cxx(a) for a in node.args)})"
    if isinstance(node, ast.Compare) and len(node.ops) == 1:
        return f"({cxx(node.left)} {_CMP[type(node.ops[0])]} {cxx(node.comparators[0])})"
    if isinstance(node, ast.BoolOp):
        op = " && " if isinstance(node.op, ast.And) else " || "
        return "(" + op.join(cxx(v) for v in node.values) + ")"
    raise NmodlError(f"unsupported expression {ast.dump(node)}")


def _cond(e: str) -> ast.expr:
    e = re.sub(r"(?<![=!<>])=(?!=)", "==", e)
    e = e.replace("&&", " and ").replace("||", " or ")
    return _py(e)


def _mentions(node: ast.expr, name: str) -> bool:
    return any(isinstance(n, ast.Name) and n.id == name for n in ast.walk(node))


def _is(node: ast.expr, c: float) -> bool:
    return isinstance(node, ast.Constant) and float(node.value) == c


def _add(a: ast.expr, op: ast.operator, b: ast.expr) -> ast.expr:
    """a + b or a - b, dropping an exact zero operand."""
    if _is(b, 0.0):
        return a
    if _is(a, 0.0):
        return b if isinstance(op, ast.Add) else ast.UnaryOp(ast.USub(), b)
    return ast.BinOp(a, op, b)


def _mul(a: ast.expr, op: ast.operator, b: ast.expr) -> ast.expr:
    """a * b or a / b, keeping an exact zero and dropping a unit factor."""
    if _is(a, 0.0) or (isinstance(op, ast.Mult) and _is(b, 0.0)):
        return ast.Constant(0.0)
    if _is(b, 1.0):
        return a
    if isinstance(op, ast.Mult) and _is(a, 1.0):
        return b
    return ast.BinOp(a, op, b)


def linear(node: ast.expr, x: str) -> Tuple[ast.expr, ast.expr]:
    """(A, B) with node = A + B x, for node linear in x (exact zeros and unit
    factors dropped, which changes no rounding)."""
    zero = ast.Constant(0.0)
    if not _mentions(node, x):
        return node, zero
    if isinstance(node, ast.Name):
        return zero, ast.Constant(1.0)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        a, b = linear(node.operand, x)
        if isinstance(node.op, ast.UAdd):
            return a, b
        return ast.UnaryOp(ast.USub(), a), ast.UnaryOp(ast.USub(), b)
    if isinstance(node, ast.BinOp):
        if isinstance(node.op, (ast.Add, ast.Sub)):
            a1, b1 = linear(node.left, x)
            a2, b2 = linear(node.right, x)
            return _add(a1, node.op, a2), _add(b1, node.op, b2)
        if isinstance(node.op, ast.Mult):
            if not _mentions(node.left, x):
                a, b = linear(node.right, x)
                return _mul(node.left, ast.Mult(), a), _mul(node.left, ast.Mult(), b)
            if not _mentions(node.right, x):
                a, b = linear(node.left, x)
                return _mul(a, ast.Mult(), node.right), _mul(b, ast.Mult(), node.right)
        if isinstance(node.op, ast.Div) and not _mentions(node.right, x):
            a, b = linear(node.left, x)
            return _mul(a, ast.Div(), node.right), _mul(b, ast.Div(), node.right)
    raise NmodlError(f"{ast.unparse(node)} is not linear in {x} (cnexp needs x' = A + B x)")


# ------------------------------------------------------------------ fast forms
# The fast emission (emit(..., fast=constants)) evaluates the same equations
# with the arrangements mechs.h's hand-written sets use on the FP64 VALU,
# where an IEEE division costs about ten instructions and an exp about twenty;
# results differ from the reference's operation order by rounding only:
#   - subexpressions of literals and known constants (celsius, and LOCALs
#     assigned such a value) are evaluated here;
#   - a division by a constant is a product with its reciprocal, any other
#     division a product with the refined reciprocal (ng_div);
#   - exp is ng_expc (no special-case branches), c^y for a constant c > 0 is
#     ng_expc(y log c);
#   - cnexp of x' = (X - x) / T is x += (1 - exp(-dt / T)) (X - x), with the
#     quotients that cancel in -A/B taken out.
_FOLD = {"exp": math.exp, "log": math.log, "fabs": abs, "sqrt": math.sqrt, "tanh": math.tanh, "sin": math.sin,
         "cos": math.cos, "ng_rcp": lambda x: 1.0 / x}


def _c(x: float) -> ast.expr:
    return ast.Constant(float(x))


def _num(node: ast.expr) -> Optional[float]:
    return float(node.value) if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) else None


def _finite(x) -> bool:
    return isinstance(x, float) and math.isfinite(x)


def _names_in(node: ast.expr) -> frozenset:
    return frozenset(n.id for n in ast.walk(node) if isinstance(n, ast.Name))


def _affine(node: ast.expr) -> Optional[Tuple[Dict[str, float], float, str, frozenset]]:
    """(coefficients, constant, divisor, the divisor's names) when node is an
    affine function of names with literal coefficients, or such a function
    divided by one expression (ng_div(N, D): the divisor is D, as its AST
    dump; "" for none), else None."""
    c = _num(node)
    if c is not None:
        return {}, c, "", frozenset()
    if isinstance(node, ast.Name):
        return {node.id: 1.0}, 0.0, "", frozenset()
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        f = _affine(node.operand)
        if f is None or isinstance(node.op, ast.UAdd):
            return f
        return {k: -c for k, c in f[0].items()}, -f[1], f[2], f[3]
    if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == "ng_div"
            and len(node.args) == 2):
        f = _affine(node.args[0])
        if f is None or f[2]:
            return None
        return f[0], f[1], ast.dump(node.args[1]), _names_in(node.args[1])
    if isinstance(node, ast.BinOp):
        if isinstance(node.op, (ast.Add, ast.Sub)):
            a, b = _affine(node.left), _affine(node.right)
            if a is None or b is None:
                return None
            if a[2] != b[2]:
                # a divided and an undivided part only combine if one is zero
                if not a[0] and a[1] == 0.0:
                    a = ({}, 0.0, b[2], b[3])
                elif not b[0] and b[1] == 0.0:
                    b = ({}, 0.0, a[2], a[3])
                else:
                    return None
            sg = 1.0 if isinstance(node.op, ast.Add) else -1.0
            co = dict(a[0])
            for k, c in b[0].items():
                co[k] = co.get(k, 0.0) + sg * c
            return co, a[1] + sg * b[1], a[2], a[3]
        if isinstance(node.op, ast.Mult):
            for x, y in ((node.left, node.right), (node.right, node.left)):
                if _num(y) is not None:
                    f = _affine(x)
                    if f is None:
                        return None
                    return {k: c * _num(y) for k, c in f[0].items()}, f[1] * _num(y), f[2], f[3]
    return None


def _exp_key(arg: ast.expr):
    """(key, key of the negated argument, names) of an exp argument that
    _affine reads and that depends on a name, else None."""
    f = _affine(arg)
    if f is None or not f[0]:
        return None
    key = (tuple(sorted(f[0].items())), f[1], f[2])
    neg = (tuple(sorted((k, -c) for k, c in f[0].items())), -f[1], f[2])
    return key, neg, set(f[0]) | set(f[3])


class _ExpShare:
    """exp(x) and exp(-x) of one statement list share one ng_expc: the second
    is the refined reciprocal of the first.  Arguments are matched as affine
    functions of names with literal coefficients, possibly over a common
    divisor (_affine); an entry ends when one of its names is assigned.
    ``pre`` collects the declarations of the shared values to emit before the
    statement being converted."""

    def __init__(self, parent: Optional["_ExpShare"] = None):
        self.table = dict(parent.table) if parent else {}
        self.count = parent.count if parent else [0]
        self.pre: List[str] = []

    def has(self, arg: ast.expr) -> bool:
        k = _exp_key(arg)
        return k is not None and (k[0] in self.table or k[1] in self.table)

    def exp(self, arg: ast.expr) -> ast.expr:
        k = _exp_key(arg)
        if k is None:
            return ast.Call(ast.Name("ng_expc"), [arg], [])
        key, neg, names = k
        if key in self.table:
            return ast.Name(self.table[key][0])
        if neg in self.table:
            return ast.Call(ast.Name("ng_rcp"), [ast.Name(self.table[neg][0])], [])
        name = f"exp_{self.count[0]}"
        self.count[0] += 1
        self.pre.append(f"const double {name} = ng_expc({cxx(arg)});")
        self.table[key] = (name, names)
        return ast.Name(name)

    def kill(self, target: str) -> None:
        self.table = {k: e for k, e in self.table.items() if target not in e[1]}


def fast_expr(node: ast.expr, consts: Dict[str, float], share: Optional[_ExpShare] = None) -> ast.expr:
    """The fast form of an expression tree (see above)."""
    def fe(n):
        return fast_expr(n, consts, share)

    if isinstance(node, ast.Name):
        return _c(consts[node.id]) if node.id in consts else node
    if isinstance(node, ast.Constant):
        return node
    if isinstance(node, ast.UnaryOp):
        o = fe(node.operand)
        if _num(o) is not None and isinstance(node.op, (ast.USub, ast.UAdd)):
            return _c(-_num(o) if isinstance(node.op, ast.USub) else _num(o))
        return ast.UnaryOp(node.op, o)
    if isinstance(node, ast.BinOp):
        l, r = fe(node.left), fe(node.right)
        lv, rv = _num(l), _num(r)
        if lv is not None and rv is not None:
            try:
                val = {ast.Add: lambda: lv + rv, ast.Sub: lambda: lv - rv, ast.Mult: lambda: lv * rv,
                       ast.Div: lambda: lv / rv, ast.Pow: lambda: math.pow(lv, rv)}[type(node.op)]()
            except (ZeroDivisionError, ValueError, OverflowError):
                val = None
            if _finite(val):
                return _c(val)
        if isinstance(node.op, ast.Div):
            if rv is not None and rv != 0.0 and _finite(1.0 / rv):
                return ast.BinOp(l, ast.Mult(), _c(1.0 / rv))
            if lv == 1.0:
                return ast.Call(ast.Name("ng_rcp"), [r], [])
            return ast.Call(ast.Name("ng_div"), [l, r], [])
        if isinstance(node.op, ast.Pow) and lv is not None and lv > 0.0 and rv is None:
            return ast.Call(ast.Name("ng_expc"), [ast.BinOp(r, ast.Mult(), _c(math.log(lv)))], [])
        if isinstance(node.op, ast.Pow) and lv is None and rv is not None and rv not in (0.0, 1.0, 2.0, 3.0, 4.0):
            # x^c = exp(c log x) for x > 0 (a concentration ratio, a rate)
            return ast.Call(ast.Name("ng_expc"), [ast.BinOp(_c(rv), ast.Mult(), ast.Call(ast.Name("ng_logc"), [l], []))], [])
        return ast.BinOp(l, node.op, r)
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
        args = [fe(a) for a in node.args]
        vals = [_num(a) for a in args]
        if node.func.id in _FOLD and len(args) == 1 and vals[0] is not None:
            try:
                val = _FOLD[node.func.id](vals[0])
            except (ValueError, OverflowError, ZeroDivisionError):
                val = None
            if _finite(val):
                return _c(val)
        if node.func.id == "exp":
            return share.exp(args[0]) if share is not None and len(args) == 1 else ast.Call(ast.Name("ng_expc"), args, [])
        return ast.Call(node.func, args, [])
    if isinstance(node, ast.Compare):
        return ast.Compare(fe(node.left), node.ops, [fe(c) for c in node.comparators])
    if isinstance(node, ast.BoolOp):
        return ast.BoolOp(node.op, [fe(v) for v in node.values])
    return node


def _neg_unit_over(node: ast.expr) -> Optional[ast.expr]:
    """T when node is -1/T (as (-1)/T or -(1/T)), else None."""
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Div):
        n = node.left
        if _num(n) == -1.0 or (isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub) and _num(n.operand) == 1.0):
            return node.right
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        n = node.operand
        if isinstance(n, ast.BinOp) and isinstance(n.op, ast.Div) and _num(n.left) == 1.0:
            return n.right
    return None


def _reciprocal_of(d: ast.expr) -> Optional[ast.expr]:
    """1/d without a division when d (a fast form) is a quotient, a refined
    reciprocal or a refined reciprocal times a constant, else None."""
    if isinstance(d, ast.Call) and isinstance(d.func, ast.Name):
        if d.func.id == "ng_div" and len(d.args) == 2:
            n, den = d.args
            if _num(n) is not None and _num(n) != 0.0:
                return ast.BinOp(den, ast.Mult(), _c(1.0 / _num(n)))
            return ast.Call(ast.Name("ng_div"), [den, n], [])
        if d.func.id == "ng_rcp" and len(d.args) == 1:
            return d.args[0]
    if isinstance(d, ast.BinOp) and isinstance(d.op, ast.Mult):
        for r, c in ((d.left, d.right), (d.right, d.left)):
            if (_num(c) not in (None, 0.0) and isinstance(r, ast.Call) and isinstance(r.func, ast.Name)
                    and r.func.id == "ng_rcp" and len(r.args) == 1):
                return ast.BinOp(r.args[0], ast.Mult(), _c(1.0 / _num(c)))
    return None


def fast_cnexp(x: str, a: ast.expr, b: ast.expr, consts: Dict[str, float],
               defs: Optional[Dict[str, ast.expr]] = None) -> str:
    """C++ of one cnexp step of x' = A + B x in the fast form; ``defs``: the
    fast forms last assigned to names of the statement list (a time constant
    tau = c / (a + b) gives the rate (a + b) / c without a reciprocal)."""
    T = _neg_unit_over(b)
    if T is not None:
        r = None
        if defs and isinstance(T, ast.Name) and T.id in defs:
            r = _reciprocal_of(defs[T.id])
        rate = cxx(fast_expr(r if r is not None else ast.Call(ast.Name("ng_rcp"), [T], []), consts))
        if isinstance(a, ast.BinOp) and isinstance(a.op, ast.Div) and ast.dump(a.right) == ast.dump(T):
            inf = fast_expr(a.left, consts)                      # A = X / T: -A/B = X
        else:
            inf = fast_expr(ast.BinOp(a, ast.Mult(), T), consts)  # -A/B = A T
        return f"{x} = {x} + (1.0 - ng_expc(-(dt * {rate}))) * ({cxx(inf)} - {x});"
    A, B = fast_expr(a, consts), fast_expr(b, consts)
    q = fast_expr(ast.BinOp(ast.UnaryOp(ast.USub(), a), ast.Div(), b), consts)
    return f"{x} = {x} + (1.0 - ng_expc(dt * ({cxx(B)}))) * ({cxx(q)} - {x});"


# ------------------------------------------------------------------ inlining
class _Subst(ast.NodeTransformer):
    def __init__(self, mapping: Dict[str, ast.expr]):
        self.mapping = mapping

    def visit_Name(self, node):
        return copy.deepcopy(self.mapping[node.id]) if node.id in self.mapping else node


def _subst(text: str, mapping: Dict[str, ast.expr], cond: bool = False) -> str:
    return ast.unparse(_Subst(mapping).visit(_cond(text) if cond else _py(text)))


def _rename(stmts: List[Stmt], mapping: Dict[str, ast.expr]) -> List[Stmt]:
    out = []
    for s in stmts:
        tgt = s.target
        if tgt in mapping:
            if not isinstance(mapping[tgt], ast.Name):
                raise NmodlError(f"cannot inline an assignment to {tgt}")
            tgt = mapping[tgt].id
        if s.kind == "local":
            out.append(Stmt("local", tgt))
        elif s.kind == "assign":
            out.append(Stmt("assign", tgt, _subst(s.expr, mapping)))
        elif s.kind == "call":
            args = [a.strip() for a in s.expr.split(",") if a.strip()]
            out.append(Stmt("call", s.target, ", ".join(_subst(x, mapping) for x in args)))
        elif s.kind == "if":
            out.append(Stmt("if", branches=[(None if c is None else _subst(c, mapping, True), _rename(body, mapping))
                                            for c, body in s.branches]))
        else:
            raise NmodlError(f"cannot inline a {s.kind} statement")
    return out


def _inlinable(fn: Block) -> bool:
    body = statements(fn.body)
    assigned = set(_assigned_names(body))

    def plain(st):
        return all(x.kind in ("assign", "local", "call") or (x.kind == "if" and all(plain(b) for _, b in x.branches))
                   for x in st)
    return plain(body) and not (assigned & set(fn.args))


def _inline_functions(stmts: List[Stmt], mech: Mechanism, counter: List[int]) -> List[Stmt]:
    """The statement list with each call of one of the mechanism's FUNCTIONs
    in an assignment, and each PROCEDURE call statement, replaced by the
    block's statements (its arguments substituted as expressions, its LOCALs
    and a function's value renamed apart), so that exps inside them take
    part in the sharing and the cnexp sees the assignments of the time
    constants.  Blocks that assign their arguments stay calls."""
    funcs = {b.name: b for b in mech.blocks.values() if b.kind == "FUNCTION" and _inlinable(b)}
    procs = {b.name: b for b in mech.blocks.values() if b.kind == "PROCEDURE" and _inlinable(b)}
    if not funcs and not procs:
        return stmts

    def inline_expr(node: ast.expr, pre: List[Stmt]) -> ast.expr:
        class T(ast.NodeTransformer):
            def visit_Call(self, c):
                self.generic_visit(c)
                if not (isinstance(c.func, ast.Name) and c.func.id in funcs):
                    return c
                fn = funcs[c.func.id]
                if len(c.args) != len(fn.args):
                    raise NmodlError(f"{fn.name} takes {len(fn.args)} arguments")
                k = counter[0]
                counter[0] += 1
                body = statements(fn.body)
                mapping = dict(zip(fn.args, c.args))
                val = f"{fn.name}_v{k}"
                mapping[fn.name] = ast.Name(val)
                for st in body:
                    if st.kind == "local":
                        mapping[st.target] = ast.Name(f"{st.target}_l{k}")
                pre.append(Stmt("local", val))
                pre.extend(_inline_functions(_rename(body, mapping), mech, counter))
                return ast.Name(val)
        return T().visit(node)

    out = []
    for s in stmts:
        if s.kind == "assign":
            pre: List[Stmt] = []
            node = inline_expr(_py(s.expr), pre)
            out += pre
            out.append(Stmt("assign", s.target, ast.unparse(node)) if pre else s)
        elif s.kind == "call" and s.target in procs:
            pr = procs[s.target]
            args = [a.strip() for a in s.expr.split(",") if a.strip()]
            if len(args) != len(pr.args):
                raise NmodlError(f"{pr.name} takes {len(pr.args)} arguments")
            pre = []
            argn = [inline_expr(_py(a), pre) for a in args]
            k = counter[0]
            counter[0] += 1
            body = statements(pr.body)
            mapping = dict(zip(pr.args, argn))
            for st in body:
                if st.kind == "local":
                    mapping[st.target] = ast.Name(f"{st.target}_l{k}")
            out += pre
            out += _inline_functions(_rename(body, mapping), mech, counter)
        elif s.kind == "if":
            out.append(Stmt("if", branches=[(c, _inline_functions(body, mech, counter)) for c, body in s.branches]))
        else:
            out.append(s)
    return out


def _exp_args(stmts: List[Stmt]) -> List[ast.expr]:
    """The arguments of the exp calls in a statement list, at any depth."""
    out = []
    for s in stmts:
        nodes = []
        if s.kind == "assign":
            nodes.append(_py(s.expr))
        elif s.kind == "if":
            for c, body in s.branches:
                if c is not None:
                    nodes.append(_cond(c))
                out += _exp_args(body)
        for n in nodes:
            out += [c.args[0] for c in ast.walk(n) if isinstance(c, ast.Call) and isinstance(c.func, ast.Name)
                    and c.func.id == "exp" and len(c.args) == 1]
    return out


# ------------------------------------------------------------------ emission
def _emit(stmts: List[Stmt], mech: Mechanism, ind: str, locals_: List[str], ret: Optional[str] = None,
          fast: Optional[Dict[str, float]] = None, share: Optional[_ExpShare] = None,
          counter: Optional[List[int]] = None) -> List[str]:
    """C++ of a statement list; ``fast``: None for the reference's operation
    order, else the known constants of the fast forms (LOCALs assigned a
    constant at the top level of the list join them; FUNCTION calls inlined;
    an exp inside a branch whose argument the branches do not change, and
    that shares with an exp after the branch, is taken before it);
    ``share``: the shared exps of the enclosing list."""
    out = []
    env = None if fast is None else dict(fast)
    sh = None if fast is None else _ExpShare(share)
    defs: Dict[str, ast.expr] = {}            # fast form last assigned to a name of this list
    if fast is not None and counter is None:
        counter = [0]
        stmts = _inline_functions(stmts, mech, counter)

    def ex(e: ast.expr, keep: Optional[list] = None) -> str:
        if env is None:
            return cxx(e)
        sh.pre = []
        fe = fast_expr(e, env, sh)
        if keep is not None:
            keep.append(fe)
        out.extend(ind + line for line in sh.pre)
        return cxx(fe)

    def forget(name: str) -> None:
        for k in [k for k, d in defs.items() if k == name or name in _names_in(d)]:
            del defs[k]

    for i, s in enumerate(stmts):
        if s.kind == "local":
            if s.target not in locals_:
                locals_.append(s.target)
                out.append(f"{ind}double {s.target} = 0.0;")
        elif s.kind == "assign":
            tgt = "ret_" if ret is not None and s.target == ret else s.target
            e = _py(s.expr)
            keep: list = []
            code = ex(e, keep)
            if env is not None:
                val = _num(fast_expr(e, env))
                env.pop(s.target, None)
                sh.kill(s.target)
                forget(s.target)
                if tgt != "ret_":
                    defs[s.target] = keep[0]
                # a LOCAL assigned a constant, or a derived constant assigned its value
                if val is not None and tgt != "ret_" and (s.target in locals_ or fast.get(s.target) == val):
                    env[s.target] = val
            out.append(f"{ind}{tgt} = {code};")
        elif s.kind == "call":
            args = [a.strip() for a in s.expr.split(",") if a.strip()]
            code = ", ".join([ex(_py(a)) for a in args])
            out.append(f"{ind}{s.target}({code});")
            if env is not None:
                # a procedure may assign any field
                sh.table = {}
                defs.clear()
        elif s.kind == "deriv":
            a, b = linear(_py(s.expr), s.target)
            x = s.target
            if env is not None:
                out.append(ind + fast_cnexp(x, a, b, env, defs))
                sh.kill(x)
                forget(x)
            else:
                A, B = cxx(a), cxx(b)
                out.append(f"{ind}{x} = {x} + (1.0 - ng_exp(dt * ({B}))) * (-({A}) / ({B}) - {x});")
        elif s.kind == "if":
            if env is not None:
                changed = set(_assigned_names([s]))
                later = set()
                for x in _exp_args(stmts[i + 1:]):
                    k = _exp_key(fast_expr(x, env))
                    if k is not None:
                        later |= {k[0], k[1]}
                sh.pre = []
                for x in _exp_args([s]):
                    fx = fast_expr(x, env)
                    k = _exp_key(fx)
                    if k is not None and not (k[2] & changed) and k[0] in later and not sh.has(fx):
                        sh.exp(fx)
                out.extend(ind + line for line in sh.pre)
            for k, (cond, body) in enumerate(s.branches):
                head = "if" if k == 0 else ("} else if" if cond is not None else "} else")
                if cond is not None and env is not None and k > 0:
                    # an else-if condition is evaluated inside the chain: no shared exps there
                    c = cxx(fast_expr(_cond(cond), env))
                else:
                    c = None if cond is None else ex(_cond(cond))
                out.append(f"{ind}{head}" + (f" ({c}) {{" if cond is not None else " {"))
                sub = None if env is None else {n: v for n, v in env.items()}
                out += _emit(body, mech, ind + "    ", locals_, ret, sub, sh, counter)
            out.append(f"{ind}}}")
            if env is not None:
                # a name assigned in a branch is no longer a known constant
                for _, body in s.branches:
                    for t in _assigned_names(body):
                        env.pop(t, None)
                        sh.kill(t)
                        forget(t)
    return out


def _assigned_names(stmts: List[Stmt]) -> List[str]:
    out = []
    for s in stmts:
        if s.kind in ("assign", "deriv"):
            out.append(s.target)
        elif s.kind == "if":
            for _, body in s.branches:
                out += _assigned_names(body)
    return out


def derived_constants(mech: Mechanism, known: Dict[str, float]) -> Dict[str, float]:
    """ASSIGNED variables that every assignment in the mechanism's blocks sets
    to one and the same constant, given the known constants (celsius, the
    model's constant parameters, CONSTANTs and UNITS constants): for example
    Mainen's tadj = q10^((celsius - temp)/10), assigned in INITIAL and in the
    rates procedure and read in BREAKPOINT.  NEURON keeps such a variable
    between blocks; here it is a constant of the mechanism."""
    assigns: Dict[str, list] = {}

    def walk(stmts):
        for st in stmts:
            if st.kind in ("assign", "deriv"):
                assigns.setdefault(st.target, []).append(st.expr if st.kind == "assign" else None)
            elif st.kind == "if":
                for _, body in st.branches:
                    walk(body)

    for b in mech.blocks.values():
        if b.kind in ("INITIAL", "BREAKPOINT", "DERIVATIVE", "PROCEDURE", "FUNCTION"):
            walk(statements(b.body))
    env = dict(known, **{k: float(c) for k, c in mech.constants.items()})
    out: Dict[str, float] = {}
    changed = True
    while changed:
        changed = False
        for name, exprs in assigns.items():
            if name in out or name not in mech.assigned or name in mech.states or any(e is None for e in exprs):
                continue
            vals = {_num(fast_expr(_py(e), env)) for e in exprs}
            if len(vals) == 1 and None not in vals:
                out[name] = env[name] = vals.pop()
                changed = True
    return out


def emit(mech: Mechanism, source: str = "", fast: Optional[Dict[str, float]] = None,
         known: Optional[Dict[str, float]] = None) -> str:
    """C++ for one mechanism: a struct of its variables with its blocks as
    member functions (initial(), states() -- the solved DERIVATIVE block --,
    breakpoint(), and every PROCEDURE and FUNCTION).  ``fast``: None for the
    reference's operation order, else the constants of the fast forms (they
    need mechs.h's ng_expc, ng_rcp and ng_div).  ``known``: constants known
    at generation (celsius, the model's constant parameters): the fields of
    derived constants (derived_constants) start at their value, and the fast
    forms fold all of them."""
    known = dict(known or {})
    derived = derived_constants(mech, dict(known, **(fast or {})))
    if fast is not None:
        fast = dict(fast, **known, **derived, **{k: float(c) for k, c in mech.constants.items()})
    ionvars = []
    for ion, (rd, wr) in mech.ions.items():
        for v in rd + wr:
            if v not in ionvars:
                ionvars.append(v)
    fields = ["v", "dt"]
    lines = [f"/* {mech.suffix}: generated by neurogpu_amd.nmodl" + (f" from {source}" if source else "") + " */",
             f"struct Mech_{mech.suffix} {{", "    double v = 0.0, dt = 0.0;"]
    for name, val in mech.params.items():
        if name in ("v", "dt") or name in fields:
            continue
        fields.append(name)
        lines.append(f"    double {name} = {repr(float(val)) if val is not None else '0.0'};")
    for name, val in mech.constants.items():
        lines.append(f"    static constexpr double {name} = {repr(float(val))};")
    for name in mech.states + mech.assigned + ionvars:
        if name in fields or name in mech.constants:
            continue
        fields.append(name)
        lines.append(f"    double {name} = {repr(derived[name]) if name in derived else '0.0'};")
    for key, b in mech.blocks.items():
        if b.kind == "FUNCTION":
            lines.append(f"    NG_HD_M double {b.name}({', '.join('double ' + a for a in b.args)})")
            lines.append("    {")
            lines.append("        double ret_ = 0.0;")
            lines += _emit(statements(b.body), mech, "        ", list(b.args), ret=b.name, fast=fast)
            lines.append("        return ret_;")
            lines.append("    }")
    for key, b in mech.blocks.items():
        if b.kind == "PROCEDURE":
            lines.append(f"    NG_HD_M void {b.name}({', '.join('double ' + a for a in b.args)})")
            lines.append("    {")
            lines += _emit(statements(b.body), mech, "        ", list(b.args), fast=fast)
            lines.append("    }")
    ini = mech.blocks.get("INITIAL")
    lines.append("    NG_HD_M void initial()")
    lines.append("    {")
    if ini is not None:
        lines += _emit(statements(ini.body), mech, "        ", [], fast=fast)
    lines.append("    }")
    lines.append("    NG_HD_M void states()")
    lines.append("    {")
    if mech.solve:
        der = mech.blocks.get(f"DERIVATIVE {mech.solve}")
        if der is None:
            raise NmodlError(f"BREAKPOINT solves {mech.solve}, which is not a DERIVATIVE block")
        lines += _emit(statements(der.body), mech, "        ", [], fast=fast)
    lines.append("    }")
    bp = mech.blocks.get("BREAKPOINT")
    lines.append("    NG_HD_M void breakpoint()")
    lines.append("    {")
    if bp is not None:
        lines += _emit(statements(bp.body), mech, "        ", [], fast=fast)
    lines.append("    }")
    lines.append("};")
    return "\n".join(lines)


def interface(mech: Mechanism) -> dict:
    """Names of a generated mechanism's parameters, states, assigned and ion
    variables, in the order the harness entry passes them."""
    ionvars = []
    for ion, (rd, wr) in mech.ions.items():
        for v in rd + wr:
            if v not in ionvars:
                ionvars.append(v)
    params = [p for p in mech.params if p not in ("v", "dt")]
    seen = set(params) | {"v", "dt"}
    states = [s for s in mech.states if s not in seen]
    seen |= set(states)
    assigned = [a for a in mech.assigned if a not in seen and a not in mech.constants]
    seen |= set(assigned)
    ions = [x for x in ionvars if x not in seen]
    return {"suffix": mech.suffix, "params": params, "states": states, "assigned": assigned, "ions": ions,
            "currents": mech.currents}


def harness(mech: Mechanism) -> str:
    """extern "C" int ng_gen_<suffix>(block, v, dt, params, states, assigned, ions):
    loads the arrays into an instance, runs block 0 (initial), 1 (states) or
    2 (breakpoint), and stores every array back."""
    f = interface(mech)
    s = mech.suffix
    out = [f"extern \"C\" int ng_gen_{s}(int block, double v, double dt, double *p, double *st, double *as, double *io)",
           "{", f"    Mech_{s} m;", "    m.v = v; m.dt = dt;"]
    for arr, names in (("p", f["params"]), ("st", f["states"]), ("as", f["assigned"]), ("io", f["ions"])):
        for k, n in enumerate(names):
            out.append(f"    m.{n} = {arr}[{k}];")
    out.append("    if (block == 0) m.initial(); else if (block == 1) m.states(); else if (block == 2) m.breakpoint(); else return -1;")
    for arr, names in (("p", f["params"]), ("st", f["states"]), ("as", f["assigned"]), ("io", f["ions"])):
        for k, n in enumerate(names):
            out.append(f"    {arr}[{k}] = m.{n};")
    out.append("    return 0;")
    out.append("}")
    return "\n".join(out)


# NEURON's built-in passive mechanism, as .mod text (it has no .mod file)
PAS_MOD = """NEURON {
    SUFFIX pas
    NONSPECIFIC_CURRENT i
    RANGE g, e
}
PARAMETER {
    g = .001 (S/cm2)
    e = -70 (mV)
}
ASSIGNED {
    v (mV)
    i (mA/cm2)
}
BREAKPOINT {
    i = g*(v - e)
}
"""


def set_parameters(mech: Mechanism, exclude=()) -> List[str]:
    """The parameter-matrix columns a mechanism owns: its PARAMETERs in
    declaration order, without v, dt, celsius, its ion variables and the
    names in ``exclude`` (the model's constant parameters) -- the layout of
    the reference's generated parameter list, e.g. Ih's gIhbar, ehcn, or
    Mainen's na: gbar, tha, qa, Ra, Rb, thi1, thi2, qi, thinf, qinf, Rg, Rd
    with vshift, temp, q10, vmin and vmax constant."""
    ionvars = {x for rd, wr in mech.ions.values() for x in rd + wr}
    return [p for p in mech.params if p not in ("v", "dt", "celsius") and p not in ionvars and p not in exclude]


def model_constants(mech: Mechanism, exclude=(), overrides: Optional[Dict[str, float]] = None) -> Dict[str, float]:
    """Values of a mechanism's constant parameters (``exclude``): the
    PARAMETER default, or overrides["<name>_<suffix>"] (NEURON's name of a
    GLOBAL, e.g. vshift_na)."""
    overrides = overrides or {}
    out = {}
    for p in mech.params:
        if p in exclude:
            key = f"{p}_{mech.suffix}"
            val = overrides.get(key, mech.params[p])
            if val is None:
                raise NmodlError(f"constant parameter {key} has no value")
            out[p] = float(val)
    return out


def emit_set(name: str, mechs: List[Mechanism], celsius: float, ion_values: Dict[str, float],
             masks: List[int], nernst_R: float = 8.3134, exclude=(),
             overrides: Optional[Dict[str, float]] = None, fast: bool = False) -> str:
    """A mechanism set for the simulation kernel (the interface of mechs.h's
    SetTTPC) over generated mechanisms, in the given order (bit k = mechanism
    k): state slots in mechanism order, then the calcium concentration, its
    reversal and the calcium current handed from BREAKPOINT to the next
    DERIVATIVE pass; parameters in mechanism order (set_parameters); each
    block call fills an instance of the mechanism's struct from the slots,
    the parameters and the model's ion values, runs the block and stores
    the states back.  ``ion_values``: constant ion variables (ek, ena, cao,
    cai0 = the initial inside calcium); eca follows cai by Nernst
    (1000 R (celsius + 273.15) / (2 F) ln(cao / cai)) where a mechanism
    writes cai.  ``masks``: mechanism masks that get straight-line code.
    ``exclude``, ``overrides``: the model's constant parameters and their
    values (model_constants).  ``fast``: eca by the branch-free log and the
    refined reciprocal, as the fast forms of the mechanisms."""
    slots, nslot = [], 0
    for m in mechs:
        own = [st for st in interface(m)["states"] if not (st.endswith("i") and st[:-1] in m.ions)]
        slots.append({st: nslot + k for k, st in enumerate(own)})
        nslot += len(own)
    ca = any("ca" in m.ions for m in mechs)
    cai_writer = any("cai" in m.ions.get("ca", ([], []))[1] for m in mechs)
    CAI, ECA, ICA = (nslot, nslot + 1, nslot + 2) if ca else (31, 31, 0)
    nstate = nslot + 3 if ca else max(nslot, 1)
    params, pbase = [], []
    for m in mechs:
        pbase.append(len(params))
        params += [(m.suffix, p) for p in set_parameters(m, exclude)]
    pbase.append(len(params))
    ca_users = sum(1 << k for k, m in enumerate(mechs) if "ca" in m.ions)
    ica_readers = sum(1 << k for k, m in enumerate(mechs) if "ica" in m.ions.get("ca", ([], []))[0])
    cao = ion_values.get("cao", 2.0)
    cai0 = ion_values.get("cai0", 5e-5)
    ktf = 1000.0 * nernst_R * (celsius + 273.15) / UNIT_CONSTANTS["faraday"]
    S = f"SetGen_{name}"

    def fill(k: int, m: Mechanism, ica_expr: str, ind: str) -> List[str]:
        out = [f"{ind}Mech_{m.suffix} M;", f"{ind}M.v = v;"]
        if "celsius" in m.params or "celsius" in m.assigned:
            out.append(f"{ind}M.celsius = {celsius!r};")
        for p, val in model_constants(m, exclude, overrides).items():
            out.append(f"{ind}M.{p} = {val!r};")
        for j, p in enumerate(set_parameters(m, exclude)):
            out.append(f"{ind}M.{p} = P({pbase[k] + j});")
        for st, sl in slots[k].items():
            out.append(f"{ind}M.{st} = st[{sl}];")
        for ion, (rd, wr) in m.ions.items():
            for x in dict.fromkeys(rd + wr):
                if x == "e" + ion:
                    out.append(f"{ind}M.{x} = " + ("st[ECA];" if ion == "ca" and cai_writer else f"{ion_values[x]!r};"))
                elif x == ion + "i":
                    out.append(f"{ind}M.{x} = st[CAI];" if ion == "ca" else f"{ind}M.{x} = {ion_values[x]!r};")
                elif x == ion + "o":
                    out.append(f"{ind}M.{x} = {ion_values[x]!r};")
                elif x == "i" + ion and x in rd:
                    out.append(f"{ind}M.{x} = {ica_expr};")
        return out

    def store(k: int, m: Mechanism, ind: str) -> List[str]:
        out = [f"{ind}st[{sl}] = M.{st};" for st, sl in slots[k].items()]
        if "cai" in m.ions.get("ca", ([], []))[1]:
            out.append(f"{ind}st[CAI] = M.cai;")
        return out

    L = [f"/* {S}: the simulation kernel's mechanism set over the generated mechanisms",
         f" * {', '.join(m.suffix for m in mechs)} (bit k = mechanism k); generated by neurogpu_amd.nmodl. */",
         f"struct {S} {{",
         f"    static constexpr int NMECH = {len(mechs)}, NPARAM = {len(params)}, NSTATE = {nstate};",
         f"    static constexpr int CAI = {CAI}, ECA = {ECA}, ICA = {ICA};",
         f"    /* first parameter-matrix column of each mechanism, then the total */",
         f"    static constexpr int PBASE[{len(pbase)}] = {{{', '.join(map(str, pbase))}}};",
         "    template <class F> NG_DEV static void dispatch(unsigned m, F &&f)",
         "    {",
         f"        ng_dispatch(m, f, NgMasks<{', '.join(hex(x) + 'u' for x in masks)}>{{}});",
         "    }",
         f"    NG_DEV static unsigned ica_bits(unsigned m) {{ return (m & {hex(ica_readers)}u) ? 1u << ICA : 0u; }}",
         (f"    NG_DEV static double eca_of(double cai) {{ return {ktf / 2!r} * ng_logc(ng_div({cao!r}, cai)); }}"
          if fast else
          f"    NG_DEV static double eca_of(double cai) {{ return {ktf / 2!r} * ng_log({cao!r} / cai); }}"),
         "    NG_DEV static void eca_after_cnexp(unsigned m, double *st)",
         "    {",
         (f"        if (m & {hex(ca_users)}u) st[ECA] = eca_of(st[CAI]);" if cai_writer else "        (void)m; (void)st;"),
         "    }",
         "    NG_DEV static unsigned state_bits(unsigned m)",
         "    {",
         "        unsigned s = 0;"]
    for k, m in enumerate(mechs):
        if slots[k]:
            bits = sum(1 << sl for sl in slots[k].values())
            L.append(f"        if (m & (1u << {k})) s |= {hex(bits)}u;")
    if ca:
        L.append(f"        if (m & {hex(ca_users)}u) s |= (1u << CAI) | (1u << ECA);")
    L += ["        return s;", "    }",
          "    /* INITIAL blocks in mechanism order.  A calcium mechanism that writes cai",
          "     * or does not read it first sets it to its resting value; every calcium",
          "     * mechanism sets eca from cai (CuInitModel_*, Code/OneKernel/src/AllModels.cu:202-296,",
          "     * Figures/FigureS4_mainen/CParsed/AllModels.cu) */",
          "    template <class PF> NG_DEV static void init(unsigned m, double v, double *st, const PF &P)",
          "    {"]
    for k, m in enumerate(mechs):
        L.append(f"        if (m & (1u << {k})) {{")
        if "ca" in m.ions:
            if "cai" not in m.ions["ca"][0] or "cai" in m.ions["ca"][1]:
                L.append(f"            st[CAI] = {cai0!r};")
            if cai_writer:
                L.append("            st[ECA] = eca_of(st[CAI]);")
        L += fill(k, m, "0.0", "            ")
        L.append("            M.initial();")
        L += store(k, m, "            ")
        L.append("        }")
    L += ["    }",
          "    /* DERIVATIVE (cnexp) blocks at the new v, in mechanism order; ica is the",
          "     * calcium current of the last BREAKPOINT */",
          "    template <class PF> NG_DEV static void cnexp(unsigned m, double v, double dt, double *st, double ica, const PF &P)",
          "    {",
          "        (void)ica;"]
    for k, m in enumerate(mechs):
        if m.solve is None:
            continue
        L.append(f"        if (m & (1u << {k})) {{")
        L += fill(k, m, "ica", "            ")
        L.append("            M.dt = dt;")
        L.append("            M.states();")
        L += store(k, m, "            ")
        L.append("        }")
    L += ["    }",
          "    /* BREAKPOINT blocks: the sum of the currents the mechanisms write, and",
          "     * the calcium currents into ica when asked */",
          "    template <class PF>",
          "    NG_DEV static void current(unsigned m, double v, const double *stc, const PF &P, double &sum, double &ica, bool add_ica)",
          "    {",
          "        double *st = const_cast<double *>(stc);   /* read only: no block below stores a slot */",
          "        double s = sum;",
          "        (void)ica; (void)add_ica;"]
    for k, m in enumerate(mechs):
        cur = m.currents
        if not cur:
            continue
        L.append(f"        if (m & (1u << {k})) {{")
        L += fill(k, m, "0.0", "            ")
        L.append("            M.breakpoint();")
        for c in cur:
            L.append(f"            s += M.{c};")
            if c == "ica":
                L.append("            if (add_ica) ica += M.ica;")
        L.append("        }")
    L += ["        sum = s;", "    }", "};"]
    return "\n".join(L)


PROLOGUE = """/* Generated by neurogpu_amd.nmodl -- do not edit. */
#pragma once
#include "ngmath.h"
#include <math.h>
#ifdef __HIPCC__
#define NG_HD_M __host__ __device__ inline
#else
#define NG_HD_M inline
#endif
"""


def generate(sources: Dict[str, str], with_harness: bool = False,
             fast: Optional[Dict[str, float]] = None,
             known: Optional[Dict[str, Dict[str, float]]] = None,
             namespace: str = "", extra: str = "", export: str = "") -> Tuple[str, List[dict]]:
    """C++ for several .mod texts (name -> text) and their interfaces
    (``fast`` as in :func:`emit`; the header then includes mechs.h;
    ``known``: per mechanism suffix, constants known at generation, with the
    entry "" applying to every mechanism).  With ``namespace`` the structs,
    and ``extra`` after them, are in that namespace (so that several
    generated sets can share a translation unit), and ``export`` is made
    visible outside it."""
    known = known or {}
    parts, faces = [PROLOGUE if fast is None else PROLOGUE.replace('"ngmath.h"', '"mechs.h"')], []
    if namespace:
        parts.append(f"namespace {namespace} {{")
    for name, text in sources.items():
        mech = parse(text)
        parts.append(emit(mech, name, fast, dict(known.get("", {}), **known.get(mech.suffix, {}))))
        faces.append(interface(mech))
        if with_harness:
            parts.append(harness(mech))
    if extra:
        parts.append(extra)
    if namespace:
        parts.append(f"}} // namespace {namespace}")
        if export:
            parts.append(f"using {namespace}::{export};")
    return "\n\n".join(parts) + "\n", faces


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mods", nargs="+")
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("--harness", action="store_true", help="also emit the extern \"C\" evaluation entries")
    ap.add_argument("--interface", default="", help="write the mechanisms' variable lists as JSON here")
    ap.add_argument("--set", default="", help="also emit SetGen_<NAME>, a mechanism set for the simulation kernel")
    ap.add_argument("--pas", action="store_true", help="add NEURON's built-in pas mechanism")
    ap.add_argument("--celsius", type=float, default=6.3)
    ap.add_argument("--ion", action="append", default=[], help="constant ion value, e.g. ek=-85 (repeatable)")
    ap.add_argument("--masks", default="", help="comma-separated mechanism masks with straight-line code")
    ap.add_argument("--fast", action="store_true",
                    help="emit the fast forms (constants folded, refined reciprocals, branch-free exp, "
                         "cnexp of (X - x) / T without the cancelling quotients) instead of the reference's "
                         "operation order")
    ap.add_argument("--constant", default="",
                    help="comma-separated PARAMETER names that are constants of the model, not parameter-matrix "
                         "columns (e.g. temp,q10,vmin,vmax,vshift)")
    ap.add_argument("--global", dest="globals_", action="append", default=[],
                    help="value of a constant parameter, by NEURON's GLOBAL name, e.g. vshift_na=-5 (repeatable)")
    args = ap.parse_args(argv)
    import os
    srcs = {os.path.basename(p): open(p).read() for p in args.mods}
    if args.pas:
        srcs["pas (built in)"] = PAS_MOD
    if args.set:
        # mechanism order: suffixes sorted case-insensitively (the reference's
        # model order, PYS/extractModel.py:327)
        order = sorted(srcs, key=lambda k: parse(srcs[k]).suffix.lower())
        srcs = {k: srcs[k] for k in order}
    exclude = tuple(x.strip() for x in args.constant.split(",") if x.strip())
    overrides = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in args.globals_}
    known = {"": {"celsius": args.celsius}}
    for text in srcs.values():
        m = parse(text)
        known[m.suffix] = model_constants(m, exclude, overrides)
    extra, ns = "", ""
    if args.set:
        mechs = [parse(t) for t in srcs.values()]
        ions = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in args.ion}
        masks = [int(x, 0) for x in args.masks.split(",") if x.strip()]
        extra = emit_set(args.set, mechs, args.celsius, ions, masks, exclude=exclude, overrides=overrides,
                         fast=args.fast)
        ns = f"ng_gen_{args.set}"
    code, faces = generate(srcs, args.harness, {"celsius": args.celsius} if args.fast else None, known,
                           namespace=ns, extra=extra, export=f"SetGen_{args.set}" if args.set else "")
    if args.set:
        code = code.replace('#include "ngmath.h"', '#include "mechs.h"', 1)
    with open(args.output, "w") as fh:
        fh.write(code)
    if args.interface:
        with open(args.interface, "w") as fh:
            json.dump(faces, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
