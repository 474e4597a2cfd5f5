"""Model compiler: log-density formulas -> stack-machine bytecode.

Restates the reference's DSL compiler (R/dsl.R:13-134) and ``gpum_model``
(R/api.R:87-116) so the HIP backend can be driven without R.  The output
(``code`` as (op, arg) int pairs, a deduplicated f64 constant pool and the
peak stack depth) is the exact bytecode the reference feeds its backends.

Formulas are written as R one-sided formulas, e.g. ``"~ -((y - mu)^2) / 8"``.
R's operator precedence is honoured: ``^`` (right-associative) binds tighter
than unary minus, which binds tighter than ``*`` and ``/``, which bind tighter
than binary ``+`` and ``-`` (left-associative).  Parentheses are transparent.
Supported: ``+ - * /`` (binary), unary ``-``, ``^``, ``exp``, ``log``,
``sqrt``, ``lgamma``; anything else is rejected, as in R/dsl.R:105-111.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

# R/dsl.R:13-17
OP = {
    "PUSH_CONST": 0, "PUSH_PARAM": 1, "PUSH_DATA": 2,
    "ADD": 3, "SUB": 4, "MUL": 5, "DIV": 6, "NEG": 7,
    "EXP": 8, "LOG": 9, "SQRT": 10, "POW": 11, "LGAMMA": 12,
}
# R/dsl.R:20
STACK_SIZE = 32


class DSLError(ValueError):
    """A model expression the DSL does not accept."""


# --------------------------------------------------------------- parsing ---

@dataclass
class Num:
    value: float


@dataclass
class Sym:
    name: str


@dataclass
class Call:
    fn: str
    args: List["Expr"] = field(default_factory=list)


Expr = Union[Num, Sym, Call]

_TOKEN = re.compile(
    r"\s*(?:"
    r"(?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?L?)"
    r"|(?P<name>[A-Za-z.][A-Za-z0-9._]*)"
    r"|(?P<op>\*\*|[-+*/^(),~])"
    r")")


def _tokenize(src: str):
    pos, out = 0, []
    src = src.strip()
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m or m.end() == pos:
            raise DSLError(f"cannot parse the log-density near {src[pos:pos + 10]!r}")
        pos = m.end()
        if m.group("num") is not None:
            txt = m.group("num")
            out.append(("num", float(txt[:-1] if txt.endswith("L") else txt)))
        elif m.group("name") is not None:
            out.append(("name", m.group("name")))
        elif m.group("op") is not None:
            op = m.group("op")
            out.append(("op", "^" if op == "**" else op))
    out.append(("end", None))
    return out


class _Parser:
    """Recursive-descent parser with R's precedence for the DSL subset."""

    def __init__(self, src: str):
        self.toks = _tokenize(src)
        self.i = 0

    def peek(self):
        return self.toks[self.i]

    def take(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def expect(self, op):
        t = self.take()
        if t != ("op", op):
            raise DSLError(f"expected '{op}' in the log-density")

    def parse(self) -> Expr:
        e = self.additive()
        if self.peek()[0] != "end":
            raise DSLError("unexpected trailing input in the log-density")
        return e

    def additive(self) -> Expr:
        e = self.multiplicative()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            e = Call(op, [e, self.multiplicative()])
        return e

    def multiplicative(self) -> Expr:
        e = self.unary()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            e = Call(op, [e, self.unary()])
        return e

    def unary(self) -> Expr:
        t = self.peek()
        if t == ("op", "-"):
            self.take()
            return Call("-", [self.unary()])
        if t == ("op", "+"):
            self.take()
            return Call("+", [self.unary()])
        return self.power()

    def power(self) -> Expr:
        base = self.primary()
        if self.peek() == ("op", "^"):
            self.take()
            # right-associative; the exponent may carry a unary sign
            return Call("^", [base, self.unary_power()])
        return base

    def unary_power(self) -> Expr:
        t = self.peek()
        if t == ("op", "-"):
            self.take()
            return Call("-", [self.unary_power()])
        if t == ("op", "+"):
            self.take()
            return Call("+", [self.unary_power()])
        return self.power()

    def primary(self) -> Expr:
        kind, val = self.take()
        if kind == "num":
            return Num(val)
        if kind == "name":
            if self.peek() == ("op", "("):
                self.take()
                args = []
                if self.peek() != ("op", ")"):
                    args.append(self.additive())
                    while self.peek() == ("op", ","):
                        self.take()
                        args.append(self.additive())
                self.expect(")")
                return Call(val, args)
            return Sym(val)
        if (kind, val) == ("op", "("):
            e = self.additive()
            self.expect(")")
            return Call("(", [e])
        raise DSLError("unsupported term in the log-density expression.")


def parse_rhs(formula: str) -> Expr:
    """Right-hand side of a one-sided formula (R/dsl.R:126-134)."""
    s = formula.strip()
    if s.startswith("~"):
        s = s[1:]
    elif "~" in s:
        raise DSLError("expected a one-sided formula such as `~ expr`.")
    return _Parser(s).parse()


# -------------------------------------------------------------- compiler ---

@dataclass
class Program:
    code: List[int]
    consts: List[float]
    depth: int

    @property
    def n_instr(self) -> int:
        return len(self.code) // 2


def compile_expr(expr: Union[str, Expr], params: Sequence[str],
                 data: Sequence[str] = ()) -> Program:
    """R/dsl.R:28-122: post-order emission, const-pool dedupe by value,
    stack-depth tracking and the depth limit."""
    if isinstance(expr, str):
        expr = parse_rhs(expr)
    params = list(params)
    data = list(data)
    code: List[int] = []
    consts: List[float] = []
    depth = 0
    max_depth = 0

    def push_depth(d):
        nonlocal depth, max_depth
        depth += d
        max_depth = max(max_depth, depth)

    def emit_op(op, arg=0):
        code.extend([OP[op], int(arg)])

    def const_index(v):
        for i, c in enumerate(consts):  # R/dsl.R:41-46: first match wins
            if c == v:
                return i
        consts.append(float(v))
        return len(consts) - 1

    def emit(node: Expr):
        if isinstance(node, Num):
            emit_op("PUSH_CONST", const_index(node.value))
            push_depth(1)
        elif isinstance(node, Sym):
            nm = node.name
            if nm in params:
                emit_op("PUSH_PARAM", params.index(nm))
            elif nm in data:
                emit_op("PUSH_DATA", data.index(nm))
            else:
                raise DSLError(f"unknown symbol in the log-density: '{nm}'. "
                               "Declare it in `params` or `data`.")
            push_depth(1)
        elif isinstance(node, Call):
            fn, args = node.fn, node.args

            def binary(op):
                emit(args[0])
                emit(args[1])
                emit_op(op)
                push_depth(-1)

            def unary_fn(op):
                if len(args) != 1:
                    raise DSLError(f"`{fn}` takes one argument in the log-density.")
                emit(args[0])
                emit_op(op)

            if fn == "(":
                emit(args[0])
            elif fn == "+" and len(args) == 2:
                binary("ADD")
            elif fn == "-" and len(args) == 2:
                binary("SUB")
            elif fn == "-" and len(args) == 1:
                emit(args[0])
                emit_op("NEG")
            elif fn == "*" and len(args) == 2:
                binary("MUL")
            elif fn == "/" and len(args) == 2:
                binary("DIV")
            elif fn == "^" and len(args) == 2:
                binary("POW")
            elif fn == "exp":
                unary_fn("EXP")
            elif fn == "log":
                unary_fn("LOG")
            elif fn == "sqrt":
                unary_fn("SQRT")
            elif fn == "lgamma":
                unary_fn("LGAMMA")
            else:
                raise DSLError(f"unsupported function in the log-density: '{fn}'. "
                               "Supported: + - * / ^ exp log sqrt lgamma.")
        else:
            raise DSLError("unsupported term in the log-density expression.")

    emit(expr)
    if max_depth > STACK_SIZE:
        raise DSLError("the log-density expression is too deep for the interpreter "
                       f"(needs {max_depth} stack slots, limit {STACK_SIZE}).")
    return Program(code=code, consts=consts, depth=max_depth)


@dataclass
class Model:
    """R/api.R:87-116 (gpum_model)."""
    loglik: Program
    prior: Program
    params: List[str]
    data: List[str]

    @property
    def n_params(self) -> int:
        return len(self.params)

    @property
    def n_cols(self) -> int:
        return len(self.data)


def gpum_model(loglik: str, params: Sequence[str], data: Sequence[str] = (),
               prior: Optional[str] = None) -> Model:
    params = [str(p) for p in params]
    data = [str(d) for d in data]
    if len(params) < 1:
        raise DSLError("`params` must name at least one parameter.")
    names = params + data
    if len(set(names)) != len(names):
        raise DSLError("parameter and data names must be distinct.")
    ll = compile_expr(loglik, params, data)
    if prior is None:
        pr = Program(code=[], consts=[], depth=0)
    else:
        pr = compile_expr(prior, params, [])
    return Model(loglik=ll, prior=pr, params=params, data=data)
