"""ORACLE — test infrastructure only. Nothing in the product imports this
module; tests/ use it as the checker and tests/golden/ to write fixtures.

An independent restatement of the variation language's semantics: a
tokenizer and recursive-descent parser written from the grammar
(src/librefrakt/depend/flang/src/grammar_defs.h:1-112) and an interpreter
that evaluates a variation program in float64 over numpy arrays of points
(one lane per point), with the meaning the reference's generated C++ gives
each construct (src/librefrakt/src/flame_compiler.cpp:108-391,
assets/templates/flame.tpl:46-87):

  * `p` is the point after the vlink affine, `result` the vlink's output
    accumulator (starting at 0), `weight` the variation weight, `param.X` the
    variation's parameters, `aff.X` the vlink affine (x' = a x + b y + c,
    y' = d x + e y + f), `common.X` the shared per-point values of common.yml,
    `math.pi` / `math.inv_pi`;
  * integer literals are C ints and decimals are the float type, with C's
    usual arithmetic conversions (int op int stays int, `/` truncates);
  * every library call is `fl::f<FloatT>(...)`: `floor` returns the float
    type, `trunc` and `rint` return int (flamelib.h:392-417), `modf` is C's
    fmod, `eps(x)` = x == 0 ? 1e-20 : x (flamelib.h:481-484), `sincos` is
    (sin, cos), `cossin` is (cos, sin), and the `pi` forms take multiples of
    pi (flamelib.h:419-450);
  * `break` leaves the variation's code (it is emitted inside
    `do { ... } while (0)`);
  * `rand01()` = u / 4294967295 in float (the divisor rounds to 2^32),
    `randbit()` = bit 24 of a word and `randgauss()` draws the angle first,
    then four uniforms (random.h:8-28); the words come from a caller-supplied
    per-lane stream, consumed only by lanes that execute the call.

Control flow runs as masked lanes: an `if` evaluates both branches with the
lanes where its condition holds (or fails), assignments only change active
lanes, and `break` retires lanes for the rest of the program. The FMA fusion
of the reference's emitter changes only roundings, which float64 makes
negligible next to the float32 results this checks.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

# ------------------------------------------------------------------ lexer

_TOKEN = re.compile(r"""
    (?P<ws>\s+|\#[^\n]*)
  | (?P<num>\d+\.\d+|\d+)
  | (?P<id>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>\*\*|==|!=|<=|>=|&&|\|\||\+=|-=|\*=|/=|:=|[-+*/<>=!(){},;.?:])
""", re.VERBOSE)


@dataclass
class Tok:
    kind: str   # num, id, op, eof
    text: str
    pos: int


def tokenize(src: str) -> list[Tok]:
    out, i = [], 0
    while i < len(src):
        m = _TOKEN.match(src, i)
        if not m:
            raise SyntaxError(f"unexpected character {src[i]!r} at {i}")
        i = m.end()
        if m.lastgroup == "ws":
            continue
        out.append(Tok(m.lastgroup, m.group(), m.start()))
    out.append(Tok("eof", "", len(src)))
    return out


# ------------------------------------------------------------------ parser
# AST nodes are tuples: (kind, ...)

class Parser:
    def __init__(self, src: str):
        self.t = tokenize(src)
        self.i = 0

    def peek(self, k: int = 0) -> Tok:
        return self.t[min(self.i + k, len(self.t) - 1)]

    def eat(self, text: str | None = None, kind: str | None = None) -> Tok:
        tok = self.peek()
        if (text is not None and tok.text != text) or (kind is not None and tok.kind != kind):
            raise SyntaxError(f"expected {text or kind} at {tok.pos}, found {tok.text!r}")
        self.i += 1
        return tok

    def accept(self, text: str) -> bool:
        if self.peek().text == text and self.peek().kind != "eof":
            self.i += 1
            return True
        return False

    # program := statement+
    def program(self):
        stmts = []
        while self.peek().kind != "eof":
            stmts.append(self.statement())
        if not stmts:
            raise SyntaxError("empty program")
        return ("block", stmts)

    def statement(self):
        tok = self.peek()
        if tok.kind == "id" and tok.text == "if":
            self.i += 1
            cond = self.expression()
            then = self.statement()
            other = None
            if self.peek().kind == "id" and self.peek().text == "else":
                self.i += 1
                other = self.statement()
            return ("if", cond, then, other)
        if tok.kind == "id" and tok.text == "break":
            self.i += 1
            self.eat(";")
            return ("break",)
        if tok.text == "{":
            self.i += 1
            stmts = []
            while not self.accept("}"):
                stmts.append(self.statement())
            if not stmts:
                raise SyntaxError("empty scope")
            return ("block", stmts)
        target = self.lvalue()
        op = self.eat(kind="op").text
        if op == ":=":
            if target[0] != "var":
                raise SyntaxError("declaration of a member")
            e = self.expression()
            self.eat(";")
            return ("decl", target[1], e)
        if op not in ("=", "+=", "-=", "*=", "/="):
            raise SyntaxError(f"bad assignment operator {op}")
        e = self.expression()
        self.eat(";")
        return ("assign", target, op, e)

    def lvalue(self):
        name = self.eat(kind="id").text
        if self.peek().text == "." and self.peek(1).kind == "id":
            self.i += 1
            member = self.eat(kind="id").text
            return ("member", name, member)
        return ("var", name)

    # expression := '(' boolean ')' '?' boolean ':' boolean | boolean
    def expression(self):
        if self.peek().text == "(":
            save = self.i
            try:
                self.i += 1
                c = self.boolean()
                self.eat(")")
                if self.accept("?"):
                    a = self.boolean()
                    self.eat(":")
                    b = self.boolean()
                    return ("ternary", c, a, b)
            except SyntaxError:
                pass
            self.i = save
        return self.boolean()

    def _left(self, sub, ops):
        e = sub()
        while self.peek().kind in ("op", "id") and self.peek().text in ops:
            op = self.peek().text
            self.i += 1
            e = ("bin", {"and": "&&", "or": "||"}.get(op, op), e, sub())
        return e

    def boolean(self):
        return self._left(self.equality, ("&&", "||", "and", "or"))

    def equality(self):
        return self._left(self.comparison, ("==", "!="))

    def comparison(self):
        return self._left(self.sum, ("<", "<=", ">", ">="))

    def sum(self):
        return self._left(self.product, ("+", "-"))

    def product(self):
        return self._left(self.power, ("*", "/"))

    def power(self):
        return self._left(self.unary, ("**",))

    def unary(self):
        if self.peek().text in ("-", "!"):
            op = self.eat().text
            return ("neg" if op == "-" else "not", self.value())
        return self.value()

    def value(self):
        tok = self.peek()
        if tok.kind == "num":
            self.i += 1
            return ("dec", float(tok.text)) if "." in tok.text else ("int", int(tok.text))
        if tok.kind == "id" and tok.text in ("true", "false"):
            self.i += 1
            return ("bool", tok.text == "true")
        if tok.text == "(":
            self.i += 1
            e = self.expression()
            self.eat(")")
            return e
        lv = self.lvalue()
        if self.peek().text == "(":
            self.i += 1
            args = []
            if not self.accept(")"):
                args.append(self.expression())
                while self.accept(","):
                    args.append(self.expression())
                self.eat(")")
            fname = lv[1] if lv[0] == "var" else f"{lv[1]}.{lv[2]}"
            return ("call", fname, args)
        return lv


def parse(src: str):
    return Parser(src).program()


# ------------------------------------------------------------------ values

@dataclass
class Val:
    kind: str          # 'f' float, 'i' int, 'b' bool, 'v' vec2
    x: np.ndarray      # the float type (ints and bools held exactly as floats/bools)
    y: np.ndarray | None = None


# the float type of the evaluation in progress: float64 normally, float32 to
# measure how a program behaves in single precision (evaluate(dtype=...))
_FT = [np.float64]


def _f(a) -> Val:
    return Val("f", np.asarray(a, dtype=_FT[0]))


def _v(x, y) -> Val:
    return Val("v", np.asarray(x, dtype=_FT[0]), np.asarray(y, dtype=_FT[0]))


class Random:
    """per-lane word streams: words[lane, k] is lane's k-th 32-bit word.
    rand01 is u / 4294967295 in the float type of the kernel being checked:
    in float the divisor rounds to 2^32 (the default), in double it does not
    (double_divide=True, random.h:17-18)."""

    def __init__(self, words: np.ndarray, double_divide: bool = False):
        self.words = np.asarray(words, dtype=np.uint64)
        self.cur = np.zeros(self.words.shape[0], dtype=np.int64)
        self.double_divide = double_divide

    def next(self, mask: np.ndarray) -> np.ndarray:
        n = self.words.shape[0]
        if np.any(self.cur[mask] >= self.words.shape[1]):
            raise IndexError("random word stream exhausted")
        idx = np.minimum(self.cur, self.words.shape[1] - 1)
        w = self.words[np.arange(n), idx]
        self.cur = self.cur + mask.astype(np.int64)
        return w

    def rand01(self, mask):
        u = self.next(mask)
        if self.double_divide:
            return (u.astype(np.float64) / 4294967295.0).astype(_FT[0])
        # float(u) rounded to float32, times 2^-32 (exact): random.h:16-22
        return (u.astype(np.float32).astype(np.float64) * 2.0 ** -32).astype(_FT[0])

    def randbit(self, mask):
        return ((self.next(mask) >> 24) & 1).astype(_FT[0])

    def randgauss(self, mask):
        a = self.rand01(mask) * 2.0
        s = sum(self.rand01(mask) for _ in range(4)) - 2.0
        return s * np.cos(np.pi * a), s * np.sin(np.pi * a)


class _Break(Exception):
    pass


def _eps(x):
    return np.where(x == 0.0, 1e-20, x)


_UNARY = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh,
    "asin": np.arcsin, "acos": np.arccos, "atan": np.arctan, "exp": np.exp, "log": np.log, "log2": np.log2,
    "log10": np.log10, "sqrt": np.sqrt, "abs": np.abs, "rsqrt": lambda a: 1.0 / np.sqrt(a), "eps": _eps,
    "sinpi": lambda a: np.sin(np.pi * a), "cospi": lambda a: np.cos(np.pi * a),
    "tanpi": lambda a: np.tan(np.pi * a), "exp2": np.exp2, "exp10": lambda a: 10.0 ** a,
    "floor": np.floor,
}
_BINARY = {"copysign": np.copysign, "modf": np.fmod, "pow": np.power, "atan2": np.arctan2, "hypot": np.hypot}
_DUAL = {
    "sincos": lambda a: (np.sin(a), np.cos(a)), "cossin": lambda a: (np.cos(a), np.sin(a)),
    "sincospi": lambda a: (np.sin(np.pi * a), np.cos(np.pi * a)),
    "cossinpi": lambda a: (np.cos(np.pi * a), np.sin(np.pi * a)),
}


@dataclass
class Context:
    n: int
    p: Val | None = None
    weight: float = 0.0
    params: dict = field(default_factory=dict)
    aff: dict = field(default_factory=dict)
    commons: dict = field(default_factory=dict)   # name -> AST
    rng: Random | None = None
    precalc: bool = False
    mca: "Perturb | None" = None


@dataclass
class Perturb:
    """Monte Carlo arithmetic (D. S. Parker, "Monte Carlo Arithmetic",
    UCLA CSD-970002, 1997): every float result of an operator is multiplied
    by (1 + op_rel * u) and every library function's result by
    (1 + fn_rel * u) plus fn_abs * u for the periodic functions, with u
    uniform in [-1, 1] per lane and operation. A float64 evaluation so
    perturbed behaves like one in a lower precision with less exact
    functions; the spread of several such runs bounds how far a float32
    kernel with hardware transcendentals may land from the exact result."""
    rng: np.random.Generator
    op_rel: float = 2.0 ** -24
    fn_rel: float = 2.0 ** -21
    fn_abs: float = 2.0 ** -21

    def op(self, x):
        x = np.asarray(x)
        return x * (1.0 + self.op_rel * self.rng.uniform(-1.0, 1.0, x.shape))

    def fn(self, x, periodic: bool):
        x = np.asarray(x)
        y = x * (1.0 + self.fn_rel * self.rng.uniform(-1.0, 1.0, x.shape))
        if periodic:
            y = y + self.fn_abs * self.rng.uniform(-1.0, 1.0, x.shape)
        return y


_PERIODIC = {"sin", "cos", "tan", "sinpi", "cospi", "tanpi", "sincos", "cossin", "sincospi", "cossinpi"}


class Interp:
    def __init__(self, ctx: Context):
        self.c = ctx
        self.scopes: list[dict] = [{}]
        self.common_cache: dict = {}
        n = ctx.n
        self.alive = np.ones(n, dtype=bool)
        self.result = _v(np.zeros(n), np.zeros(n))
        self.ft = _FT[0]

    # ---- statements
    def run(self, ast, mask=None):
        mask = np.ones(self.c.n, dtype=bool) if mask is None else mask
        try:
            self.stmt(ast, mask)
        except _Break:
            pass

    def stmt(self, s, mask):
        kind = s[0]
        m = mask & self.alive
        if kind == "block":
            self.scopes.append({})
            for t in s[1]:
                self.stmt(t, mask)
            self.scopes.pop()
        elif kind == "decl":
            self.scopes[-1][s[1]] = self.expr(s[2], m)
        elif kind == "assign":
            self.assign(s[1], s[2], s[3], m)
        elif kind == "if":
            cond = self.expr(s[1], m)
            cb = np.broadcast_to(np.asarray(cond.x, dtype=bool), (self.c.n,))
            self.scopes.append({})
            self.stmt(s[2], m & cb)
            self.scopes.pop()
            if s[3] is not None:
                self.scopes.append({})
                self.stmt(s[3], m & ~cb)
                self.scopes.pop()
        elif kind == "break":
            self.alive = self.alive & ~m
        else:
            raise ValueError(kind)

    def _lookup_scope(self, name):
        for sc in reversed(self.scopes):
            if name in sc:
                return sc
        return None

    def assign(self, target, op, e, m):
        rhs = self.expr(e, m)
        if target[0] == "var":
            name = target[1]
            if name == "result":
                old = self.result
            else:
                sc = self._lookup_scope(name)
                if sc is None:
                    raise NameError(name)
                old = sc[name]
            new = self.binop(op[0], old, rhs) if op != "=" else self.convert(rhs, old.kind)
            new = self.merge(old, new, m)
            if name == "result":
                self.result = new
            else:
                sc[name] = new
            return
        grp, mem = target[1], target[2]
        if grp == "param":
            if not self.c.precalc:
                raise NameError("parameters are read-only outside precalc")
            old = _f(self.c.params[mem])
            new = self.binop(op[0], old, rhs) if op != "=" else self.convert(rhs, "f")
            self.c.params[mem] = self.merge(old, new, m).x
            return
        if grp == "result":
            holder = self.result
        else:
            sc = self._lookup_scope(grp)
            if sc is None:
                raise NameError(grp)
            holder = sc[grp]
        if holder.kind != "v" or mem not in ("x", "y"):
            raise TypeError(f"member {mem} of a non-vector")
        old = _f(holder.x if mem == "x" else holder.y)
        new = self.binop(op[0], old, rhs) if op != "=" else self.convert(rhs, "f")
        val = self.merge(old, new, m).x
        nv = _v(val, holder.y) if mem == "x" else _v(holder.x, val)
        if grp == "result":
            self.result = nv
        else:
            sc[grp] = nv

    def merge(self, old: Val, new: Val, m) -> Val:
        if new.kind == "v":
            ox = old.x if old.kind == "v" else np.zeros(self.c.n, self.ft)
            oy = old.y if old.kind == "v" else np.zeros(self.c.n, self.ft)
            return Val("v", np.where(m, new.x, ox), np.where(m, new.y, oy))
        return Val(new.kind, np.where(m, new.x, old.x))

    def convert(self, v: Val, kind: str) -> Val:
        if kind == "i" and v.kind == "f":
            return Val("i", np.trunc(v.x))
        if kind == "f" and v.kind == "i":
            return Val("f", v.x)
        return v

    # ---- expressions
    def expr(self, e, m) -> Val:
        k = e[0]
        if k == "dec":
            return Val("f", self.ft(e[1]))
        if k == "int":
            return Val("i", self.ft(e[1]))
        if k == "bool":
            return Val("b", np.bool_(e[1]))
        if k == "var":
            return self.var(e[1])
        if k == "member":
            return self.member(e[1], e[2], m)
        if k == "neg":
            v = self.expr(e[1], m)
            return _v(-v.x, -v.y) if v.kind == "v" else Val(v.kind, -v.x)
        if k == "not":
            v = self.expr(e[1], m)
            return Val("b", ~np.asarray(v.x, dtype=bool))
        if k == "bin":
            op = e[1]
            a = self.expr(e[2], m)
            if op in ("&&", "||"):
                ab = np.asarray(a.x, dtype=bool)
                sub = m & ab if op == "&&" else m & ~ab
                b = np.asarray(self.expr(e[3], sub).x, dtype=bool)
                return Val("b", (ab & b) if op == "&&" else (ab | b))
            return self.binop(op, a, self.expr(e[3], m))
        if k == "ternary":
            c = np.broadcast_to(np.asarray(self.expr(e[1], m).x, dtype=bool), (self.c.n,))
            a = self.expr(e[2], m & c)
            b = self.expr(e[3], m & ~c)
            if a.kind == "v" or b.kind == "v":
                return _v(np.where(c, a.x, b.x), np.where(c, a.y, b.y))
            kind = "f" if "f" in (a.kind, b.kind) else a.kind
            return Val(kind, np.where(c, a.x, b.x))
        if k == "call":
            return self.call(e[1], e[2], m)
        raise ValueError(k)

    def var(self, name) -> Val:
        if name == "p":
            return self.c.p
        if name == "result":
            return self.result
        if name == "weight":
            return Val("f", self.ft(self.c.weight))
        sc = self._lookup_scope(name)
        if sc is None:
            raise NameError(name)
        return sc[name]

    def member(self, grp, mem, m) -> Val:
        if grp == "param":
            return Val("f", np.asarray(self.c.params[mem], dtype=self.ft))
        if grp == "aff":
            return Val("f", self.ft(self.c.aff[mem]))
        if grp == "math":
            return Val("f", self.ft({"pi": math.pi, "inv_pi": 1.0 / math.pi}[mem]))
        if grp == "common":
            if mem not in self.common_cache:
                self.common_cache[mem] = self.expr(self.c.commons[mem], m)
            return self.common_cache[mem]
        base = self.var(grp)
        if base.kind != "v":
            raise TypeError(f"{grp} is not a vector")
        return _f(base.x if mem == "x" else base.y)

    def binop(self, op, a: Val, b: Val) -> Val:
        if op in ("<", "<=", ">", ">=", "==", "!="):
            f = {"<": np.less, "<=": np.less_equal, ">": np.greater, ">=": np.greater_equal,
                 "==": np.equal, "!=": np.not_equal}[op]
            return Val("b", f(a.x, b.x))
        if a.kind == "v" or b.kind == "v":
            if op in "+-":
                if a.kind != "v" or b.kind != "v":
                    raise TypeError("vector +/- scalar")
                f = np.add if op == "+" else np.subtract
                return _v(f(a.x, b.x), f(a.y, b.y))
            if op == "*":
                if a.kind == "v" and b.kind == "v":
                    raise TypeError("vector * vector")
                v, s = (a, b) if a.kind == "v" else (b, a)
                return _v(v.x * s.x, v.y * s.x)
            if op == "/":
                if a.kind != "v" or b.kind == "v":
                    raise TypeError("only vector / scalar")
                return _v(a.x / b.x, a.y / b.x)
            raise TypeError(op)
        both_int = a.kind in ("i", "b") and b.kind in ("i", "b")
        if op == "+":
            r = a.x + b.x
        elif op == "-":
            r = a.x - b.x
        elif op == "*":
            r = a.x * b.x
        elif op == "/":
            if both_int:
                bb = np.where(b.x == 0, 1, b.x)
                return Val("i", np.trunc(a.x / bb))
            r = a.x / b.x
        elif op == "**":
            r = np.power(a.x, b.x)
        else:
            raise ValueError(op)
        if self.c.mca is not None and not both_int:
            r = self.c.mca.op(r)
        return Val("i" if both_int else "f", np.asarray(r, dtype=self.ft))

    def call(self, name, args, m) -> Val:
        rng = self.c.rng
        if name == "rand01":
            return _f(rng.rand01(m))
        if name == "randbit":
            return Val("i", rng.randbit(m))
        if name == "randgauss":
            gx, gy = rng.randgauss(m)
            return _v(gx, gy)
        vals = [self.expr(a, m) for a in args]
        if name == "vec2":
            return _v(np.broadcast_to(vals[0].x, (self.c.n,)), np.broadcast_to(vals[1].x, (self.c.n,)))
        if name == "elmul":
            a, b = vals
            return _v(a.x * b.x, a.y * b.y)
        if name in ("trunc", "rint"):
            f = np.trunc if name == "trunc" else np.rint
            return Val("i", f(vals[0].x))
        if name == "is_even":
            return Val("b", np.fmod(vals[0].x, 2.0) == 0.0)
        mca = self.c.mca
        per = name in _PERIODIC

        def fz(x):
            return x if mca is None or name in ("abs", "floor", "eps", "copysign") else mca.fn(x, per)

        if name in _DUAL:
            s, c = _DUAL[name](vals[0].x)
            return _v(fz(s), fz(c))
        if name in _BINARY:
            return _f(fz(_BINARY[name](vals[0].x, vals[1].x)))
        if name in _UNARY:
            f = _UNARY[name]
            v = vals[0]
            if v.kind == "v":
                return _v(fz(f(v.x)), fz(f(v.y)))
            return _f(fz(f(v.x)))
        raise NameError(f"unknown function {name}")


# ------------------------------------------------------------------ database

@dataclass
class Variation:
    name: str
    src: str
    params: dict          # name -> {"default": float, "identity": float, "precalc": bool}
    precalc_src: str | None
    ast: tuple = None
    precalc_ast: tuple | None = None

    @property
    def user_params(self) -> list[str]:
        return [k for k, v in self.params.items() if not v["precalc"]]


def load_db(variations_yaml: str, common_yaml: str) -> tuple[dict, dict]:
    """(variations by name, common ASTs by name) from the two YAML texts"""
    import yaml
    vy = yaml.safe_load(variations_yaml)
    cy = yaml.safe_load(common_yaml)
    commons = {k: parse(f"__c := {v};")[1][0][2] for k, v in cy.items()}
    out = {}
    for name, ent in vy.items():
        if isinstance(ent, str):
            src, params, pre = ent, {}, None
        else:
            src = ent["src"]
            pre = ent.get("precalc")
            params = {}
            for pn, pd in (ent.get("param") or {}).items():
                pd = pd or {}
                params[pn] = {"default": float(pd.get("default", 0.0)),
                              "identity": float(pd.get("identity", pd.get("default", 0.0))),
                              "precalc": "precalc" in (pd.get("tags") or [])}
        v = Variation(name, src, params, pre)
        v.ast = parse(src)
        v.precalc_ast = parse(pre) if pre else None
        out[name] = v
    return out, commons


def precalc(v: Variation, user: dict, aff: dict) -> dict:
    """the variation's full parameter set: user values plus precalc slots"""
    prm = {k: float(user.get(k, d["default"])) for k, d in v.params.items() if not d["precalc"]}
    for k, d in v.params.items():
        if d["precalc"]:
            prm[k] = 0.0
    if v.precalc_ast is not None:
        ctx = Context(n=1, params={k: np.float64(x) for k, x in prm.items()}, aff=aff, precalc=True)
        it = Interp(ctx)
        it.run(v.precalc_ast)
        prm = {k: float(np.asarray(x).reshape(-1)[0]) for k, x in ctx.params.items()}
    return prm


def evaluate(v: Variation, commons: dict, px, py, weight: float, params: dict, aff: dict,
             words: np.ndarray | None = None, rng: Random | None = None, with_precalc: bool = True,
             dtype=np.float64, mca: Perturb | None = None):
    """result (x, y) of one variation at points p (after the vlink affine);
    `params` holds the user parameters (precalc runs here unless
    with_precalc is False, in which case it must hold every slot). With
    dtype=np.float32 every value and operation is single precision (numpy's
    float32 functions), which measures the program's own float32 error;
    with `mca` every operation is perturbed (Perturb)."""
    n = np.asarray(px).shape[0]
    prm = precalc(v, params, aff) if with_precalc else dict(params)
    if rng is None:
        rng = Random(words if words is not None else np.zeros((n, 0), np.uint32))
    _FT[0] = dtype
    try:
        px = np.asarray(px, dtype=dtype)
        py = np.asarray(py, dtype=dtype)
        ctx = Context(n=n, p=_v(px, py), weight=float(weight), params={k: dtype(x) for k, x in prm.items()},
                      aff=aff, commons=commons, rng=rng, mca=mca)
        it = Interp(ctx)
        with np.errstate(all="ignore"):
            it.run(v.ast)
        r = it.result
        return (np.broadcast_to(r.x, (n,)).astype(np.float64), np.broadcast_to(r.y, (n,)).astype(np.float64))
    finally:
        _FT[0] = np.float64


def uses_random(v: Variation) -> bool:
    return bool(re.search(r"\brand(01|bit|gauss)\s*\(", v.src))
