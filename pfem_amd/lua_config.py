"""Reader of the reference's problem files: the subset of Lua 5 that its examples use.

The reference reads its problem description with Lua through sol2 (``srcs/simulation/utility/
SolTable.hpp``, ``srcs/simulation/Problem.cpp``): nested tables of parameters, and functions
such as ``Problem.IC:initStates(pos)`` and ``Problem.Solver.MomContEq.BC:<Tag>V(pos, t)`` that
it calls per node.  Neither Lua nor sol2 is part of this package, so this module interprets
the language subset those files are written in:

* statements: ``[local] name = expr``, assignments to ``a.b[c]``, ``function a.b:c(args)``
  and ``function a.b.c(args)`` definitions, ``if / elseif / else / end``, ``return``, calls;
* expressions: numbers, strings, ``true / false / nil``, table constructors (positional,
  ``name = v`` and ``[k] = v`` fields), ``+ - * / % ^``, unary ``-`` and ``not``,
  comparisons, ``and / or`` with Lua's short-circuit values, ``..``, indexing, calls and
  method calls, ``math.*`` (``sin cos tan asin acos atan exp log sqrt abs floor ceil max
  min pi huge``);

with Lua's semantics for these (1-based array parts, ``nil``/``false`` as the only false
values, ``a:f(x)`` passing ``a`` as ``self``).  ``SolTable`` mirrors the reference's class of
that name (``SolTable.hpp:10-117``): ``call``, ``checkCall``, ``checkCallNoThrow``, ``get``,
``checkAndGet``, ``doesVarExist``, ``for_each``, ``getName``.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

__all__ = ["LuaError", "LuaTable", "LuaFunction", "LuaState", "SolTable", "load_problem"]


class LuaError(RuntimeError):
    pass


class LuaTable:
    """A Lua table: string and number keys in one dict (numbers normalised to float/int)."""

    def __init__(self, items=None):
        self.d = {}
        if items:
            for k, v in items.items():
                self[k] = v

    @staticmethod
    def _key(k):
        if isinstance(k, float) and k.is_integer():
            return int(k)
        return k

    def __getitem__(self, k):
        return self.d.get(self._key(k))

    def __setitem__(self, k, v):
        k = self._key(k)
        if v is None:
            self.d.pop(k, None)
        else:
            self.d[k] = v

    def length(self):
        n = 0
        while (n + 1) in self.d:
            n += 1
        return n

    def to_list(self):
        return [to_python(self.d[i]) for i in range(1, self.length() + 1)]

    @staticmethod
    def from_list(seq):
        t = LuaTable()
        for i, v in enumerate(seq):
            t[i + 1] = from_python(v)
        return t


def to_python(v):
    """Lua value -> Python: a table with only an array part becomes a list, any other table a dict."""
    if isinstance(v, LuaTable):
        n = v.length()
        if n == len(v.d):
            return v.to_list()
        return {k: to_python(x) for k, x in v.d.items()}
    return v


def from_python(v):
    if isinstance(v, LuaTable) or v is None or isinstance(v, (bool, str, LuaFunction)) or callable(v):
        return v
    if isinstance(v, dict):
        return LuaTable({k: from_python(x) for k, x in v.items()})
    if isinstance(v, (list, tuple)):
        return LuaTable.from_list(v)
    try:
        import numpy as np

        if isinstance(v, np.ndarray):
            return LuaTable.from_list(v.tolist())
        if isinstance(v, np.generic):
            return v.item()
    except ImportError:  # pragma: no cover
        pass
    return v


def truthy(v):
    return not (v is None or v is False)


# ------------------------------------------------------------------------------------------
# tokens
# ------------------------------------------------------------------------------------------
_TOKEN = re.compile(r"""
    (?P<ws>\s+)
  | (?P<comment>--\[(?P<eq>=*)\[.*?\](?P=eq)\]|--[^\n]*)
  | (?P<num>0[xX][0-9a-fA-F]+|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)
  | (?P<str>"(?:[^"\\\n]|\\.)*"|'(?:[^'\\\n]|\\.)*')
  | (?P<name>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<op>\.\.\.|\.\.|==|~=|<=|>=|[-+*/%^#<>=(){}\[\];:,.])
""", re.S | re.X)

_KEYWORDS = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "if", "in", "local", "nil",
             "not", "or", "repeat", "return", "then", "true", "until", "while"}


@dataclass
class Tok:
    kind: str  # num str name kw op eof
    val: object
    line: int


def tokenize(src: str):
    out, pos, line = [], 0, 1
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise LuaError(f"line {line}: unexpected character {src[pos]!r}")
        text = m.group(0)
        kind = m.lastgroup if m.lastgroup != "eq" else "comment"
        if m.group("comment") is not None:
            kind = "comment"
        if kind == "num":
            out.append(Tok("num", float(int(text, 16)) if text.lower().startswith("0x") else float(text), line))
        elif kind == "str":
            out.append(Tok("str", bytes(text[1:-1], "utf-8").decode("unicode_escape"), line))
        elif kind == "name":
            out.append(Tok("kw" if text in _KEYWORDS else "name", text, line))
        elif kind == "op":
            out.append(Tok("op", text, line))
        line += text.count("\n")
        pos = m.end()
    out.append(Tok("eof", None, line))
    return out


# ------------------------------------------------------------------------------------------
# parser (recursive descent; nodes are tuples)
# ------------------------------------------------------------------------------------------
_BINPRI = {"or": (1, 1), "and": (2, 2), "<": (3, 3), ">": (3, 3), "<=": (3, 3), ">=": (3, 3), "~=": (3, 3),
           "==": (3, 3), "..": (9, 8), "+": (10, 10), "-": (10, 10), "*": (11, 11), "/": (11, 11), "%": (11, 11),
           "^": (14, 13)}
_UNARY_PRI = 12


class Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def check(self, kind, val=None):
        tok = self.peek()
        return tok.kind == kind and (val is None or tok.val == val)

    def accept(self, kind, val=None):
        if self.check(kind, val):
            return self.next()
        return None

    def expect(self, kind, val=None):
        tok = self.next()
        if tok.kind != kind or (val is not None and tok.val != val):
            raise LuaError(f"line {tok.line}: expected {val or kind}, got {tok.val!r}")
        return tok

    def block(self, enders=("end",)):
        stmts = []
        while not (self.peek().kind == "eof" or (self.peek().kind == "kw" and self.peek().val in enders)):
            if self.check("kw", "return"):
                line = self.next().line
                exprs = []
                if not (self.check("kw") and self.peek().val in enders) and not self.check("eof") and \
                        not self.check("op", ";"):
                    exprs = self.exprlist()
                self.accept("op", ";")
                stmts.append(("return", exprs, line))
                break
            s = self.statement()
            if s is not None:
                stmts.append(s)
        return stmts

    def statement(self):
        tok = self.peek()
        if self.accept("op", ";"):
            return None
        if tok.kind == "kw" and tok.val == "local":
            self.next()
            if self.accept("kw", "function"):
                name = self.expect("name").val
                return ("local", [name], [self.funcbody(False)], tok.line)
            names = [self.expect("name").val]
            while self.accept("op", ","):
                names.append(self.expect("name").val)
            exprs = self.exprlist() if self.accept("op", "=") else []
            return ("local", names, exprs, tok.line)
        if tok.kind == "kw" and tok.val == "function":
            self.next()
            target = ("name", self.expect("name").val)
            method = False
            while True:
                if self.accept("op", "."):
                    target = ("index", target, ("const", self.expect("name").val))
                elif self.accept("op", ":"):
                    target = ("index", target, ("const", self.expect("name").val))
                    method = True
                    break
                else:
                    break
            return ("assign", [target], [self.funcbody(method)], tok.line)
        if tok.kind == "kw" and tok.val == "if":
            self.next()
            clauses = []
            cond = self.expr()
            self.expect("kw", "then")
            body = self.block(("elseif", "else", "end"))
            clauses.append((cond, body))
            orelse = []
            while True:
                if self.accept("kw", "elseif"):
                    cond = self.expr()
                    self.expect("kw", "then")
                    clauses.append((cond, self.block(("elseif", "else", "end"))))
                elif self.accept("kw", "else"):
                    orelse = self.block(("end",))
                    self.expect("kw", "end")
                    break
                else:
                    self.expect("kw", "end")
                    break
            return ("if", clauses, orelse, tok.line)
        if tok.kind == "kw" and tok.val == "do":
            self.next()
            body = self.block(("end",))
            self.expect("kw", "end")
            return ("do", body, tok.line)
        if tok.kind == "kw":
            raise LuaError(f"line {tok.line}: '{tok.val}' is outside the supported subset")
        # assignment or call
        first = self.suffixed()
        if self.check("op", "=") or self.check("op", ","):
            targets = [first]
            while self.accept("op", ","):
                targets.append(self.suffixed())
            self.expect("op", "=")
            for tgt in targets:
                if tgt[0] not in ("name", "index"):
                    raise LuaError(f"line {tok.line}: cannot assign to this expression")
            return ("assign", targets, self.exprlist(), tok.line)
        if first[0] not in ("call", "mcall"):
            raise LuaError(f"line {tok.line}: syntax error")
        return ("expr", first, tok.line)

    def funcbody(self, method):
        self.expect("op", "(")
        params = ["self"] if method else []
        if not self.check("op", ")"):
            while True:
                if self.accept("op", "..."):
                    params.append("...")
                    break
                params.append(self.expect("name").val)
                if not self.accept("op", ","):
                    break
        self.expect("op", ")")
        body = self.block(("end",))
        self.expect("kw", "end")
        return ("function", params, body)

    def exprlist(self):
        out = [self.expr()]
        while self.accept("op", ","):
            out.append(self.expr())
        return out

    def expr(self, limit=0):
        tok = self.peek()
        if (tok.kind == "kw" and tok.val == "not") or (tok.kind == "op" and tok.val in ("-", "#")):
            self.next()
            left = ("unop", tok.val, self.expr(_UNARY_PRI))
        else:
            left = self.simple()
        while True:
            tok = self.peek()
            op = tok.val if tok.kind in ("op", "kw") else None
            if op not in _BINPRI or _BINPRI[op][0] <= limit:
                break
            self.next()
            right = self.expr(_BINPRI[op][1])
            left = ("binop", op, left, right)
        return left

    def simple(self):
        tok = self.peek()
        if tok.kind == "num":
            self.next()
            return ("const", tok.val)
        if tok.kind == "str":
            self.next()
            return ("const", tok.val)
        if tok.kind == "kw" and tok.val in ("true", "false", "nil"):
            self.next()
            return ("const", {"true": True, "false": False, "nil": None}[tok.val])
        if tok.kind == "kw" and tok.val == "function":
            self.next()
            return self.funcbody(False)
        if tok.kind == "op" and tok.val == "{":
            return self.table()
        return self.suffixed()

    def table(self):
        self.expect("op", "{")
        fields, pos = [], 0
        while not self.check("op", "}"):
            if self.accept("op", "["):
                key = self.expr()
                self.expect("op", "]")
                self.expect("op", "=")
                fields.append(("kv", key, self.expr()))
            elif self.check("name") and self.peek(1).kind == "op" and self.peek(1).val == "=":
                key = self.next().val
                self.next()
                fields.append(("kv", ("const", key), self.expr()))
            else:
                pos += 1
                fields.append(("pos", pos, self.expr()))
            if not (self.accept("op", ",") or self.accept("op", ";")):
                break
        self.expect("op", "}")
        return ("table", fields)

    def primary(self):
        tok = self.peek()
        if tok.kind == "name":
            self.next()
            return ("name", tok.val)
        if self.accept("op", "("):
            e = self.expr()
            self.expect("op", ")")
            return ("paren", e)
        raise LuaError(f"line {tok.line}: unexpected {tok.val!r}")

    def args(self):
        if self.check("str"):
            return [("const", self.next().val)]
        if self.check("op", "{"):
            return [self.table()]
        self.expect("op", "(")
        out = [] if self.check("op", ")") else self.exprlist()
        self.expect("op", ")")
        return out

    def suffixed(self):
        e = self.primary()
        while True:
            if self.accept("op", "."):
                e = ("index", e, ("const", self.expect("name").val))
            elif self.accept("op", "["):
                k = self.expr()
                self.expect("op", "]")
                e = ("index", e, k)
            elif self.accept("op", ":"):
                name = self.expect("name").val
                e = ("mcall", e, name, self.args())
            elif self.check("op", "(") or self.check("str") or self.check("op", "{"):
                e = ("call", e, self.args())
            else:
                return e


# ------------------------------------------------------------------------------------------
# evaluation
# ------------------------------------------------------------------------------------------
class _Return(Exception):
    def __init__(self, values):
        self.values = values


@dataclass
class Scope:
    vars: dict = field(default_factory=dict)
    parent: "Scope | None" = None

    def find(self, name):
        s = self
        while s is not None:
            if name in s.vars:
                return s
            s = s.parent
        return None


class LuaFunction:
    def __init__(self, params, body, scope, interp):
        self.params, self.body, self.scope, self.interp = params, body, scope, interp

    def __call__(self, *args):
        local = Scope({}, self.scope)
        for k, p in enumerate(self.params):
            local.vars[p] = args[k] if k < len(args) else None
        try:
            self.interp.exec_block(self.body, local)
        except _Return as r:
            return r.values
        return []


def _num(v, what):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        if isinstance(v, str):
            try:
                return float(v)
            except ValueError:
                pass
        raise LuaError(f"attempt to perform arithmetic on a {type_name(v)} value ({what})")
    return v


def type_name(v):
    if v is None:
        return "nil"
    if isinstance(v, bool):
        return "boolean"
    if isinstance(v, (int, float)):
        return "number"
    if isinstance(v, str):
        return "string"
    if isinstance(v, LuaTable):
        return "table"
    if callable(v):
        return "function"
    return "userdata"


def _tostr(v):
    if isinstance(v, float) and v.is_integer() and abs(v) < 1e15:
        return str(int(v))
    return str(v)


def _make_math():
    m = LuaTable()
    for name in ("sin", "cos", "tan", "asin", "acos", "atan", "exp", "sqrt", "floor", "ceil"):
        f = getattr(math, name)
        m[name] = (lambda f: lambda x: [float(f(_num(x, name)))])(f)
    m["log"] = lambda x, b=None: [math.log(_num(x, "log")) if b is None else math.log(_num(x, "log"), _num(b, "log"))]
    m["abs"] = lambda x: [abs(_num(x, "abs"))]
    m["max"] = lambda *a: [max(_num(x, "max") for x in a)]
    m["min"] = lambda *a: [min(_num(x, "min") for x in a)]
    m["pi"] = math.pi
    m["huge"] = math.inf
    return m


class Interp:
    def __init__(self):
        self.globals = Scope({"math": _make_math(),
                              "print": lambda *a: [],
                              "type": lambda v=None: [type_name(v)],
                              "tostring": lambda v=None: [_tostr(v)],
                              "tonumber": lambda v=None: [_tonumber(v)]})

    def exec_block(self, stmts, scope):
        for s in stmts:
            self.exec_stmt(s, scope)

    def exec_stmt(self, s, scope):
        kind = s[0]
        if kind == "local":
            _, names, exprs, _line = s
            vals = self.eval_list(exprs, scope)
            for k, n in enumerate(names):
                scope.vars[n] = vals[k] if k < len(vals) else None
                if isinstance(scope.vars[n], LuaFunction) and exprs and exprs[-1][0] == "function":
                    scope.vars[n].scope = Scope({n: scope.vars[n]}, scope)
        elif kind == "assign":
            _, targets, exprs, line = s
            vals = self.eval_list(exprs, scope)
            for k, tgt in enumerate(targets):
                v = vals[k] if k < len(vals) else None
                if tgt[0] == "name":
                    sc = scope.find(tgt[1]) or self.globals
                    sc.vars[tgt[1]] = v
                else:
                    obj = self.eval(tgt[1], scope)
                    if not isinstance(obj, LuaTable):
                        raise LuaError(f"line {line}: attempt to index a {type_name(obj)} value")
                    obj[self.eval(tgt[2], scope)] = v
        elif kind == "if":
            _, clauses, orelse, _line = s
            for cond, body in clauses:
                if truthy(self.eval(cond, scope)):
                    self.exec_block(body, Scope({}, scope))
                    return
            self.exec_block(orelse, Scope({}, scope))
        elif kind == "do":
            self.exec_block(s[1], Scope({}, scope))
        elif kind == "return":
            raise _Return(self.eval_list(s[1], scope))
        elif kind == "expr":
            self.eval_multi(s[1], scope)

    def eval_list(self, exprs, scope):
        out = []
        for k, e in enumerate(exprs):
            if k == len(exprs) - 1 and e[0] in ("call", "mcall"):
                out.extend(self.eval_multi(e, scope))
            else:
                out.append(self.eval(e, scope))
        return out

    def eval_multi(self, e, scope):
        if e[0] == "call":
            fn = self.eval(e[1], scope)
            args = self.eval_list(e[2], scope)
        elif e[0] == "mcall":
            obj = self.eval(e[1], scope)
            if not isinstance(obj, LuaTable):
                raise LuaError(f"attempt to index a {type_name(obj)} value (method {e[2]})")
            fn = obj[e[2]]
            args = [obj] + self.eval_list(e[3], scope)
        else:
            return [self.eval(e, scope)]
        if not callable(fn):
            raise LuaError(f"attempt to call a {type_name(fn)} value")
        res = fn(*args)
        return list(res) if isinstance(res, (list, tuple)) else [res]

    def eval(self, e, scope):
        kind = e[0]
        if kind == "const":
            return e[1]
        if kind == "name":
            sc = scope.find(e[1])
            return sc.vars[e[1]] if sc else None
        if kind == "paren":
            return self.eval(e[1], scope)
        if kind == "index":
            obj = self.eval(e[1], scope)
            if not isinstance(obj, LuaTable):
                raise LuaError(f"attempt to index a {type_name(obj)} value")
            return obj[self.eval(e[2], scope)]
        if kind in ("call", "mcall"):
            vals = self.eval_multi(e, scope)
            return vals[0] if vals else None
        if kind == "function":
            return LuaFunction(e[1], e[2], scope, self)
        if kind == "table":
            t = LuaTable()
            npos = sum(1 for f in e[1] if f[0] == "pos")
            for f in e[1]:
                if f[0] == "pos":
                    if f[1] == npos and f[2][0] in ("call", "mcall"):
                        for k, v in enumerate(self.eval_multi(f[2], scope)):
                            t[f[1] + k] = v
                    else:
                        t[f[1]] = self.eval(f[2], scope)
                else:
                    t[self.eval(f[1], scope)] = self.eval(f[2], scope)
            return t
        if kind == "unop":
            v = self.eval(e[2], scope)
            if e[1] == "-":
                return -_num(v, "unary minus")
            if e[1] == "not":
                return not truthy(v)
            if isinstance(v, str):
                return len(v)
            if isinstance(v, LuaTable):
                return v.length()
            raise LuaError(f"attempt to get length of a {type_name(v)} value")
        if kind == "binop":
            op = e[1]
            if op == "and":
                a = self.eval(e[2], scope)
                return self.eval(e[3], scope) if truthy(a) else a
            if op == "or":
                a = self.eval(e[2], scope)
                return a if truthy(a) else self.eval(e[3], scope)
            a, b = self.eval(e[2], scope), self.eval(e[3], scope)
            if op == "==":
                return a == b and type_name(a) == type_name(b)
            if op == "~=":
                return not (a == b and type_name(a) == type_name(b))
            if op == "..":
                return _tostr(a) + _tostr(b)
            if op in ("<", ">", "<=", ">="):
                if type_name(a) != type_name(b) or type_name(a) not in ("number", "string"):
                    raise LuaError(f"attempt to compare {type_name(a)} with {type_name(b)}")
                return {"<": a < b, ">": a > b, "<=": a <= b, ">=": a >= b}[op]
            x, y = _num(a, op), _num(b, op)
            if op == "+":
                return x + y
            if op == "-":
                return x - y
            if op == "*":
                return x * y
            if op == "/":
                if y == 0:
                    return math.copysign(math.inf, x) if x != 0 else math.nan
                return x / y
            if op == "%":
                return x - math.floor(x / y) * y
            if op == "^":
                return math.pow(x, y)
        raise LuaError(f"cannot evaluate {kind}")


def _tonumber(v):
    if isinstance(v, (int, float)) and not isinstance(v, bool):
        return v
    try:
        return float(v)
    except (TypeError, ValueError):
        return None


class LuaState:
    """A Lua state that has run one chunk (the problem file)."""

    def __init__(self):
        self.interp = Interp()

    @classmethod
    def from_string(cls, src: str):
        st = cls()
        st.run(src)
        return st

    @classmethod
    def from_file(cls, path: str):
        with open(path) as f:
            return cls.from_string(f.read())

    def run(self, src: str):
        stmts = Parser(tokenize(src)).block(())
        try:
            self.interp.exec_block(stmts, self.interp.globals)
        except _Return:
            pass

    def __getitem__(self, name):
        return self.interp.globals.vars.get(name)


class SolTable:
    """SolTable of the reference (SolTable.hpp:10-117) over this reader's tables.

    ``SolTable(name, state)`` takes a global table of the state, ``SolTable(name, parent)`` a
    field of another SolTable, ``SolTable(table)`` an anonymous table."""

    def __init__(self, name, where=None):
        if where is None:
            if not isinstance(name, LuaTable):
                raise LuaError("SolTable needs a table")
            self.table, self.name = name, "anonymous"
            return
        self.name = name
        obj = where[name] if isinstance(where, LuaState) else where.table[name]
        if not isinstance(obj, LuaTable):
            where_name = "sol::state" if isinstance(where, LuaState) else where.getName()
            raise LuaError(f"Could not find table {name} inside {where_name}!")
        self.table = obj

    def getName(self):  # noqa: N802
        return self.name

    def doesVarExist(self, prop):  # noqa: N802
        return self.table[prop] is not None

    def get(self, prop):
        return to_python(self.table[prop])

    def checkAndGet(self, prop):  # noqa: N802
        v = self.table[prop]
        if v is None:
            raise LuaError(f"Property {prop} was not found in table {self.name}!")
        return to_python(v)

    def call(self, fn, *args):
        """Calls table:fn(args...) (the table is passed as self) and returns the first result
        converted to Python (a Lua array becomes a list)."""
        f = self.table[fn]
        if not callable(f):
            raise LuaError(f"Cannot find method {fn} inside table {self.name}!")
        res = f(self.table, *[from_python(a) for a in args])
        res = list(res) if isinstance(res, (list, tuple)) else [res]
        return to_python(res[0]) if res else None

    def checkCall(self, fn, *args):  # noqa: N802
        if not callable(self.table[fn]):
            raise LuaError(f"Cannot find method {fn} inside table {self.name}!")
        self.call(fn, *args)

    def checkCallNoThrow(self, fn, *args):  # noqa: N802
        if not callable(self.table[fn]):
            return False
        try:
            self.call(fn, *args)
            return True
        except LuaError:
            return False

    def for_each(self, f):
        for k, v in list(self.table.d.items()):
            f(k, v)


def load_problem(path: str) -> SolTable:
    """The ``Problem`` table of a problem file, as Problem.cpp reads it."""
    return SolTable("Problem", LuaState.from_file(path))
