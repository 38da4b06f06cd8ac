"""OpenQASM 2.0 programs on the GPU register (src/openqasm.rs:52-190).

``run_openqasm(path)`` parses an OpenQASM 2.0 file, creates one GPU
:class:`~quaru_amd.register.Register` per ``qreg`` and one list of bools per
``creg``, and applies the program's basic operations in order, as the
reference's ``run_openqasm`` does:

* ``U(theta, phi, lambda) q`` -> :func:`quaru_amd.operation.u` with each angle
  passed through float32 first (the reference's parser hands f32 angles to
  ``u``, src/openqasm.rs:98);
* ``CX a, b`` -> ``cnot(a, b)``; both qubits must be in one register
  (src/openqasm.rs:102-104, the same error otherwise);
* ``measure q -> c`` -> ``measure``; ``reset q`` -> a fresh |0...0> register;
  ``reset c`` -> all false;
* ``if (c == n) op`` -> op only when the creg, read with c[0] as the least
  significant bit, equals n (src/openqasm.rs:165-189).

The parser is this module's own: OpenQASM 2.0 as published (Cross et al.,
"Open Quantum Assembly Language", arXiv:1707.03429) -- ``qreg``, ``creg``,
``gate`` definitions with parameters, ``opaque`` (rejected when applied),
``barrier`` (ignored), ``include`` (files relative to the program;
``qelib1.inc`` is built in, written from the gate definitions in that paper),
register-wide arguments (broadcast over the register), and real expressions
with ``pi``, ``+ - * / ^``, unary minus and sin, cos, tan, exp, ln, sqrt.
The reference uses the third-party openqasm-parser 1.0.0 crate, which is not
available here; expressions are evaluated in FP64 and rounded to f32 once,
at the U gate.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import operation as ops
from .register import Register

# The standard header, from the OpenQASM 2.0 paper's qelib1.inc listing
# (gate names, parameters and decompositions into U and CX).
QELIB1 = """
gate u3(theta,phi,lambda) q { U(theta,phi,lambda) q; }
gate u2(phi,lambda) q { U(pi/2,phi,lambda) q; }
gate u1(lambda) q { U(0,0,lambda) q; }
gate cx c,t { CX c,t; }
gate id a { U(0,0,0) a; }
gate u0(gamma) q { U(0,0,0) q; }
gate x a { u3(pi,0,pi) a; }
gate y a { u3(pi,pi/2,pi/2) a; }
gate z a { u1(pi) a; }
gate h a { u2(0,pi) a; }
gate s a { u1(pi/2) a; }
gate sdg a { u1(-pi/2) a; }
gate t a { u1(pi/4) a; }
gate tdg a { u1(-pi/4) a; }
gate rx(theta) a { u3(theta,-pi/2,pi/2) a; }
gate ry(theta) a { u3(theta,0,0) a; }
gate rz(phi) a { u1(phi) a; }
gate cz a,b { h b; cx a,b; h b; }
gate cy a,b { sdg b; cx a,b; s b; }
gate swap a,b { cx a,b; cx b,a; cx a,b; }
gate ch a,b { h b; sdg b; cx a,b; h b; t b; cx a,b; t b; h b; s b; x b; s a; }
gate ccx a,b,c { h c; cx b,c; tdg c; cx a,c; t c; cx b,c; tdg c; cx a,c; t b; t c; h c;
                 cx a,b; t a; tdg b; cx a,b; }
gate crz(lambda) a,b { u1(lambda/2) b; cx a,b; u1(-lambda/2) b; cx a,b; }
gate cu1(lambda) a,b { u1(lambda/2) a; cx a,b; u1(-lambda/2) b; cx a,b; u1(lambda/2) b; }
gate cu3(theta,phi,lambda) c,t { u1((lambda-phi)/2) t; cx c,t;
                                 u3(-theta/2,0,-(phi+lambda)/2) t; cx c,t;
                                 u3(theta/2,phi,0) t; }
"""


class OpenQASMError(Exception):
    """OpenQASMError (src/openqasm.rs:24-32): ``kind`` is "parse" for a
    syntax or semantic error, "run" for an error while running."""

    def __init__(self, kind: str, message: str):
        super().__init__(f"{kind}: {message}")
        self.kind = kind


# ---- lexer ----------------------------------------------------------------------

_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*)
  | (?P<real>(?:\d+\.\d*|\.\d+)(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+)
  | (?P<int>\d+)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<str>"[^"]*")
  | (?P<arrow>->)
  | (?P<eq>==)
  | (?P<sym>[;,()\[\]{}+\-*/^])
""", re.X)


def _tokens(text: str) -> List[Tuple[str, str]]:
    out, pos = [], 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise OpenQASMError("parse", f"unexpected character {text[pos]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group()))
    return out


# ---- expressions ------------------------------------------------------------------

_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp,
          "ln": math.log, "sqrt": math.sqrt}


@dataclass
class _Expr:
    """A parsed real expression: a tree of tuples, evaluated with a parameter
    environment."""
    tree: tuple

    def eval(self, env: Dict[str, float]) -> float:
        return _eval(self.tree, env)


def _eval(t, env):
    k = t[0]
    if k == "num":
        return t[1]
    if k == "pi":
        return math.pi
    if k == "var":
        if t[1] not in env:
            raise OpenQASMError("parse", f"unknown parameter {t[1]}")
        return env[t[1]]
    if k == "neg":
        return -_eval(t[1], env)
    if k == "fn":
        return _FUNCS[t[1]](_eval(t[2], env))
    a, b = _eval(t[1], env), _eval(t[2], env)
    if k == "+":
        return a + b
    if k == "-":
        return a - b
    if k == "*":
        return a * b
    if k == "/":
        return a / b
    if k == "^":
        return a ** b
    raise OpenQASMError("parse", f"bad expression node {k}")


# ---- program ------------------------------------------------------------------------

@dataclass
class _GateDef:
    params: List[str]
    args: List[str]
    body: Optional[list]  # None for opaque


@dataclass
class _Stmt:
    kind: str            # "U", "CX", "call", "measure", "reset", "barrier"
    name: str = ""
    exprs: List[_Expr] = field(default_factory=list)
    args: List[tuple] = field(default_factory=list)  # (reg, index or None)
    cond: Optional[Tuple[str, int]] = None


class _Parser:
    def __init__(self, toks, base_dir):
        self.t = toks
        self.i = 0
        self.base = base_dir

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else ("eof", "")

    def take(self, value=None, kind=None):
        tok = self.peek()
        if (value is not None and tok[1] != value) or (kind is not None and tok[0] != kind):
            want = value if value is not None else kind
            raise OpenQASMError("parse", f"expected {want}, got {tok[1]!r}")
        self.i += 1
        return tok

    # expression grammar: sum := term (('+'|'-') term)*; term := unary (('*'|'/') unary)*;
    # unary := '-' unary | power; power := atom ('^' unary)?
    def expr(self):
        node = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            node = (op, node, self.term())
        return node

    def term(self):
        node = self.unary()
        while self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            node = (op, node, self.unary())
        return node

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            return ("neg", self.unary())
        if self.peek()[1] == "+":
            self.take()
            return self.unary()
        return self.power()

    def power(self):
        node = self.atom()
        if self.peek()[1] == "^":
            self.take()
            node = ("^", node, self.unary())
        return node

    def atom(self):
        kind, val = self.take()
        if kind in ("real", "int"):
            return ("num", float(val))
        if kind == "id":
            if val == "pi":
                return ("pi",)
            if val in _FUNCS:
                self.take("(")
                e = self.expr()
                self.take(")")
                return ("fn", val, e)
            return ("var", val)
        if val == "(":
            e = self.expr()
            self.take(")")
            return e
        raise OpenQASMError("parse", f"unexpected {val!r} in expression")

    def argument(self):
        name = self.take(kind="id")[1]
        if self.peek()[1] == "[":
            self.take("[")
            idx = int(self.take(kind="int")[1])
            self.take("]")
            return (name, idx)
        return (name, None)

    def arglist(self):
        args = [self.argument()]
        while self.peek()[1] == ",":
            self.take(",")
            args.append(self.argument())
        return args

    def exprlist(self):
        out = []
        if self.peek()[1] == ")":
            return out
        out.append(_Expr(self.expr()))
        while self.peek()[1] == ",":
            self.take(",")
            out.append(_Expr(self.expr()))
        return out

    def qop(self, cond=None):
        kind, val = self.peek()
        if val == "U":
            self.take()
            self.take("(")
            ex = self.exprlist()
            self.take(")")
            args = [self.argument()]
            self.take(";")
            if len(ex) != 3:
                raise OpenQASMError("parse", "U takes three parameters")
            return _Stmt("U", "U", ex, args, cond)
        if val == "CX":
            self.take()
            args = self.arglist()
            self.take(";")
            if len(args) != 2:
                raise OpenQASMError("parse", "CX takes two arguments")
            return _Stmt("CX", "CX", [], args, cond)
        if val == "measure":
            self.take()
            q = self.argument()
            self.take(kind="arrow")
            c = self.argument()
            self.take(";")
            return _Stmt("measure", "measure", [], [q, c], cond)
        if val == "reset":
            self.take()
            a = self.argument()
            self.take(";")
            return _Stmt("reset", "reset", [], [a], cond)
        if val == "barrier":
            self.take()
            args = self.arglist()
            self.take(";")
            return _Stmt("barrier", "barrier", [], args, cond)
        if kind == "id":
            name = self.take()[1]
            ex = []
            if self.peek()[1] == "(":
                self.take("(")
                ex = self.exprlist()
                self.take(")")
            args = self.arglist()
            self.take(";")
            return _Stmt("call", name, ex, args, cond)
        raise OpenQASMError("parse", f"unexpected {val!r}")

    def program(self, prog):
        while self.peek()[0] != "eof":
            kind, val = self.peek()
            if val == "OPENQASM":
                self.take()
                ver = self.take()[1]
                if not ver.startswith("2"):
                    raise OpenQASMError("parse", f"OpenQASM version {ver} is not 2.x")
                self.take(";")
            elif val == "include":
                self.take()
                fname = self.take(kind="str")[1].strip('"')
                self.take(";")
                prog.include(fname, self.base)
            elif val in ("qreg", "creg"):
                self.take()
                name = self.take(kind="id")[1]
                self.take("[")
                size = int(self.take(kind="int")[1])
                self.take("]")
                self.take(";")
                if name in prog.qregs or name in prog.cregs:
                    raise OpenQASMError("parse", f"register {name} declared twice")
                (prog.qregs if val == "qreg" else prog.cregs)[name] = size
            elif val in ("gate", "opaque"):
                self.take()
                name = self.take(kind="id")[1]
                params = []
                if self.peek()[1] == "(":
                    self.take("(")
                    if self.peek()[1] != ")":
                        params.append(self.take(kind="id")[1])
                        while self.peek()[1] == ",":
                            self.take(",")
                            params.append(self.take(kind="id")[1])
                    self.take(")")
                args = [self.take(kind="id")[1]]
                while self.peek()[1] == ",":
                    self.take(",")
                    args.append(self.take(kind="id")[1])
                body = None
                if val == "gate":
                    self.take("{")
                    body = []
                    while self.peek()[1] != "}":
                        body.append(self.qop())
                    self.take("}")
                else:
                    self.take(";")
                prog.gates[name] = _GateDef(params, args, body)
            elif val == "if":
                self.take()
                self.take("(")
                creg = self.take(kind="id")[1]
                self.take(kind="eq")
                n = int(self.take(kind="int")[1])
                self.take(")")
                prog.body.append(self.qop((creg, n)))
            else:
                prog.body.append(self.qop())


class OpenQASMProgram:
    """Registers, gate definitions and the statement list of a program."""

    def __init__(self):
        self.qregs: Dict[str, int] = {}
        self.cregs: Dict[str, int] = {}
        self.gates: Dict[str, _GateDef] = {}
        self.body: List[_Stmt] = []

    def include(self, fname, base):
        if fname == "qelib1.inc":
            text = QELIB1
            sub_base = base
        else:
            path = os.path.join(base, fname)
            try:
                text = open(path).read()
            except OSError as e:
                raise OpenQASMError("parse", f"cannot include {fname}: {e}") from None
            sub_base = os.path.dirname(path)
        _Parser(_tokens(text), sub_base).program(self)

    # -- expansion to basic operations --------------------------------------------
    def _check_arg(self, a, quantum: bool):
        regs = self.qregs if quantum else self.cregs
        name, idx = a
        if name not in regs:
            raise OpenQASMError("parse", f"unknown {'qreg' if quantum else 'creg'} {name}")
        if idx is not None and not 0 <= idx < regs[name]:
            raise OpenQASMError("parse", f"index {idx} out of range for {name}")

    def _broadcast(self, args, quantum=True):
        """Expand register-wide arguments: all whole registers must have one
        size; single qubits are repeated."""
        sizes = {(self.qregs if quantum else self.cregs)[n] for n, i in args if i is None}
        if len(sizes) > 1:
            raise OpenQASMError("parse", "register arguments of different sizes")
        if not sizes:
            return [args]
        k = sizes.pop()
        return [[(n, j if i is None else i) for n, i in args] for j in range(k)]

    def basic_operations(self):
        """[(condition, op)] with op one of ("U", (t, p, l), q), ("CX", a, b),
        ("measure", q, c), ("resetq", reg), ("resetc", reg)."""
        out = []
        for st in self.body:
            self._expand(st, {}, {}, st.cond, out, depth=0)
        return out

    def _expand(self, st, env, qmap, cond, out, depth):
        if depth > 64:
            raise OpenQASMError("parse", "gate definitions nest too deeply")

        def resolve(a):
            # inside a gate body, arguments are names bound to (reg, idx)
            if a[1] is None and a[0] in qmap:
                return qmap[a[0]]
            return a

        if st.kind == "barrier":
            return
        if st.kind == "measure":
            q, c = st.args
            self._check_arg(q, True)
            self._check_arg(c, False)
            if (q[1] is None) != (c[1] is None):
                raise OpenQASMError("parse", "measure needs two registers or two bits")
            if q[1] is None:
                if self.qregs[q[0]] != self.cregs[c[0]]:
                    raise OpenQASMError("parse", "measure of registers of different sizes")
                for j in range(self.qregs[q[0]]):
                    out.append((cond, ("measure", (q[0], j), (c[0], j))))
            else:
                out.append((cond, ("measure", q, c)))
            return
        if st.kind == "reset":
            a = st.args[0]
            if a[0] in self.qregs:
                out.append((cond, ("resetq", a[0])))
            elif a[0] in self.cregs:
                out.append((cond, ("resetc", a[0])))
            else:
                raise OpenQASMError("parse", f"unknown register {a[0]}")
            return
        args = [resolve(a) for a in st.args]
        for a in args:
            self._check_arg(a, True)
        vals = [e.eval(env) for e in st.exprs]
        for inst in self._broadcast(args):
            if st.kind == "U":
                out.append((cond, ("U", tuple(vals), inst[0])))
            elif st.kind == "CX":
                if inst[0] == inst[1]:
                    raise OpenQASMError("parse", "CX on one qubit")
                out.append((cond, ("CX", inst[0], inst[1])))
            else:
                g = self.gates.get(st.name)
                if g is None:
                    raise OpenQASMError("parse", f"unknown gate {st.name}")
                if g.body is None:
                    raise OpenQASMError("run", f"opaque gate {st.name} cannot be simulated")
                if len(vals) != len(g.params) or len(inst) != len(g.args):
                    raise OpenQASMError("parse", f"wrong number of arguments to {st.name}")
                sub_env = dict(zip(g.params, vals))
                sub_map = dict(zip(g.args, inst))
                for inner in g.body:
                    self._expand(inner, sub_env, sub_map, cond, out, depth + 1)


def parse_openqasm(path: str) -> OpenQASMProgram:
    """Parse an OpenQASM 2.0 file (includes resolved relative to it)."""
    try:
        text = open(path).read()
    except OSError as e:
        raise OpenQASMError("parse", str(e)) from None
    return parse_openqasm_str(text, os.path.dirname(os.path.abspath(path)))


def parse_openqasm_str(text: str, base_dir: str = ".") -> OpenQASMProgram:
    prog = OpenQASMProgram()
    _Parser(_tokens(text), base_dir).program(prog)
    return prog


@dataclass
class OpenQASMResult:
    """OpenQASMResult (src/openqasm.rs:12-21)."""
    qregs: Dict[str, Register]
    cregs: Dict[str, List[bool]]
    program: OpenQASMProgram


def _f32(x: float) -> float:
    return float(np.float32(x))


def creg_to_int(creg: Sequence[bool]) -> int:
    """creg_to_u32 (src/openqasm.rs:175-189): creg[0] is the least
    significant bit; more than 32 bits is an error."""
    if len(creg) > 32:
        raise OpenQASMError("run", "Registers larger than 32 bits are not allowed")
    v = 0
    for bit in reversed(creg):
        v = (v << 1) | (1 if bit else 0)
    return v


def run_program(prog: OpenQASMProgram, dtype: str = "c128",
                seed: Optional[int] = None) -> OpenQASMResult:
    """run_openqasm (src/openqasm.rs:52-142) on a parsed program."""
    basic = prog.basic_operations()
    res = OpenQASMResult(
        qregs={name: Register.new([False] * size, dtype, seed=None if seed is None else seed + k)
               for k, (name, size) in enumerate(sorted(prog.qregs.items()))},
        cregs={name: [False] * size for name, size in prog.cregs.items()},
        program=prog)
    for cond, op in basic:
        if cond is not None and creg_to_int(res.cregs[cond[0]]) != cond[1]:
            continue
        kind = op[0]
        if kind == "U":
            (t, p, l), (reg, q) = op[1], op[2]
            res.qregs[reg].apply(ops.u(_f32(t), _f32(p), _f32(l), q))
        elif kind == "CX":
            (ra, a), (rb, b) = op[1], op[2]
            if ra != rb:
                raise OpenQASMError("run", "The simulator currently does not support applying "
                                           "CX on qubits in two different registers")
            res.qregs[ra].apply(ops.cnot(a, b))
        elif kind == "measure":
            (qr, q), (cr, c) = op[1], op[2]
            res.cregs[cr][c] = res.qregs[qr].measure(q)
        elif kind == "resetq":
            # a fresh |0...0> register (src/openqasm.rs:127-133); the uniform
            # stream carries on, as the reference's thread-local generator does
            old = res.qregs[op[1]]
            fresh = Register.new([False] * old.size(), old.dtype)
            fresh._rng = old._rng
            res.qregs[op[1]] = fresh
            old.close()
        elif kind == "resetc":
            res.cregs[op[1]] = [False] * len(res.cregs[op[1]])
    return res


def run_openqasm(path: str, dtype: str = "c128", seed: Optional[int] = None) -> OpenQASMResult:
    """Read an OpenQASM 2.0 file and run it on GPU registers."""
    return run_program(parse_openqasm(path), dtype, seed)
