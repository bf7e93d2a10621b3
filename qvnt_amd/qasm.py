"""OpenQASM 2.0 front end: source text -> ``Program`` (qvnt_amd/program.py).

Follows the interpreter of qvnt 0.4.4 (src/qasm/int/mod.rs, gates.rs,
macros.rs, parse.rs):
  * ``qreg`` / ``creg`` append their bits in declaration order; ``a[i]`` is
    the i-th bit of register a, ``a`` alone all of its bits (mod.rs:188-200,
    324-375);
  * a gate call on registers ORs their masks (x/y/z/s/t/h/qft act on every
    bit of the mask); rx/ry/rz need one bit, rxx/ryy/rzz and the swap family
    two; u1/u2/u3 one bit with 1/2/3 angles; ``sdg``/``tdg`` map to s/t as
    in the reference (gates.rs:14-23); a leading ``c`` makes the first
    argument a control of the rest (gates.rs:76-93);
  * ``gate`` defines a macro of gate calls on its formal registers, with
    angle expressions over its parameters (macros.rs); ``opaque`` and
    ``barrier`` do nothing; ``include`` lines are skipped;
  * ``measure q -> c`` needs as many bits on both sides, ``reset q``, and
    ``if (c == v) gate`` (the whole classical register c) become the
    program's separators (mod.rs:207-230, 280-298);
  * angle expressions: numbers, pi, + - * / % ^ (right-associative, above
    unary minus), parentheses, sqrt exp ln abs floor ceil round, atan2,
    max, min (the reference's meval context, parse.rs:7-25).
"""
import math
import re

from .ops import op as _op
from .program import Program

_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*)
  | (?P<num>(\d+\.\d*|\.\d+|\d+)([eE][+-]?\d+)?)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<str>"[^"]*")
  | (?P<sym>->|==|[;,()\[\]{}+\-*/%^])
""", re.VERBOSE)


class QasmError(ValueError):
    pass


def _lex(src):
    out, pos = [], 0
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise QasmError("unexpected character %r at offset %d" % (src[pos], pos))
        pos = m.end()
        if m.lastgroup != "ws":
            out.append((m.lastgroup, m.group()))
    return out


# ------------------------------------------------------------ expressions
_FUNCS1 = {"sqrt": math.sqrt, "exp": math.exp, "ln": math.log, "abs": abs, "floor": math.floor,
           "ceil": math.ceil, "round": lambda x: math.floor(abs(x) + 0.5) * (1 if x >= 0 else -1)}


def evaluate(expr, variables=None):
    """An angle expression in the reference's meval context (parse.rs:7-25)."""
    toks = _lex(expr)
    env = {"pi": math.pi}
    env.update(variables or {})
    pos = [0]

    def peek():
        return toks[pos[0]][1] if pos[0] < len(toks) else None

    def take(want=None):
        if pos[0] >= len(toks):
            raise QasmError("unexpected end of expression %r" % expr)
        t = toks[pos[0]]
        if want is not None and t[1] != want:
            raise QasmError("expected %r in expression %r" % (want, expr))
        pos[0] += 1
        return t

    def atom():
        kind, text = take()
        if kind == "num":
            return float(text)
        if text == "(":
            v = add()
            take(")")
            return v
        if kind == "id":
            if peek() == "(":
                take("(")
                args = [add()]
                while peek() == ",":
                    take(",")
                    args.append(add())
                take(")")
                if text in _FUNCS1 and len(args) == 1:
                    return float(_FUNCS1[text](args[0]))
                if text == "atan2" and len(args) == 2:
                    return math.atan2(args[0], args[1])
                if text in ("max", "min") and args:
                    return float((max if text == "max" else min)(args))
                raise QasmError("unknown function %s/%d in %r" % (text, len(args), expr))
            if text not in env:
                raise QasmError("unknown variable %r in %r" % (text, expr))
            return float(env[text])
        raise QasmError("unexpected %r in expression %r" % (text, expr))

    def power():
        base = atom()
        if peek() == "^":
            take("^")
            return base ** unary()      # right-associative
        return base

    def unary():
        if peek() in ("-", "+"):
            sign = -1.0 if take()[1] == "-" else 1.0
            return sign * unary()
        return power()

    def mul():
        v = unary()
        while peek() in ("*", "/", "%"):
            o = take()[1]
            r = unary()
            v = v * r if o == "*" else v / r if o == "/" else math.fmod(v, r)
        return v

    def add():
        v = mul()
        while peek() in ("+", "-"):
            o = take()[1]
            r = mul()
            v = v + r if o == "+" else v - r
        return v

    v = add()
    if pos[0] != len(toks):
        raise QasmError("trailing tokens in expression %r" % expr)
    return v


# ------------------------------------------------------------------ gates
def _bits(mask):
    return bin(mask).count("1")


def _gate(name, regs, args):
    """gates.rs:76-129: one gate call on register masks with evaluated angles."""
    if name[:1] in ("c", "C"):
        if not regs:
            raise QasmError("%s: wrong number of registers (0)" % name)
        ctrl, rest = regs[0], regs[1:]
        inner = _gate(name[1:], rest, args)
        out = inner.c(ctrl)
        if out is None:
            raise QasmError("%s: control mask %#x overlaps the gate's qubits %#x" % (name, ctrl, inner.act_on()))
        return out
    low = name.lower()
    mask = 0
    for r in regs:
        mask |= r
    table_any = {"x": _op.x, "y": _op.y, "z": _op.z, "s": _op.s, "sdg": _op.s, "t": _op.t, "tdg": _op.t,
                 "h": _op.h, "qft": _op.qft}
    if name not in (low, low.upper()):
        raise QasmError("unknown gate %s" % name)
    if low in table_any:
        if mask == 0:
            raise QasmError("%s: wrong number of registers (0)" % name)
        if args:
            raise QasmError("%s: wrong number of arguments (%d)" % (name, len(args)))
        return table_any[low](mask)
    rot = {"rx": (1, _op.rx), "ry": (1, _op.ry), "rz": (1, _op.rz), "rxx": (2, _op.rxx), "ryy": (2, _op.ryy),
           "rzz": (2, _op.rzz)}
    if low in rot:
        nbits, ctor = rot[low]
        if _bits(mask) != nbits:
            raise QasmError("%s: wrong number of registers (%d)" % (name, _bits(mask)))
        if len(args) != 1:
            raise QasmError("%s: wrong number of arguments (%d)" % (name, len(args)))
        return ctor(args[0], mask)
    swaps = {"swap": _op.swap, "sqrt_swap": _op.sqrt_swap, "i_swap": _op.i_swap, "sqrt_i_swap": _op.sqrt_i_swap}
    if low in swaps:
        if _bits(mask) != 2:
            raise QasmError("%s: wrong number of registers (%d)" % (name, _bits(mask)))
        if args:
            raise QasmError("%s: wrong number of arguments (%d)" % (name, len(args)))
        return swaps[low](mask)
    us = {"u1": (1, _op.u1), "u2": (2, _op.u2), "u3": (3, _op.u3)}
    if low in us:
        nargs, ctor = us[low]
        if _bits(mask) != 1:
            raise QasmError("%s: wrong number of registers (%d)" % (name, _bits(mask)))
        if len(args) != nargs:
            raise QasmError("%s: wrong number of arguments (%d)" % (name, len(args)))
        return ctor(*args, mask)
    raise QasmError("unknown gate %s" % name)


class _Macro:
    """A ``gate`` definition (macros.rs): calls on formal registers."""

    def __init__(self, name, regs, params, body):
        self.name, self.regs, self.params, self.body = name, regs, params, body
        for gname, gregs, gargs in body:
            for r in gregs:
                if isinstance(r, tuple):
                    raise QasmError("indexing qubits (%s[%d]) isn't allowed in gate %s" % (r[0], r[1], name))
                if r not in regs:
                    raise QasmError("no such register %r in gate %s" % (r, name))

    def expand(self, regs, args, macros):
        if len(regs) != len(self.regs):
            raise QasmError("%s: wrong number of registers (%d)" % (self.name, len(regs)))
        if len(args) != len(self.params):
            raise QasmError("%s: wrong number of arguments (%d)" % (self.name, len(args)))
        rmap = dict(zip(self.regs, regs))
        amap = dict(zip(self.params, args))
        out = _op.id()
        for gname, gregs, gargs in self.body:
            masks = [rmap[r] for r in gregs]
            vals = [evaluate(e, amap) for e in gargs]
            if gname in macros:
                if gname == self.name:
                    raise QasmError("recursive macro call %s" % gname)
                out *= macros[gname].expand(masks, vals, macros)
            else:
                out *= _gate(gname, masks, vals)
        return out


# ----------------------------------------------------------------- parser
class _Parser:
    def __init__(self, src):
        self.t = _lex(src)
        self.i = 0

    def peek(self, k=0):
        j = self.i + k
        return self.t[j] if j < len(self.t) else (None, None)

    def take(self, want=None):
        if self.i >= len(self.t):
            raise QasmError("unexpected end of source")
        kind, text = self.t[self.i]
        if want is not None and text != want:
            raise QasmError("expected %r, found %r" % (want, text))
        self.i += 1
        return kind, text

    def ident(self):
        kind, text = self.take()
        if kind != "id":
            raise QasmError("expected an identifier, found %r" % text)
        return text

    def integer(self):
        kind, text = self.take()
        if kind != "num" or not text.isdigit():
            raise QasmError("expected an integer, found %r" % text)
        return int(text)

    def argument(self):
        name = self.ident()
        if self.peek()[1] == "[":
            self.take("[")
            idx = self.integer()
            self.take("]")
            return (name, idx)
        return name

    def arguments(self):
        args = [self.argument()]
        while self.peek()[1] == ",":
            self.take(",")
            args.append(self.argument())
        return args

    def expressions(self):
        """Raw expression strings between parentheses, split at top-level commas."""
        self.take("(")
        exprs, depth, cur = [], 0, []
        while True:
            kind, text = self.take()
            if text == "(":
                depth += 1
            elif text == ")":
                if depth == 0:
                    break
                depth -= 1
            elif text == "," and depth == 0:
                exprs.append(" ".join(cur))
                cur = []
                continue
            cur.append(text)
        if cur:
            exprs.append(" ".join(cur))
        return exprs

    def gate_call(self):
        name = self.ident()
        exprs = self.expressions() if self.peek()[1] == "(" else []
        regs = self.arguments()
        self.take(";")
        return name, regs, exprs

    def statements(self):
        out = []
        while self.i < len(self.t):
            kind, text = self.peek()
            if text == "OPENQASM":
                self.take()
                self.take()
                self.take(";")
            elif text == "include":
                self.take()
                self.take()
                self.take(";")
            elif text in ("qreg", "creg"):
                self.take()
                name = self.ident()
                self.take("[")
                size = self.integer()
                self.take("]")
                self.take(";")
                out.append((text, name, size))
            elif text == "gate":
                self.take()
                name = self.ident()
                params = []
                if self.peek()[1] == "(":
                    self.take("(")
                    if self.peek()[1] != ")":
                        params.append(self.ident())
                        while self.peek()[1] == ",":
                            self.take(",")
                            params.append(self.ident())
                    self.take(")")
                regs = [self.ident()]
                while self.peek()[1] == ",":
                    self.take(",")
                    regs.append(self.ident())
                self.take("{")
                body = []
                while self.peek()[1] != "}":
                    if self.peek()[1] in ("measure", "reset", "barrier", "if", "gate", "qreg", "creg", "opaque"):
                        raise QasmError("operation %r isn't allowed in gate definition %s" % (self.peek()[1], name))
                    body.append(self.gate_call())
                self.take("}")
                out.append(("gate", name, params, regs, body))
            elif text == "opaque":
                while self.take()[1] != ";":
                    pass
            elif text == "barrier":
                self.take()
                self.arguments()
                self.take(";")
            elif text == "reset":
                self.take()
                arg = self.argument()
                self.take(";")
                out.append(("reset", arg))
            elif text == "measure":
                self.take()
                q = self.argument()
                self.take("->")
                c = self.argument()
                self.take(";")
                out.append(("measure", q, c))
            elif text == "if":
                self.take()
                self.take("(")
                c = self.ident()
                self.take("==")
                v = self.integer()
                self.take(")")
                if self.peek()[1] in ("measure", "reset", "barrier", "if"):
                    raise QasmError("operation %r isn't allowed in an if block" % self.peek()[1])
                out.append(("if", c, v, self.gate_call()))
            elif kind == "id":
                out.append(("call",) + self.gate_call())
            else:
                raise QasmError("unexpected %r" % text)
        return out


def parse(source, measure="set"):
    """OpenQASM 2.0 source -> Program (measure: "set" or "xor", MeasureOp)."""
    stmts = _Parser(source).statements()
    if not stmts:
        raise QasmError("empty source")
    qbits, cbits, macros = [], [], {}
    for s in stmts:
        if s[0] in ("qreg", "creg"):
            name, size = s[1], s[2]
            if len(name) >= 32:
                raise QasmError("identifier %s is too long" % name)
            if size >= 64:
                raise QasmError("register %s is too large (%d)" % (name, size))
            if name in qbits or name in cbits:
                raise QasmError("duplicate register %s" % name)
            (qbits if s[0] == "qreg" else cbits).extend([name] * size)
    prog = Program(len(qbits), len(cbits), measure=measure)

    def mask_of(bits, arg, kind):
        name, idx = (arg if isinstance(arg, tuple) else (arg, None))
        m = sum(1 << k for k, b in enumerate(bits) if b == name)
        if m == 0:
            raise QasmError("no %s register %s" % (kind, name))
        if idx is None:
            return m
        sel = [1 << k for k, b in enumerate(bits) if b == name]
        if idx >= len(sel):
            raise QasmError("index %s[%d] out of range" % (name, idx))
        return sel[idx]

    def call(name, regs, exprs):
        masks = [mask_of(qbits, r, "quantum") for r in regs]
        vals = [evaluate(e) for e in exprs]
        if name in macros:
            return macros[name].expand(masks, vals, macros)
        return _gate(name, masks, vals)

    for s in stmts:
        kind = s[0]
        if kind == "gate":
            _, name, params, regs, body = s
            if name in macros:
                raise QasmError("macro %s is already defined" % name)
            macros[name] = _Macro(name, regs, params, body)
        elif kind == "call":
            prog.apply(call(*s[1:]))
        elif kind == "measure":
            q, c = mask_of(qbits, s[1], "quantum"), mask_of(cbits, s[2], "classical")
            if _bits(q) != _bits(c):
                raise QasmError("measure: %d qubits into %d bits" % (_bits(q), _bits(c)))
            prog.measure(q, c)
        elif kind == "reset":
            prog.reset(mask_of(qbits, s[1], "quantum"))
        elif kind == "if":
            c = mask_of(cbits, s[1], "classical")
            prog.if_eq(c, s[2], call(*s[3]))
    return prog
