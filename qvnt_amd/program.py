"""The QASM driver loop on a device register (qvnt 0.4.4 src/qasm/sym.rs).

``Program`` holds what the reference's interpreter builds from a source
file: a list of (MultiOp, separator) steps and a trailing MultiOp
(src/qasm/int/ext_op.rs:5-16).  ``Sym`` runs it the way ``Sym::finish``
does (sym.rs:40-74): apply each step's op, then measure into the classical
register, apply the op only if the classical bits match, or reset qubits by
mask.  Measurements take seeded draws (the reference's generator is
unseeded).  The OpenQASM lexer/parser itself is not part of this package.
"""
import ctypes

import numpy as np

from ._lib import SEP_IF, SEP_MEASURE, SEP_NOP, SEP_RESET, Step, check, lib
from .ops import MultiOp, op
from .register import DEFAULT_SEED, CReg, QReg, _bits, uniform


class Program:
    """(op, separator) steps plus a tail op; q_num qubits, c_num classical bits,
    measurement mode "set" or "xor" (MeasureOp, qasm/int/mod.rs:27-31)."""

    def __init__(self, q_num, c_num, measure="set"):
        if measure not in ("set", "xor"):
            raise ValueError("measure is 'set' or 'xor'")
        self.q_num, self.c_num, self.measure_op = q_num, c_num, measure
        self.steps = []          # (MultiOp, sep, a, b)
        self.tail = op.id()

    def _flush(self, sep, a=0, b=0):
        self.steps.append((self.tail, sep, a, b))
        self.tail = op.id()
        return self

    def apply(self, o):
        """Queue gates (ExtOp::push: they join the pending op)."""
        if not isinstance(o, MultiOp):
            raise TypeError("apply expects a MultiOp")
        self.tail = self.tail * o
        return self

    def measure(self, q_arg, c_arg):
        """measure q -> c after the pending gates (Sep::Measure(q_arg, c_arg))."""
        return self._flush(SEP_MEASURE, q_arg, c_arg)

    def if_eq(self, c_mask, value, o):
        """if (c[c_mask] == value) o; (Sep::IfBranch(c, v)).  Pending gates run first."""
        if len(self.tail):
            self._flush(SEP_NOP)
        self.steps.append((o, SEP_IF, c_mask, value))
        return self

    def reset(self, q_mask):
        """reset q after the pending gates (Sep::Reset(q))."""
        return self._flush(SEP_RESET, q_mask)


class Sym:
    """Sym (qasm/sym.rs:7-13): a program with its quantum and classical registers."""

    def __init__(self, program, prec="f64", seed=DEFAULT_SEED, stream=7):
        self.program = program
        self.q_reg = QReg(program.q_num, prec=prec, seed=seed, stream=stream)
        self.c_reg = CReg(program.c_num)
        self._draws = 0

    def reset(self):
        """Sym::reset (sym.rs:35-38)."""
        self.q_reg.reset(0)
        self.c_reg.reset(0)

    def finish(self, draws=None):
        """Sym::finish (sym.rs:40-74); draws default to the register's seeded stream."""
        p = self.program
        n_meas = sum(1 for _, sep, a, _ in p.steps if sep == SEP_MEASURE and a & ((1 << p.q_num) - 1))
        if draws is None:
            draws = [self.q_reg.next_draw() for _ in range(n_meas)]
        d = np.ascontiguousarray(draws, np.float64)
        arr = (Step * max(1, len(p.steps)))()
        for k, (o, sep, a, b) in enumerate(p.steps):
            arr[k].op, arr[k].sep, arr[k].a, arr[k].b = o.handle, sep, a, b
        cval = ctypes.c_uint64(self.c_reg.value)
        used = ctypes.c_size_t(0)
        check(lib.qvnt_reg_run_program(self.q_reg.handle, arr, len(p.steps), p.tail.handle,
                                       1 if p.measure_op == "xor" else 0, ctypes.byref(cval), p.c_num,
                                       d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(d), ctypes.byref(used)))
        self.c_reg.value = cval.value
        self._draws += used.value
        return self

    def measure(self, q_arg, c_arg, u=None):
        """Sym::measure (sym.rs:76-87)."""
        m = self.q_reg.measure_mask(q_arg, u=u).get()
        for q, c in zip(_bits(q_arg), _bits(c_arg)):
            if self.program.measure_op == "set":
                self.c_reg.set(m & q != 0, c)
            else:
                self.c_reg.xor(m & q != 0, c)

    def get_class(self):
        return CReg(self.c_reg.q_num, self.c_reg.value)

    def get_polar_wavefunction(self):
        return self.q_reg.get_polar()

    def get_probabilities(self):
        return self.q_reg.get_probabilities()
