"""The OpenQASM 2.0 front end (qvnt_amd/qasm.py) against the reference
interpreter's rules (src/qasm/int/{mod,gates,macros,parse}.rs): register
masks, the gate table, controls, macros, separators and expressions."""
import math

import pytest

from qvnt_amd import op, qasm
from qvnt_amd.program import SEP_IF, SEP_MEASURE, SEP_NOP, SEP_RESET

HEADER = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'


def test_expressions():
    """parse.rs:50-63 and meval's operators."""
    assert qasm.evaluate("2 * pi / 16") == 2.0 * math.pi / 16.0
    assert qasm.evaluate("-2^2") == -4.0            # ^ binds above unary minus
    assert qasm.evaluate("2^3^2") == 2.0 ** 9       # right-associative
    assert qasm.evaluate("7 % 4 - (1 + 2) * 3") == -6.0
    assert qasm.evaluate("sqrt(2.0) * sqrt(2.0)") == pytest.approx(2.0)
    assert qasm.evaluate("max(1, 3, 2) + min(4, -1)") == 2.0
    assert qasm.evaluate("atan2(1, 1)") == math.pi / 4
    assert qasm.evaluate("round(2.5) + round(-2.5) + floor(-0.5) + ceil(0.5)") == 0.0
    assert qasm.evaluate("2 * x / 16", {"x": math.pi}) == 2.0 * math.pi / 16.0
    with pytest.raises(qasm.QasmError):
        qasm.evaluate("2 * x / 16")                 # unknown variable
    with pytest.raises(qasm.QasmError):
        qasm.evaluate("sin(1)")                     # not in the reference's context


def test_registers_and_gate_table():
    p = qasm.parse(HEADER + """qreg a[1]; qreg b[2]; creg c[2];
        x b; h a[0]; cx a[0], b[1]; ccx a, b[0], b[1]; rzz(pi) a[0], b[0]; sdg b[1]; tdg a;
        crz(pi/2) b[0], a[0]; u2(0, pi) b[1];""")
    assert (p.q_num, p.c_num) == (3, 2)
    assert p.steps == []
    want = (op.x(0b110) * op.h(0b001) * op.x(0b100).c(0b001) * op.x(0b100).c(0b010).c(0b001)
            * op.rzz(math.pi, 0b011) * op.s(0b100) * op.t(0b001) * op.rz(math.pi / 2, 0b001).c(0b010)
            * op.u2(0.0, math.pi, 0b100))
    assert p.tail.name() == want.name()


def test_separators():
    p = qasm.parse(HEADER + """qreg q[3]; creg c[3]; creg d[1];
        h q; measure q[1] -> c[2]; if (c == 4) x q[0]; reset q[1]; measure q -> c; z q[2];""")
    seps = [(s[1], s[2], s[3]) for s in p.steps]
    assert seps == [(SEP_MEASURE, 0b010, 0b100), (SEP_IF, 0b111, 4), (SEP_RESET, 0b010, 0),
                    (SEP_MEASURE, 0b111, 0b111)]
    assert p.steps[0][0].name() == op.h(0b111).name()
    assert p.steps[1][0].name() == op.x(0b001).name()
    assert len(p.steps[2][0]) == 0 and len(p.steps[3][0]) == 0
    assert p.tail.name() == op.z(0b100).name()
    # pending gates before an if run as their own (Nop) step
    p = qasm.parse(HEADER + "qreg q[2]; creg c[1]; measure q[0] -> c[0]; h q[1]; if (c == 1) x q[0];")
    assert [s[1] for s in p.steps] == [SEP_MEASURE, SEP_NOP, SEP_IF]


def test_macros():
    p = qasm.parse(HEADER + """qreg q[2]; creg c[2];
        gate foo(x, y) a, b { rx(x) a; cx a, b; rz(y / 2) b; }
        gate bar a { foo(pi, 0.5) a, a; }
        foo(3.141592653589793, 0) q[0], q[1];""")
    want = op.rx(math.pi, 0b01) * op.x(0b10).c(0b01) * op.rz(0.0, 0b10)
    assert p.tail.name() == want.name()
    with pytest.raises(qasm.QasmError):           # recursion
        qasm.parse(HEADER + "qreg q[1]; gate g a { g a; } g q[0];")
    with pytest.raises(qasm.QasmError):           # indexed qubit inside a gate body
        qasm.parse(HEADER + "qreg q[2]; gate g a { x a[0]; } g q[0];")
    with pytest.raises(qasm.QasmError):           # measure inside a gate body
        qasm.parse(HEADER + "qreg q[1]; creg c[1]; gate g a { measure a -> c; }")


@pytest.mark.parametrize("src", [
    "qreg q[2]; rx q[0];",                     # missing angle
    "qreg q[2]; rx(1) q;",                     # two bits for a one-bit rotation
    "qreg q[2]; swap q[0];",                   # one bit for swap
    "qreg q[2]; cx q[0], q[0];",               # control overlaps the target
    "qreg q[2]; foo q[0];",                    # unknown gate
    "qreg q[2]; Rx(1) q[0];",                  # gate names are x or X, not mixed case
    "qreg q[2]; x q[2];",                      # index out of range
    "qreg q[2]; creg c[1]; measure q -> c;",   # sizes differ
    "qreg q[2]; qreg q[1];",                   # duplicate register
    "qreg q[64];",                             # register too large
    "",                                        # empty source
])
def test_errors(src):
    with pytest.raises(qasm.QasmError):
        qasm.parse(HEADER + src if src else src)
