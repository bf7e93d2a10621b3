"""QASM programs through the front end and the device driver loop
(Sym::finish, qasm/sym.rs:40-74): textbook algorithms with certain
outcomes, and the final state, classical value and draws against the
oracle's driver on the same program."""
import numpy as np
import pytest

import oracle
import qvnt_amd
from qvnt_amd import Sym, qasm

pytestmark = pytest.mark.gpu

HEADER = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'

DEUTSCH = HEADER + """
qreg q[2]; creg c[1];
x q[1]; h q;
cx q[0], q[1];            // balanced oracle f(x) = x
h q[0];
measure q[0] -> c[0];
"""

BERNSTEIN_VAZIRANI = HEADER + """
qreg q[5]; qreg anc[1]; creg c[5];
x anc; h q; h anc;
cx q[0], anc[0]; cx q[2], anc[0]; cx q[3], anc[0];    // hidden string 01101
h q;
measure q -> c;
"""

TELEPORT = HEADER + """
qreg q[3]; creg m0[1]; creg m1[1];
gate prep(t) a { ry(t) a; rz(t / 3) a; }
prep(0.9) q[0];
h q[1]; cx q[1], q[2];
cx q[0], q[1]; h q[0];
measure q[0] -> m0[0];
measure q[1] -> m1[0];
if (m1 == 1) x q[2];
if (m0 == 1) z q[2];
"""

RESET_LOOP = HEADER + """
qreg q[4]; creg c[4];
h q; cx q[0], q[3];
measure q[0] -> c[0];
reset q[3];
crz(pi / 3) q[1], q[2];
measure q[1] -> c[1];
reset q;
x q[2];
measure q -> c;
"""


def to_oracle(mo):
    return oracle.Op([oracle.Gate.from_buffer_copy(g) for g in mo.gates()])


def run_both(src, draws, measure="set"):
    prog = qasm.parse(src, measure=measure)
    s = Sym(prog)
    s.finish(draws=draws)
    ref = oracle.Reg(prog.q_num)
    steps = [(to_oracle(o), sep, a, b) for o, sep, a, b in prog.steps]
    c, used = ref.run_program(steps, to_oracle(prog.tail), prog.c_num, measure_xor=(measure == "xor"), draws=draws)
    return s, ref, c, used


def test_deutsch_balanced():
    s, ref, c, _ = run_both(DEUTSCH, [0.3])
    assert s.get_class().get() == 1 == c


def test_bernstein_vazirani():
    for u in (0.01, 0.5, 0.99):
        s, ref, c, _ = run_both(BERNSTEIN_VAZIRANI, [u])
        assert s.get_class().get() == 0b01101 == c


@pytest.mark.parametrize("k", range(4))
def test_teleport(k):
    """Whatever the two outcomes, q[2] ends in the prepared state."""
    draws = [qvnt_amd.uniform(17, 3, 2 * k), qvnt_amd.uniform(17, 3, 2 * k + 1)]
    s, ref, c, used = run_both(TELEPORT, draws)
    assert s.get_class().get() == c and used == 2
    psi = s.q_reg.get_state()
    assert np.max(np.abs(psi - ref.state())) <= 1e-12
    m = s.get_class().get()
    p = np.abs(psi.reshape(2, 2, 2)) ** 2                # [q2][q1][q0]
    p2 = p.sum(axis=(1, 2))
    p2 = p2 / p2.sum()                                   # measure_mask does not renormalise (quant.rs:490-501)
    want1 = np.sin(0.9 / 2) ** 2                         # ry(0.9) from |0>; rz keeps |amplitudes|
    assert abs(p2[1] - want1) < 1e-12, (m, p2)


@pytest.mark.parametrize("measure", ["set", "xor"])
def test_reset_and_measure_modes(measure):
    draws = [qvnt_amd.uniform(5, 1, k) for k in range(3)]
    s, ref, c, used = run_both(RESET_LOOP, draws, measure)
    assert s.get_class().get() == c and used == 3
    assert np.max(np.abs(s.q_reg.get_state() - ref.state())) <= 1e-12
    if measure == "set":       # after `reset q; x q[2];` the last measurement is certain
        assert s.get_class().get() == 0b0100
