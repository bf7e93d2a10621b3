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


# Programs written for these tests, with the features the reference's
# example programs use (macros, multi-controlled gates, if on a classical
# register, reset, u1/u3): each has an outcome the mathematics fixes.

ADDER = HEADER + """
// 3-bit ripple-carry addition (Cuccaro et al.): b := a + b, carry out in z
gate maj x, y, w { cx w, y; cx w, x; ccx x, y, w; }
gate uma x, y, w { ccx x, y, w; cx w, x; cx x, y; }
qreg c0[1]; qreg a[3]; qreg b[3]; qreg z[1]; creg s[4];
x a[0]; x a[2];            // a = 101 = 5
x b[1]; x b[2];            // b = 110 = 6
maj c0[0], b[0], a[0];
maj a[0], b[1], a[1];
maj a[1], b[2], a[2];
cx a[2], z[0];
uma a[1], b[2], a[2];
uma a[0], b[1], a[1];
uma c0[0], b[0], a[0];
measure b[0] -> s[0]; measure b[1] -> s[1]; measure b[2] -> s[2]; measure z[0] -> s[3];
"""

# the Fourier state of x = 0b101 on 3 qubits (qubit j carries the phase
# 2 pi x 2^j / 8), read out by the semiclassical inverse QFT: measure the
# qubit with phase pi x first, then correct each next qubit by the bits
# already read (if on the whole classical register) before its h
SEMICLASSICAL_IQFT = HEADER + """
qreg q[3]; creg c[3];
h q;
u1(2 * pi * 5 / 8) q[0]; u1(2 * pi * 10 / 8) q[1]; u1(2 * pi * 20 / 8) q[2];
h q[2];
measure q[2] -> c[0];
if (c == 1) u1(-pi / 2) q[1];
h q[1];
measure q[1] -> c[1];
if (c == 1) u1(-pi / 4) q[0];
if (c == 2) u1(-pi / 2) q[0];
if (c == 3) u1(-3 * pi / 4) q[0];
h q[0];
measure q[0] -> c[2];
"""

W_STATE = HEADER + """
// (|001> + |010> + |100>) / sqrt(3)
gate cry(t) p, r { ry(t / 2) r; cx p, r; ry(-t / 2) r; cx p, r; }
qreg q[3]; creg c[3];
ry(2 * atan2(sqrt(2), 1)) q[0];     // P(q0 = 1) = 2/3
cry(pi / 2) q[0], q[1];             // q0 = 1: q1 in an equal superposition
cx q[1], q[0];                      // |11> -> |10>
x q[0]; x q[1]; ccx q[0], q[1], q[2]; x q[0]; x q[1];   // |00> -> |100>
measure q -> c;
"""

# iterative phase estimation of the phase 2 pi * 3/8 of u1 on |1>, the
# control qubit measured between rounds and flipped back to |0> by an if
# on the bits read (qvnt's reset projects onto |0> and renormalises,
# quant.rs:207-229, which after a |1> outcome would clear the target too)
IPEA = HEADER + """
gate cp(t) ctl, tgt { u1(t / 2) ctl; cx ctl, tgt; u1(-t / 2) tgt; cx ctl, tgt; u1(t / 2) tgt; }
qreg q[2]; creg c[3];
x q[1];
h q[0]; cp(4 * 2 * pi * 3 / 8) q[0], q[1]; h q[0];
measure q[0] -> c[0];
if (c == 1) x q[0];
h q[0]; cp(2 * 2 * pi * 3 / 8) q[0], q[1];
if (c == 1) u1(-pi / 2) q[0];
h q[0];
measure q[0] -> c[1];
if (c == 2) x q[0];
if (c == 3) x q[0];
h q[0]; cp(2 * pi * 3 / 8) q[0], q[1];
if (c == 1) u1(-pi / 4) q[0];
if (c == 2) u1(-pi / 2) q[0];
if (c == 3) u1(-3 * pi / 4) q[0];
h q[0];
measure q[0] -> c[2];
"""


@pytest.mark.parametrize("k", range(3))
def test_adder(k):
    draws = [qvnt_amd.uniform(31, k, j) for j in range(4)]
    s, ref, c, used = run_both(ADDER, draws)
    assert s.get_class().get() == c == 11 and used == 4      # 5 + 6
    assert np.max(np.abs(s.q_reg.get_state() - ref.state())) <= 1e-12


@pytest.mark.parametrize("k", range(3))
def test_semiclassical_inverse_qft(k):
    draws = [qvnt_amd.uniform(37, k, j) for j in range(3)]
    s, ref, c, used = run_both(SEMICLASSICAL_IQFT, draws)
    assert s.get_class().get() == c == 0b101 and used == 3
    assert np.max(np.abs(s.q_reg.get_state() - ref.state())) <= 1e-12


def test_w_state():
    seen = set()
    for k in range(12):
        draws = [qvnt_amd.uniform(41, k, 0)]
        s, ref, c, used = run_both(W_STATE, draws)
        assert s.get_class().get() == c and c in (0b001, 0b010, 0b100)
        seen.add(c)
    assert len(seen) == 3
    p = Sym(qasm.parse(W_STATE.replace("measure q -> c;", ""))).finish().get_probabilities()
    assert np.max(np.abs(p - np.array([0, 1, 1, 0, 1, 0, 0, 0]) / 3.0)) <= 1e-12


@pytest.mark.parametrize("k", range(3))
def test_iterative_phase_estimation(k):
    draws = [qvnt_amd.uniform(43, k, j) for j in range(3)]
    s, ref, c, used = run_both(IPEA, draws)
    assert s.get_class().get() == c == 0b011 and used == 3    # 3/8 = 0.011 in binary, last bit read first
    assert np.max(np.abs(s.q_reg.get_state() - ref.state())) <= 1e-12
