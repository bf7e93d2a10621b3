"""The oracle's reset, register algebra, readback and QASM driver loop
(quant.rs:156-166, 202-229, 245-431, 513-594; qasm/sym.rs:40-74) pinned by
the reference's own tests and by closed forms, plus the host-side CReg /
VReg utilities of the package (class.rs, virtl.rs)."""
import numpy as np
import pytest

import oracle
from qvnt_amd import CReg, VReg

op = oracle.op


def rand_state(g, n, norm=True):
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    return psi / np.linalg.norm(psi) if norm else psi


def reg_with(psi):
    n = int(np.log2(len(psi)))
    r = oracle.Reg(n)
    full = np.zeros(len(r), np.complex128)
    full[: len(psi)] = psi
    r.set_state(full)
    return r


def test_tensor_reference_test():
    """quant.rs:679-711: (h on |01>) x (h on |1>) probabilities."""
    r1, r2 = oracle.Reg(2, 0b01), oracle.Reg(1, 0b1)
    r1.apply(op.h(0b01))
    r2.apply(op.h(0b01))
    p = (r1 * r2).get_probabilities()
    assert np.allclose(p, [0.25, 0.25, 0, 0, 0.25, 0.25, 0, 0], atol=1e-9)


@pytest.mark.parametrize("na,nb", [(0, 3), (1, 1), (2, 3), (4, 2), (5, 5)])
def test_tensor_is_kron(na, nb):
    """self on the low qubits: psi = kron(b, a)."""
    g = np.random.default_rng(na * 10 + nb)
    a, b = rand_state(g, na), rand_state(g, nb)
    t = reg_with(a) * reg_with(b)
    got = t.state()
    want = np.kron(b, a)
    assert np.max(np.abs(got[: len(want)] - want)) < 1e-15
    assert np.all(got[len(want):] == 0)
    assert t.q_num == na + nb


def test_virtual_regs_reference_test():
    """virtl.rs tests: index, virtual_regs (reg = new(0) * new(3) * new(2))."""
    v = VReg(8)
    assert v[3] == 0b00001000
    assert v[lambda i: i % 2 == 0] == 0b01010101
    assert v[lambda i: i % 2 != 0] == 0b10101010
    assert v[:] == 0b11111111
    assert v[lambda _: True] == 0b11111111
    assert v[[0, 7]] == 0b10000001
    reg = oracle.Reg(0) * oracle.Reg(3) * oracle.Reg(2)
    assert reg.q_num == 5
    r, m = VReg.from_mask(0b11111), VReg.from_mask(0b01101)
    x, y = VReg.from_mask(0b00111), VReg.from_mask(0b11000)
    assert (r[0], r[1], r[2]) == (0b00001, 0b00010, 0b00100)
    assert (m[0], m[1], m[2], m[:]) == (0b00001, 0b00100, 0b01000, 0b01101)
    assert (x[:], y[:], y[0], y[1]) == (0b00111, 0b11000, 0b01000, 0b10000)


def test_creg_reference_examples():
    """class.rs doc examples: tensor of 4-bit registers, Debug string."""
    assert CReg.with_state(4, 11) * CReg.with_state(4, 7) == CReg.with_state(8, 123)
    assert repr(CReg.with_state(8, 123)) == "(01111011)"
    c = CReg(6, 0b101101)
    assert c.get_by_mask(0b110100) == 0b101   # bits 2, 4, 5 of the value
    c.set(False, 0b1)
    c.xor(True, 0b10)
    assert c.get() == 0b101110


def test_combine_closed_forms():
    g = np.random.default_rng(4)
    for n in (3, 4, 6):
        a, b = rand_state(g, n), rand_state(g, n)
        c = reg_with(a).combine(reg_with(b))
        assert np.array_equal(c.state(), np.concatenate([a, b]))
        th = 0.7
        m = np.array([[np.cos(th), 1j * np.sin(th)], [1j * np.sin(th), np.cos(th)]])
        u = reg_with(a).combine_with_unitary(reg_with(b), m)
        want = np.concatenate([m[0, 0] * a + m[0, 1] * b, m[1, 0] * a + m[1, 1] * b])
        assert np.max(np.abs(u.state() - want)) < 1e-15
    assert oracle.Reg(2).combine(oracle.Reg(2)) is None      # the reference panics: 8-element halves do not fit
    assert oracle.Reg(3).combine(oracle.Reg(4)) is None      # None: sizes differ


def test_linear_composition_and_set_num():
    g = np.random.default_rng(5)
    a, b = rand_state(g, 5), rand_state(g, 5)
    r = reg_with(a)
    c0, c1 = 0.3 - 0.2j, -0.1 + 0.9j
    r.linear_composition(reg_with(b), c0, c1)
    assert np.max(np.abs(r.state() - (a * c0 + b * c1))) < 1e-15
    r = reg_with(a)
    r.set_num(3)                                   # truncated, not renormalised
    assert np.array_equal(r.state(), a[:8]) and r.q_num == 3
    r.set_num(6)                                   # zero-extended
    assert np.array_equal(r.state()[:8], a[:8]) and np.all(r.state()[8:] == 0) and len(r) == 64


def test_reset_by_mask():
    g = np.random.default_rng(6)
    a = rand_state(g, 4)
    r = reg_with(a)
    r.reset_by_mask(0b0101)                        # project bits 0 and 2 onto 0, renormalise
    keep = np.array([(i & 0b0101) == 0 for i in range(16)])
    want = np.where(keep, a, 0)
    want /= np.linalg.norm(want)
    assert np.max(np.abs(r.state() - want)) < 1e-15
    r = reg_with(a)
    r.reset_by_mask(0b1111)                        # whole register: reset(0)
    assert np.array_equal(r.state(), np.eye(16)[0])
    r = oracle.Reg(3, 0b001)
    r.reset_by_mask(0b001)                         # the only amplitude is removed: norm 0 -> reset(0)
    assert np.array_equal(r.state(), np.eye(8)[0])
    r = oracle.Reg(3, 0b110)
    r.set_state(r.state() * 1.5)                   # norm above 1 is left alone (1 - norm <= 1e-9)
    r.reset_by_mask(0b001)
    assert r.state()[6] == 1.5


def test_polar():
    g = np.random.default_rng(7)
    a = rand_state(g, 6)
    p = reg_with(a).get_polar()
    assert np.allclose(p[:, 0], np.abs(a), rtol=1e-15, atol=0)
    assert np.allclose(p[:, 1], np.angle(a), rtol=1e-15, atol=1e-300)
    assert np.max(np.abs(p[:, 0] * np.exp(1j * p[:, 1]) - a)) < 1e-15


def test_sample_all_histogram():
    """quant.rs:713-724: len 256, sum 2048; and the counts follow |psi|^2."""
    r = oracle.Reg(8, 123)
    r.apply(op.h(255))
    for k in range(10):
        h = r.sample_all(2048, 1, k)
        assert len(h) == 256 and int(h.sum()) == 2048
    g = np.random.default_rng(8)
    a = rand_state(g, 4)
    h = reg_with(a).sample_all(200000, 2, 0).astype(np.float64)
    p = np.abs(a) ** 2
    chi2 = np.sum((h - 200000 * p) ** 2 / (200000 * p))
    assert chi2 < 50.0                             # 15 degrees of freedom
    h2 = reg_with(a).sample_all(200000, 2, 0)
    assert np.array_equal(h, h2)                   # seeded: reproducible


def test_program_driver_deterministic():
    """Sym::finish on basis states: every measurement is certain."""
    # x q0; measure q0 -> c1; if (c[0b10] == 1) x q2; reset q0 (projects |101> away -> |000>)
    steps = [(op.x(1), oracle.SEP_MEASURE, 0b1, 0b10), (op.x(4), oracle.SEP_IF, 0b10, 1),
             (op.id(), oracle.SEP_RESET, 0b1, 0)]
    r = oracle.Reg(3)
    c, used = r.run_program(steps, op.id(), 2, draws=[0.25])
    assert (c, used) == (0b10, 1)
    assert np.array_equal(r.state(), np.eye(8)[0])
    # Set vs Xor: measure q1,q2 (=1,1) into c0,c2 twice
    steps = [(op.x(0b110), oracle.SEP_MEASURE, 0b110, 0b101), (op.id(), oracle.SEP_MEASURE, 0b110, 0b101)]
    c_set, _ = oracle.Reg(3).run_program(steps, op.id(), 3, draws=[0.5, 0.5])
    c_xor, _ = oracle.Reg(3).run_program(steps, op.id(), 3, measure_xor=True, draws=[0.5, 0.5])
    assert c_set == 0b101 and c_xor == 0
    # a measurement of no qubit takes no draw; the tail runs last
    steps = [(op.id(), oracle.SEP_MEASURE, 0b1000, 0b1)]
    r = oracle.Reg(3)
    c, used = r.run_program(steps, op.x(0b100), 1, c_value=1, draws=[])
    assert (c, used) == (0, 0) and r.state()[4] == 1
    # if: packed bits of c[0b1010] = (c3, c1)
    steps = [(op.x(1), oracle.SEP_IF, 0b1010, 0b10)]
    r = oracle.Reg(2)
    r.run_program(steps, op.id(), 4, c_value=0b1000)
    assert r.state()[1] == 1
