"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Plain FP64 apply is compared bit for bit (the kernels restate the
reference's arithmetic with contraction off); the stated contract is
max |dpsi| <= 1e-12.  Measured outcomes, masks and indices are compared
exactly for the same seeded draws."""
import math
import sys

import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import QReg, circuits, op
from golden_util import build_ops, load_known_answers, symbol_table

pytestmark = pytest.mark.gpu

KA = load_known_answers()
TABLE = symbol_table(KA["angle"])
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class ProductNS:
    def __getattr__(self, name):
        return getattr(op, name)

    def raw_h2(self, a, b):
        g = qvnt_amd.Gate()
        g.kind, g.act, g.a, g.b = qvnt_amd.KIND["H2"], a | b, a, b
        return qvnt_amd.from_gates([g])


def gpu_matrix(o, size):
    dim = 1 << size
    cols = []
    for col in range(dim):
        with QReg.with_state(size, col) as r:
            r.apply(o)
            cols.append(r.get_state()[:dim])
    return np.array(cols).T


@pytest.mark.parametrize("case", KA["matrix_cases"], ids=lambda c: c["ref"].split("/")[-1] + ":" + c.get("name", "Id")[:20])
def test_matrix_known_answers(case):
    say("matrix case", case["ref"], case.get("name"))
    o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    if case.get("dgr"):
        o = o.dgr()
    got = gpu_matrix(o, case["size"])
    want = np.array([[TABLE[v] for v in row] for row in case["matrix"]])
    assert np.array_equal(got, want)


def test_golden_state_vector():
    case = KA["state_cases"][0]
    o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    with QReg.with_state(case["q_num"], case["state"]) as r:
        r.apply(o)
        want = np.array([complex(re, im) for re, im in case["psi"]])
        assert np.array_equal(r.get_state(), want)
        mask = case["measure_mask"]
        assert r.measure_mask(mask).get() & ~mask == 0


@pytest.mark.parametrize("case", KA["probability_cases"], ids=lambda c: c["ref"])
def test_probabilities_known_answers(case):
    o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    with QReg.with_state(case["q_num"], case["state"]) as r:
        r.apply(o)
        p = r.get_probabilities()
    want = np.array(case["probabilities"])
    if case["tol"] == 0.0:
        assert np.array_equal(p, want)
    else:
        assert np.max(np.abs(p - want)) < case["tol"]


KINDS = ["x", "y", "z", "s", "t", "rx", "ry", "rz", "rxx", "ryy", "rzz", "phi",
         "swap", "i_swap", "sqrt_swap", "sqrt_i_swap", "h", "u3", "qft"]


def random_circuit(g, n, count):
    specs = []
    for _ in range(count):
        ctor = KINDS[int(g.integers(len(KINDS)))]
        perm = [int(v) for v in g.permutation(n)]
        if ctor in ("x", "y", "z", "s", "t"):
            k = int(g.integers(1, min(n, 3) + 1))
            args, rest = (sum(1 << q for q in perm[:k]),), perm[k:]
        elif ctor in ("rx", "ry", "rz", "phi"):
            args, rest = (float(g.uniform(0, 2 * math.pi)), 1 << perm[0]), perm[1:]
        elif ctor == "u3":
            args, rest = tuple(float(v) for v in g.uniform(0, 2 * math.pi, 3)) + (1 << perm[0],), perm[1:]
        elif ctor in ("h", "qft"):
            k = int(g.integers(1, min(n, 4) + 1))
            args, rest = (sum(1 << q for q in perm[:k]),), perm[k:]
        elif n >= 2:
            m = (1 << perm[0]) | (1 << perm[1])
            args, rest = ((float(g.uniform(0, 2 * math.pi)), m) if ctor.startswith("r") else (m,)), perm[2:]
        else:
            continue
        nc = int(g.integers(0, min(len(rest), 2) + 1))
        specs.append((ctor, args, sum(1 << q for q in rest[:nc]), bool(g.integers(0, 4) == 0)))
    return specs


def build(specs, ns):
    out = ns.id()
    for ctor, args, ctrl, dag in specs:
        o = getattr(ns, ctor)(*args)
        if dag:
            o = o.dgr()
        if ctrl:
            o = o.c(ctrl)
        out = out * o
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 9, 11, 12, 13, 16])
def test_random_circuits_bit_exact(orc, n):
    """Every gate kind, with controls and daggers, on sizes around the
    2^3 buffer minimum and the kernel tile sizes."""
    say("random circuit n =", n)
    g = np.random.default_rng(1000 + n)
    specs = random_circuit(g, n, 120)
    state = int(g.integers(0, 1 << n))
    ref = orc.Reg(n, state)
    ref.apply(build(specs, orc.op))
    with QReg.with_state(n, state) as r:
        r.apply(build(specs, op))
        got = r.get_state()
    want = ref.state()
    assert np.max(np.abs(got - want)) <= TOL
    assert np.array_equal(got, want)


def test_random_state_input_every_kind(orc):
    """Arbitrary (non-basis) input states through set_state."""
    g = np.random.default_rng(77)
    n = 10
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    specs = random_circuit(g, n, 200)
    ref = orc.Reg(n)
    ref.set_state(psi)
    ref.apply(build(specs, orc.op))
    with QReg(n) as r:
        r.set_state(psi)
        r.apply(build(specs, op))
        assert np.array_equal(r.get_state(), ref.state())


def test_plain_20q_workload(orc):
    """The 20-qubit plain workload: qft, h, 200 random gates, measure_mask(0xFF)."""
    say("plain_20q")
    state, specs, mask = circuits.plain_20q()
    ref = orc.Reg(20, state, threads=8)
    ref.apply(circuits.build(specs, orc.op))
    with QReg.with_state(20, state) as r:
        r.apply(circuits.build(specs, op))
        got = r.get_state()
        want = ref.state()
        assert np.max(np.abs(got - want)) <= TOL
        assert np.array_equal(got, want)
        u = qvnt_amd.uniform(circuits.SEED, 0, 0)
        c = r.measure_mask(mask, u=u)
        rc = ref.measure_mask(mask, u)
        assert c.get() == rc and rc & ~mask == 0
        assert np.array_equal(r.get_state(), ref.state())


@pytest.mark.parametrize("n", [0, 1, 2, 3, 6, 10, 11, 12, 15, 22, 23])
def test_measure_matches_oracle(orc, n):
    """Same state, same draw -> same outcome, same collapsed state (both
    renormalised and not), same totals."""
    say("measure n =", n)
    g = np.random.default_rng(300 + n)
    length = max(1 << n, 8)
    psi = np.zeros(length, complex)
    psi[: 1 << n] = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    psi[: 1 << n][g.random(1 << n) < 0.3] = 0
    if not psi.any():
        psi[0] = 1
    masks = [(1 << n) - 1, 0b101 & ((1 << n) - 1), 1 << max(n - 1, 0), 0]
    for k, mask in enumerate(masks):
        for renorm in (False, True):
            u = qvnt_amd.uniform(n, k, int(renorm))
            ref = orc.Reg(n)
            ref.set_state(psi)
            want_total = ref.get_absolute()
            with QReg(n) as r:
                r.set_state(psi)
                tot = r.get_absolute()
                assert tot == orc.tree_total(psi)
                assert abs(tot - want_total) <= 1e-12 * want_total
                got = r.measure_mask(mask, u=u, renorm=renorm).get()
                assert got == ref.measure_mask(mask, u, renorm=renorm)
                assert np.array_equal(r.get_state(), ref.state())


def test_probabilities_and_shots(orc):
    g = np.random.default_rng(5)
    n = 14
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    ref = orc.Reg(n)
    ref.set_state(psi)
    want = ref.get_probabilities()
    us = np.array([qvnt_amd.uniform(9, 1, k) for k in range(1024)])
    with QReg(n) as r:
        r.set_state(psi)
        p = r.get_probabilities()
        assert np.max(np.abs(p - want)) <= 1e-12 * np.max(want)
        mask = 0b1010_0110_0001
        shots = r.sample_mask(mask, us)
        assert np.array_equal(r.get_state(), psi.astype(complex))  # non-destructive
    for k in range(0, 1024, 37):
        assert shots[k] == orc.tree_sample(psi, us[k]) & mask


def test_zero_state_and_bad_gates():
    with QReg(5) as r:
        r.set_state(np.zeros(32, complex))
        with pytest.raises(qvnt_amd.QvntError) as e:
            r.measure_mask(0b11, u=0.5)
        assert e.value.code == -5
    with QReg(4) as r:
        with pytest.raises(qvnt_amd.QvntError):
            r.apply(op.x(1 << 6))       # beyond the register: the reference panics
        r.apply(op.rz(0.3, 1 << 9))     # diagonal beyond: a global phase, as in the reference
        r.apply(op.x(1).c(1 << 8))      # control beyond: never satisfied, identity
        s = r.get_state()
        assert abs(s[0] - complex(math.cos(0.15), -math.sin(0.15))) < 1e-15


def test_fp32_mode_error_bound(orc):
    """FP32 amplitudes: error against the FP64 oracle on the 20-qubit workload."""
    state, specs, mask = circuits.plain_20q()
    ref = orc.Reg(20, state, threads=8)
    ref.apply(circuits.build(specs, orc.op))
    want = ref.state()
    with QReg.with_state(20, state, prec="f32") as r:
        r.apply(circuits.build(specs, op))
        got = r.get_state().astype(np.complex128)
        assert abs(r.get_absolute() - 1.0) < 1e-5
    err = np.max(np.abs(got - want))
    infid = 1.0 - abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    say("fp32 max err", err, "infidelity", infid)
    assert err < 1e-5 and infid < 1e-5


@pytest.mark.parametrize("n", [26, 30])
def test_large_register_round_trip(n):
    """Size-independent properties at full size: a brickwork layer and its
    exact inverse (negated angles) restore the basis state; norms hold."""
    say("round trip n =", n)
    fwd, inv = [], []
    for layer in range(2):
        specs = circuits.brickwork_layer(n, layer)
        fwd += specs
    for ctor, args, ctrl in reversed(fwd):
        if ctor == "u3":
            th, ph, la, m = args
            inv += [("rz", (-ph, m), ctrl), ("ry", (-th, m), ctrl), ("rz", (-la, m), ctrl)]
        elif ctor == "rzz":
            inv.append(("rzz", (-args[0], args[1]), ctrl))
        else:
            inv.append((ctor, args, ctrl))
    start = 0x2AAAAAAA & ((1 << n) - 1)
    with QReg.with_state(n, start) as r:
        r.apply(circuits.build(fwd, op))
        assert abs(r.get_absolute() - 1.0) < 1e-12
        r.apply(circuits.build(inv, op))
        assert abs(r.get_absolute() - 1.0) < 1e-12
        draws = [qvnt_amd.uniform(4, n, k) for k in range(16)]
        outs = r.sample_mask((1 << n) - 1, draws)
        assert all(int(o) == start for o in outs)


@pytest.mark.parametrize("n,prec", [(34, "f64"), (35, "f32")])
def test_largest_register(n, prec):
    """The largest register one MI355X holds (256 GiB of amplitudes): the
    plain and fused forms on the top and bottom qubits, with outcomes fixed
    by the mathematics (x moves the basis state; h then its inverse returns
    it; a controlled x fires only where the control is set)."""
    say("largest register n=%d %s" % (n, prec))
    start = 0x155555555 & ((1 << n) - 1)
    top = 1 << (n - 1)
    with QReg.with_state(n, start, prec=prec) as r:
        r.apply(op.x(top))                       # plain: flip the top qubit
        r.apply(op.h(1) * op.h(1))               # h h = identity (exact to rounding)
        r.apply(op.x(2).c(top))                  # control set: fires
        # pair kernels on high qubits (FP32 past 32 GiB: two amplitudes per
        # access), each followed by its inverse: identity to rounding
        hi2, hi3 = top >> 1, top >> 2
        r.apply(op.ry(0.9, hi2) * op.ry(-0.9, hi2))
        r.apply(op.rxx(0.5, top | hi3) * op.rxx(-0.5, top | hi3))
        r.apply(op.swap(hi2 | hi3) * op.swap(hi2 | hi3))
        r.apply_fused(op.x(1).c(top) * op.rz(0.7, top) * op.ry(0.4, 1 << 5) * op.ry(-0.4, 1 << 5))
        want = start ^ top ^ 2 ^ 1
        assert abs(r.get_absolute() - 1.0) < (1e-12 if prec == "f64" else 1e-5)
        outs = r.sample_mask((1 << n) - 1, [qvnt_amd.uniform(5, n, k) for k in range(8)])
        assert all(int(o) == want for o in outs), [hex(int(o)) for o in outs]


def test_compare_two_registers():
    """qvnt_reg_compare against numpy on the copied-out states (FP64 vs FP32)."""
    n = 12
    g = np.random.default_rng(99)
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    specs = random_circuit(g, n, 60)
    with QReg(n) as a, QReg(n, prec="f32") as b:
        a.set_state(psi)
        b.set_state(psi.astype(np.complex64))
        a.apply(build(specs, op))
        b.apply(build(specs, op))
        err, infid, ov, na, nb = a.compare(b)
        sa, sb = a.get_state(), b.get_state().astype(np.complex128)
    assert abs(err - np.max(np.abs(sa - sb))) <= 1e-12
    assert abs(ov - np.vdot(sa, sb)) <= 1e-12
    assert abs(na - np.vdot(sa, sa).real) <= 1e-12 and abs(nb - np.vdot(sb, sb).real) <= 1e-12
    want = 1.0 - abs(np.vdot(sa, sb)) ** 2 / (np.vdot(sa, sa).real * np.vdot(sb, sb).real)
    assert abs(infid - want) <= 1e-12
    assert 0.0 <= infid < 1e-5


def test_fp32_gates_round_once_from_fp64(orc):
    """FP32 registers keep every gate parameter (the QFT ladder's phases
    among them) in FP64 and round only the output amplitude: after each
    SingleOp the FP32 state equals the FP64 oracle's result on the
    promoted FP32 input, rounded to FP32, bit for bit -- every kind, with
    controls and daggers."""
    g = np.random.default_rng(4242)
    n = 10
    psi = (g.normal(size=1 << n) + 1j * g.normal(size=1 << n)).astype(np.complex64)
    specs = random_circuit(g, n, 40) + [("qft", ((1 << n) - 1,), 0, False)]
    gates = build(specs, orc.op).gates
    dev = build(specs, op).gates()
    assert len(gates) == len(dev)
    ref = orc.Reg(n)
    with QReg(n, prec="f32") as r:
        r.set_state(psi)
        for k, (gc, gd) in enumerate(zip(gates, dev)):
            ref.set_state(psi.astype(np.complex128))
            ref.apply(orc.Op([gc]))
            want = ref.state().astype(np.complex64)
            r.apply(qvnt_amd.from_gates([gd]))
            got = r.get_state()
            assert np.array_equal(got, want), (k, qvnt_amd.KINDS[gd.kind], float(np.max(np.abs(got - want))))
            psi = got


def test_fp32_qft20_against_fp64_oracle(orc):
    """qft on all 20 qubits of an FP32 register (210 SingleOps, 190 of them
    controlled-RZ phases) against the FP64 oracle: infidelity <= 1e-5
    (reported: far below)."""
    n, state = 20, 0x5A5A5
    ref = orc.Reg(n, state, threads=8)
    ref.apply(orc.op.qft((1 << n) - 1))
    want = ref.state()
    with QReg.with_state(n, state, prec="f32") as r:
        r.apply(op.qft((1 << n) - 1))
        got = r.get_state().astype(np.complex128)
    err = float(np.max(np.abs(got - want)))
    infid = 1.0 - abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    say("fp32 qft(20): max |d| %.2e, infidelity %.2e" % (err, infid))
    assert err <= 1e-5 and infid <= 1e-5


@pytest.mark.parametrize("form", ["fused", "plain"])
def test_tfim_ring_28q_fp32_against_fp64(form):
    """The 28-qubit TFIM ring (50 Trotter steps) in an FP64 and an FP32
    register side by side: after every step the device comparison gives
    max |dpsi| <= 1e-5 and infidelity <= 1e-5; both norms stay 1."""
    n = 28
    say("tfim pair 28 q, %s" % form)
    step = circuits.build(circuits.tfim_step(n), op)
    worst = (0.0, 0.0)
    with QReg(n) as a, QReg(n, prec="f32") as b:
        for k in range(50):
            if form == "fused":
                a.apply_fused(step)
                b.apply_fused(step)
            else:
                a.apply(step)
                b.apply(step)
            err, infid, _, na, nb = a.compare(b)
            assert err <= 1e-5 and infid <= 1e-5, (k, err, infid)
            assert abs(na - 1.0) <= 1e-12 and abs(nb - 1.0) <= 1e-5
            worst = (max(worst[0], err), max(worst[1], infid))
    say("  worst step: max |dpsi| %.2e, infidelity %.2e" % worst)


def test_u2_controlled_on_its_b_qubit_is_refused_by_apply():
    """op::matrix2(a, b, m).c(b) is a valid operator (acts_on = a only,
    atomic/u2.rs:70-72) that the engine refuses to apply, leaving the state
    untouched."""
    from qvnt_amd._lib import E_CTRL_OVERLAP
    m = np.eye(4, dtype=np.complex128)[[1, 0, 3, 2]]
    cu = op.matrix2(0b001, 0b010, m).c(0b010)
    with QReg.with_state(3, 0b011) as r:
        before = r.get_state()
        with pytest.raises(qvnt_amd.QvntError) as e:
            r.apply(cu)
        assert e.value.code == E_CTRL_OVERLAP
        assert np.array_equal(r.get_state(), before)


LINE_CASES = [  # gates with qubits or controls inside one 128-byte line (FP64: qubits 0-2, FP32: 0-3)
    ("x", (1 << 12,), 1 << 0), ("x", (1 << 3,), 1 << 2), ("x", (1 << 1,), 1 << 0), ("x", (1 << 9,), (1 << 1) | (1 << 11)),
    ("rz", (0.7, 1 << 1), 1 << 0), ("rz", (0.7, 1 << 10), 1 << 3), ("ry", (0.4, 1 << 2), 1 << 3),
    ("t", (1 << 0,), 0), ("s", (1 << 2,), 1 << 5), ("z", (1 << 3,), 0), ("phi", (0.3, 1 << 1), 0),
    ("swap", ((1 << 0) | (1 << 9),), 0), ("i_swap", ((1 << 1) | (1 << 2),), 1 << 8),
    ("sqrt_swap", ((1 << 3) | (1 << 4),), 1 << 0), ("sqrt_i_swap", ((1 << 2) | (1 << 12),), 0),
    ("h", (0b11,), 0), ("h", (0b110000,), 1 << 1), ("h", ((1 << 2) | (1 << 9),), 1 << 0), ("h", (0b1100,), 0),
    ("rxx", (0.9, 0b101), 1 << 1), ("ryy", (0.9, (1 << 3) | (1 << 7)), 1 << 2), ("rzz", (0.2, 0b11), 1 << 2),
    ("y", (1 << 4,), 1 << 0), ("u3", (0.1, 0.2, 0.3, 1 << 5), 1 << 1),
]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gates_inside_a_cache_line(orc, prec):
    """Controls and one-qubit phase bits inside a 128-byte line are tested per
    amplitude rather than skipped (so every line is written whole): each
    SingleOp on a random state equals the oracle -- bit for bit in FP64, and
    in FP32 the oracle's FP64 result on the widened input, rounded once."""
    g = np.random.default_rng(5150)
    n = 13
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    dt = np.complex128 if prec == "f64" else np.complex64
    psi = psi.astype(dt)
    ref = orc.Reg(n)
    with QReg(n, prec=prec) as r:
        r.set_state(psi)
        for ctor, args, ctrl in LINE_CASES:
            o_dev = circuits.build([(ctor, args, ctrl)], op)
            o_ref = circuits.build([(ctor, args, ctrl)], orc.op)
            for gd, gc in zip(o_dev.gates(), o_ref.gates):
                ref.set_state(psi.astype(np.complex128))
                ref.apply(orc.Op([gc]))
                want = ref.state().astype(dt)
                r.apply(qvnt_amd.from_gates([gd]))
                got = r.get_state()
                assert np.array_equal(got, want), (ctor, args, ctrl, float(np.max(np.abs(got - want))))
                psi = got


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_h2_with_one_bit_in_the_lanes(orc, prec):
    """H2 with one bit in qubits 0-5 and the other above (the half-lane
    kernel), in both orientations of a_mask / b_mask, with and without a
    control inside a line: bit for bit against the oracle (FP32: the FP64
    result on the widened input, rounded once)."""
    g = np.random.default_rng(6160)
    n = 12
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    dt = np.complex128 if prec == "f64" else np.complex64
    psi = (psi / np.linalg.norm(psi)).astype(dt)
    ref = orc.Reg(n)
    with QReg(n, prec=prec) as r:
        r.set_state(psi)
        for a, b, ctrl in [(1 << 2, 1 << 9, 0), (1 << 9, 1 << 2, 0), (1 << 1, 1 << 7, 1 << 0), (1 << 11, 1 << 5, 1 << 2),
                           (1 << 0, 1 << 6, 1 << 8)]:
            gd, gc = qvnt_amd.Gate(), orc.Gate()
            for x in (gd, gc):
                x.kind, x.act, x.a, x.b, x.ctrl = qvnt_amd.KIND["H2"], a | b, a, b, ctrl
            ref.set_state(psi.astype(np.complex128))
            ref.apply(orc.Op([gc]))
            want = ref.state().astype(dt)
            r.apply(qvnt_amd.from_gates([gd]))
            got = r.get_state()
            assert np.array_equal(got, want), (a, b, ctrl, float(np.max(np.abs(got - want))))
            psi = got
