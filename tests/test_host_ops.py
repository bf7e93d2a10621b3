"""The product's host-side operator algebra (qvnt_op_* in libqvnt_hip.so)
against the oracle's restatement of the reference builders, and the Debug
names against the reference's own strings.  No GPU needed."""
import math

import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import circuits, op
from golden_util import build_ops, load_known_answers, symbol_table

KA = load_known_answers()
TABLE = symbol_table(KA["angle"])


class ProductNS:
    def __getattr__(self, name):
        return getattr(op, name)

    def raw_h2(self, a, b):
        g = qvnt_amd.Gate()
        g.kind, g.act, g.a, g.b = qvnt_amd.KIND["H2"], a | b, a, b
        return qvnt_amd.from_gates([g])


def gate_bytes(gates):
    return [bytes(g) for g in gates]


def same_gates(prod, orac):
    assert len(prod) == len(orac)
    assert gate_bytes(prod.gates()) == gate_bytes(orac.gates)


CTORS_MASK = ["x", "y", "z", "s", "t", "h", "qft", "qft_swapped"]
CTORS_2BIT = ["swap", "sqrt_swap", "i_swap", "sqrt_i_swap"]
CTORS_ROT1 = ["rx", "ry", "rz", "phi", "u1"]
CTORS_ROT2 = ["rxx", "ryy", "rzz"]


@pytest.mark.parametrize("seed", range(8))
def test_builders_match_oracle(orc, seed):
    g = np.random.default_rng(seed)
    for _ in range(50):
        n = 12
        mask = int(g.integers(0, 1 << n))
        bits = [int(v) for v in g.choice(n, 2, replace=False)]
        th = [float(v) for v in g.uniform(-10, 10, 3)]
        for c in CTORS_MASK:
            same_gates(getattr(op, c)(mask), getattr(orc.op, c)(mask))
        for c in CTORS_2BIT:
            same_gates(getattr(op, c)((1 << bits[0]) | (1 << bits[1])), getattr(orc.op, c)((1 << bits[0]) | (1 << bits[1])))
        for c in CTORS_ROT1:
            same_gates(getattr(op, c)(th[0], 1 << bits[0]), getattr(orc.op, c)(th[0], 1 << bits[0]))
        for c in CTORS_ROT2:
            m = (1 << bits[0]) | (1 << bits[1])
            same_gates(getattr(op, c)(th[0], m), getattr(orc.op, c)(th[0], m))
        same_gates(op.u2(th[0], th[1], 1 << bits[1]), orc.op.u2(th[0], th[1], 1 << bits[1]))
        same_gates(op.u3(th[0], th[1], th[2], 1 << bits[1]), orc.op.u3(th[0], th[1], th[2], 1 << bits[1]))


@pytest.mark.parametrize("seed", range(4))
def test_queue_algebra_matches_oracle(orc, seed):
    g = np.random.default_rng(50 + seed)
    for _ in range(30):
        a, b, c = (int(v) for v in g.choice(10, 3, replace=False))
        th = float(g.uniform(0, 6))
        P = op.h(1 << a) * op.ryy(th, (1 << a) | (1 << b)) * op.t(1 << b) * op.i_swap((1 << a) | (1 << b))
        O = orc.op.h(1 << a) * orc.op.ryy(th, (1 << a) | (1 << b)) * orc.op.t(1 << b) * orc.op.i_swap((1 << a) | (1 << b))
        same_gates(P.dgr(), O.dgr())
        same_gates(P.c(1 << c), O.c(1 << c))
        assert (P.c(1 << a) is None) and (O.c(1 << a) is None)
        assert P.act_on() == O.act_on()
        assert P.dgr().dgr() == P
        PP = P.clone()
        PP *= P
        assert len(PP) == 2 * len(P)


@pytest.mark.parametrize("case", [c for c in KA["matrix_cases"] if "name" in c], ids=lambda c: c["name"][:24])
def test_names_of_known_answers(case):
    o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    if case.get("dgr"):
        o = o.dgr()
    assert o.name() == "[" + case["name"] + "]"


@pytest.mark.parametrize("case", KA["name_cases"], ids=lambda c: c["name"])
def test_debug_names(case):
    o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    if "c" in case:
        o = o.c(case["c"])
    assert o.name() == case["name"]


@pytest.mark.parametrize("case", KA["invalid_cases"], ids=lambda c: c["ref"])
def test_invalid(case):
    try:
        o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    except ValueError:
        o = None
    assert o is None


@pytest.mark.parametrize("case", KA["ctrl_cases"], ids=lambda c: c["ref"] + str(c["c"]))
def test_ctrl_cases(case):
    o = build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])
    assert (o.c(case["c"]) is not None) == case["valid"]


def test_len_and_ends_with():
    case = KA["len_cases"][0]
    assert len(build_ops(case["ops"], ProductNS(), TABLE, KA["angle"])) == case["len"]
    # multi/mod.rs:210-217
    a = op.x(0x010101) * op.y(0x101010) * op.z(0x011011)
    b = op.y(0x101010) * op.z(0x011011)
    assert b.ends_with(a)
    assert len(op.x(0)) == 1 and op.x(0).name() == "[X0]"
    assert len(op.h(0)) == 0 and len(op.qft(0)) == 0
    assert op.qft(0b100) == op.h(0b100)


def test_rust_float_display():
    assert op.rx(2.0, 1).name() == "[RX1(2)]"
    assert op.rx(1.23456, 1).name() == "[RX1(1.23456)]"
    assert op.rz(0.5, 1).dgr().name().startswith("[RZ1(-")
    assert op.rx(-1.5, 4).name() == "[RX4(-1.5)]"


def test_uniform_matches_oracle(orc):
    for s in range(20):
        for k in range(50):
            assert qvnt_amd.uniform(s, 7, k) == orc.uniform(s, 7, k)


def test_workload_op_counts():
    state, specs, mask = circuits.plain_20q()
    o = circuits.build(specs, op)
    assert len(o) == 420 == circuits.single_ops(specs)
    layer = circuits.brickwork_layer(30, 0) + circuits.brickwork_layer(30, 1)
    assert circuits.single_ops(layer) == 106 + 105
    assert len(circuits.build(layer, op)) == 211
    assert circuits.single_ops(circuits.tfim_step(34)) == 68


def test_u2_controlled_on_its_b_qubit_is_accepted_by_c_and_refused_by_apply():
    """The reference's U2 reports only a_mask in acts_on (atomic/u2.rs:70-72),
    so .c(b_mask) succeeds on the operator (single/mod.rs:109-118), as here.
    Applying it would make the reference's gather read a control it also
    targets; this engine refuses it at apply time (QVNT_E_CTRL_OVERLAP,
    before any device work), a deliberate deviation."""
    import numpy as np
    from qvnt_amd._lib import E_CTRL_OVERLAP
    m = np.eye(4, dtype=np.complex128)[[1, 0, 3, 2]]      # a unitary 4x4 (swaps within pairs)
    u = op.matrix2(0b001, 0b010, m)
    assert u.act_on() == 0b001
    cu = u.c(0b010)
    assert cu is not None and cu.act_on() == 0b011
    # the compiler of the fused form refuses it too (no device needed)
    with pytest.raises(qvnt_amd.QvntError) as e:
        qvnt_amd.fused_plan_stats(cu, 4)
    assert e.value.code == E_CTRL_OVERLAP
