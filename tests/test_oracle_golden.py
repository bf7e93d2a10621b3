"""The CPU oracle against the reference's own known answers (transcribed in
tests/golden/reference_known_answers.json from qvnt's unit tests and
doctests).  These pin the oracle before it is trusted as the checker."""
import numpy as np
import pytest

from golden_util import build_ops, load_known_answers, symbol_table

KA = load_known_answers()
ANGLE = KA["angle"]
TABLE = symbol_table(ANGLE)


class OracleNS:
    """op:: of the oracle, plus the raw H2 atomic the h2.rs test builds."""

    def __init__(self, orc):
        self.orc = orc

    def __getattr__(self, name):
        return getattr(self.orc.op, name)

    def id(self):
        return self.orc.op.id()

    def matrix1(self, a, m):
        try:
            return self.orc.op.matrix1(a, m)
        except ValueError:
            return None

    def matrix2(self, a, b, m):
        try:
            return self.orc.op.matrix2(a, b, m)
        except ValueError:
            return None

    def raw_h2(self, a, b):
        g = self.orc.Gate()
        g.kind, g.act, g.a, g.b = self.orc.KIND["H2"], a | b, a, b
        return self.orc.Op([g])


def oracle_matrix(orc, o, size):
    """Applicable::matrix (applicable.rs:17-46): apply to every basis vector."""
    dim = 1 << size
    cols = []
    for col in range(dim):
        reg = orc.Reg(size, col)
        reg.apply(o)
        cols.append(reg.state()[:dim])
    return np.array(cols).T


@pytest.mark.parametrize("case", KA["matrix_cases"], ids=lambda c: c["ref"].split("/")[-1] + ":" + c.get("name", "Id"))
def test_matrix_known_answers(orc, case):
    o = build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)
    assert o is not None
    if case.get("dgr"):
        o = o.dgr()
    got = oracle_matrix(orc, o, case["size"])
    want = np.array([[TABLE[v] for v in row] for row in case["matrix"]])
    # the reference asserts exact equality (assert_eq! on f64)
    assert np.array_equal(got, want), (got, want)


@pytest.mark.parametrize("case", KA["invalid_cases"], ids=lambda c: c["ref"])
def test_invalid_constructors(orc, case):
    try:
        o = build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)
    except ValueError:
        o = None
    assert o is None


@pytest.mark.parametrize("case", KA["ctrl_cases"], ids=lambda c: c["ref"] + str(c["c"]))
def test_control_overlap(orc, case):
    o = build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)
    assert (o.c(case["c"]) is not None) == case["valid"]


def test_queue_length_drops_id(orc):
    case = KA["len_cases"][0]
    assert len(build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)) == case["len"]


def test_golden_state_vector(orc):
    case = KA["state_cases"][0]
    o = build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)
    reg = orc.Reg(case["q_num"], case["state"])
    reg.apply(o)
    want = np.array([complex(re, im) for re, im in case["psi"]])
    assert np.array_equal(reg.state(), want)
    mask = case["measure_mask"]
    for k in range(16):
        r2 = orc.Reg(case["q_num"], case["state"])
        r2.apply(o)
        assert r2.measure_mask(mask, orc.uniform(1, 2, k)) & ~mask == 0


@pytest.mark.parametrize("case", KA["probability_cases"], ids=lambda c: c["ref"])
def test_probabilities(orc, case):
    o = build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)
    reg = orc.Reg(case["q_num"], case["state"])
    reg.apply(o)
    p = reg.get_probabilities()
    want = np.array(case["probabilities"])
    if case["tol"] == 0.0:
        assert np.array_equal(p, want)
    else:
        assert np.max(np.abs(p - want)) < case["tol"]


@pytest.mark.parametrize("case", KA["measure_cases"], ids=lambda c: "q%d" % c["q_num"])
def test_bench_measure_containment(orc, case):
    o = build_ops(case["ops"], OracleNS(orc), TABLE, ANGLE)
    reg = orc.Reg(case["q_num"], 0, threads=2)
    reg.apply(o)
    assert reg.measure_mask(case["mask"], orc.uniform(3, 4, case["q_num"])) & ~case["mask"] == 0
