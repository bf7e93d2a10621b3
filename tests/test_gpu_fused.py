"""The fused form (qvnt_reg_apply_fused) against the CPU oracle.

The fused form reorders gates only past gates on disjoint qubits and
multiplies one-qubit runs into 2x2 matrices, so it agrees with the
reference to rounding: the bar is max |dpsi| <= 1e-12 in FP64 and
infidelity <= 1e-5 in FP32."""
import sys

import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import QReg, circuits, op
from test_gpu_parity import build, random_circuit

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,tq", [(6, 6), (7, 6), (8, 8), (9, 6), (10, 8), (11, 10), (12, 12), (13, 8),
                                  (14, 12), (16, 10), (17, 12), (20, 12)])
def test_random_circuits(orc, n, tq):
    say("fused random n=%d tq=%d" % (n, tq))
    g = np.random.default_rng(5000 + 31 * n + tq)
    specs = random_circuit(g, n, 200)
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    ref = orc.Reg(n, threads=8)
    ref.set_state(psi)
    ref.apply(build(specs, orc.op))
    with QReg(n) as r:
        r.set_state(psi)
        passes = r.apply_fused(build(specs, op), max_qubits=tq)
        got = r.get_state()
    err = np.max(np.abs(got - ref.state()))
    say("  passes", passes, "of", len(build(specs, op)), "gates; max |d| = %.2e" % err)
    assert err <= TOL


@pytest.mark.parametrize("n,tq,dense,prec", [(10, 10, 2, "f64"), (13, 12, 2, "f64"), (16, 11, 3, "f64"),
                                             (14, 12, 2, "f32"), (13, 13, 0, "f32"), (17, 13, 3, "f32")])
def test_dense_phases(orc, n, tq, dense, prec):
    """Phases folded into dense 16x16 unitaries on the matrix cores."""
    say("dense n=%d tq=%d min_ops=%d %s" % (n, tq, dense, prec))
    g = np.random.default_rng(7000 + n)
    specs = random_circuit(g, n, 300) + [s + (False,) for l in range(4) for s in circuits.brickwork_layer(n, l)]
    ref = orc.Reg(n, threads=8)
    ref.apply(build(specs, orc.op))
    with QReg(n, prec=prec) as r:
        r.set_dense(dense)
        r.apply_fused(build(specs, op), max_qubits=tq)
        got = r.get_state().astype(np.complex128)
    want = ref.state()
    err = np.max(np.abs(got - want))
    infid = 1.0 - abs(np.vdot(got, want)) ** 2
    say("  max |d| = %.2e, infidelity %.2e" % (err, infid))
    if prec == "f64":
        assert err <= TOL
    else:
        assert err <= 1e-5 and infid <= 1e-5


def test_small_registers_take_the_plain_kernels(orc):
    for n in range(0, 6):
        g = np.random.default_rng(n)
        specs = random_circuit(g, max(n, 1), 40) if n else []
        ref = orc.Reg(n)
        ref.apply(build(specs, orc.op))
        with QReg(n) as r:
            r.apply_fused(build(specs, op))
            assert np.array_equal(r.get_state(), ref.state())


def test_brickwork_and_measure(orc):
    n = 20
    specs = sum((circuits.brickwork_layer(n, l) for l in range(6)), [])
    ref = orc.Reg(n, threads=8)
    ref.apply(circuits.build(specs, orc.op))
    with QReg(n) as r:
        passes = r.apply_fused(circuits.build(specs, op))
        got = r.get_state()
        assert np.max(np.abs(got - ref.state())) <= TOL
        assert passes * 10 < len(circuits.build(specs, op))
        # measurement of the fused state follows the contract on that state
        u = qvnt_amd.uniform(7, 7, 7)
        want = orc.tree_sample(got, u) & 0xFF0F
        assert r.measure_mask(0xFF0F, u=u).get() == want


@pytest.mark.parametrize("n,prec", [(22, "f64"), (23, "f32")])
def test_tfim_ring_fused(orc, n, prec):
    """Trotter steps on a ring: the wrap-around rzz crosses tiles and
    phases (factored two-qubit diagonals, controlled phases, diagonals
    deferred to later passes); one call per step, then several steps in
    one call, against the oracle."""
    say("tfim ring n=%d %s" % (n, prec))
    g = np.random.default_rng(900 + n)
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    step = circuits.tfim_step(n)
    ref = orc.Reg(n, threads=8)
    ref.set_state(psi)
    for _ in range(5):
        ref.apply(circuits.build(step, orc.op))
    want = ref.state()
    three = circuits.build(step, op)
    for _ in range(2):
        three *= circuits.build(step, op)
    with QReg(n, prec=prec) as r:
        r.set_state(psi)
        for _ in range(2):
            r.apply_fused(circuits.build(step, op))
        r.apply_fused(three)
        got = r.get_state().astype(np.complex128)
    err = np.max(np.abs(got - want))
    say("  max |d| = %.2e" % err)
    assert err <= (TOL if prec == "f64" else 1e-5)


def test_brickwork_layers_single_gate_passes(orc):
    """One call per brickwork layer at 2^20 amplitudes: passes that would
    hold only the layer's controlled x run as its plain kernel."""
    n = 20
    layers = [circuits.brickwork_layer(n, l) for l in range(6)]
    ref = orc.Reg(n, threads=8)
    for specs in layers:
        ref.apply(circuits.build(specs, orc.op))
    with QReg(n) as r:
        passes = [r.apply_fused(circuits.build(specs, op)) for specs in layers]
        got = r.get_state()
    err = np.max(np.abs(got - ref.state()))
    say("  passes per layer %s, max |d| = %.2e" % (passes, err))
    assert err <= TOL


def test_plain_20q_fused(orc):
    state, specs, mask = circuits.plain_20q()
    ref = orc.Reg(20, state, threads=8)
    ref.apply(circuits.build(specs, orc.op))
    with QReg.with_state(20, state) as r:
        passes = r.apply_fused(circuits.build(specs, op))
        got = r.get_state()
    say("plain_20q fused passes", passes)
    assert np.max(np.abs(got - ref.state())) <= TOL


def test_fp32_fused_infidelity(orc):
    n = 20
    specs = sum((circuits.brickwork_layer(n, l) for l in range(10)), [])
    ref = orc.Reg(n, threads=8)
    ref.apply(circuits.build(specs, orc.op))
    want = ref.state()
    with QReg(n, prec="f32") as r:
        r.apply_fused(circuits.build(specs, op))
        got = r.get_state().astype(np.complex128)
    infid = 1.0 - abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    err = np.max(np.abs(got - want))
    say("fp32 fused: max err %.2e infidelity %.2e" % (err, infid))
    assert infid <= 1e-5 and err < 1e-5


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_auto_tile_28q(prec):
    """From 28 qubits, max_qubits = 0 picks the tile size by the run-time
    model; the program it keeps agrees with the plain form like any other."""
    n = 28
    say("auto tile n=%d %s" % (n, prec))
    specs = sum((circuits.brickwork_layer(n, l) for l in range(6)), [])
    m = circuits.build(specs, op)
    with QReg(n, prec=prec) as a, QReg(n, prec=prec) as b, QReg(n, prec=prec) as c:
        a.apply(m)
        pb = b.apply_fused(m)
        pc = c.apply_fused(m, max_qubits=12)
        err_b, infid_b = a.compare(b)[:2]
        err_c, infid_c = a.compare(c)[:2]
        say("passes auto %d, tile 12 %d; max |d| %.2e / %.2e, infidelity %.1e / %.1e"
            % (pb, pc, err_b, err_c, infid_b, infid_c))
        if prec == "f64":
            assert err_b <= TOL and err_c <= TOL
        else:
            assert infid_b <= 1e-5 and infid_c <= 1e-5


def test_fp32_tiles_of_13_at_30_qubits():
    """FP32 tiles of 13 qubits (64 KiB of LDS, 512 threads), which the tile
    choice takes for one-layer calls: three brickwork layers at 30 qubits,
    one call per layer (3 passes each: rows 0..3 and 9 free qubits a pass)
    and one call for all three, against the plain form on the device."""
    n = 30
    layers = [circuits.build(circuits.brickwork_layer(n, l), op) for l in range(3)]
    whole = layers[0].clone()
    for o in layers[1:]:
        whole *= o
    say("fp32 tiles of 13 at 30 q")
    with QReg(n, prec="f32") as a, QReg(n, prec="f32") as b, QReg(n, prec="f32") as c:
        passes = 0
        for o in layers:
            a.apply(o)
            passes += b.apply_fused(o, max_qubits=13)
        pc = c.apply_fused(whole, max_qubits=13)
        err_b, infid_b = a.compare(b)[:2]
        err_c, infid_c = a.compare(c)[:2]
    say("  passes %d per layer, %d whole; max |d| %.2e / %.2e, infidelity %.1e / %.1e"
        % (passes, pc, err_b, err_c, infid_b, infid_c))
    assert passes == 9
    assert infid_b <= 1e-5 and infid_c <= 1e-5 and err_b <= 1e-5 and err_c <= 1e-5


def test_fused_30q_round_trip():
    n = 30
    say("fused round trip n=30")
    fwd = sum((circuits.brickwork_layer(n, l) for l in range(4)), [])
    inv = []
    for ctor, args, ctrl in reversed(fwd):
        if ctor == "u3":
            th, ph, la, m = args
            inv += [("rz", (-ph, m), ctrl), ("ry", (-th, m), ctrl), ("rz", (-la, m), ctrl)]
        elif ctor == "rzz":
            inv.append(("rzz", (-args[0], args[1]), ctrl))
        else:
            inv.append((ctor, args, ctrl))
    start = 0x15555555
    with QReg.with_state(n, start) as r:
        p1 = r.apply_fused(circuits.build(fwd, op))
        assert abs(r.get_absolute() - 1.0) < 1e-12
        p2 = r.apply_fused(circuits.build(inv, op))
        outs = r.sample_mask((1 << n) - 1, [qvnt_amd.uniform(1, 2, k) for k in range(16)])
        say("passes", p1, p2)
        assert all(int(o) == start for o in outs)


def _inverse(specs):
    """The exact inverse of brickwork / TFIM / mixed specs: reversed order,
    rotations negated (a rotation's .dgr() is -U in the reference)."""
    inv = []
    for ctor, args, ctrl in reversed(specs):
        if ctor == "u3":
            th, ph, la, m = args
            inv += [("rz", (-ph, m), ctrl), ("ry", (-th, m), ctrl), ("rz", (-la, m), ctrl)]
        elif ctor in ("rx", "ry", "rz", "rxx", "ryy", "rzz", "phi"):
            inv.append((ctor, (-args[0],) + tuple(args[1:]), ctrl))
        else:
            inv.append((ctor, args, ctrl))
    return inv


@pytest.mark.parametrize("workload", ["brickwork", "mixed"])
def test_fused_equals_plain_at_30_qubits(workload):
    """At the full 30 qubits (16 GiB per register) the fused form agrees with
    the plain form -- which equals the oracle bit for bit -- to 1e-12,
    compared on the device: 3 brickwork layers, or 300 gates of the
    batched/serving gate set (swaps, iswaps, controlled x, rxx/ryy/rzz)."""
    n = 30
    say("fused vs plain at 30 q: %s" % workload)
    if workload == "brickwork":
        specs = sum((circuits.brickwork_layer(n, l) for l in range(3)), [])
    else:
        specs = circuits.random_mixed(n, 300, np.random.default_rng(3030))
    o = circuits.build(specs, op)
    with QReg(n) as a, QReg(n) as b:
        a.apply(o)
        b.apply_fused(o)
        err, infid, _, na, nb = a.compare(b)
    say("  max |d| %.2e, infidelity %.2e" % (err, infid))
    assert err <= TOL and abs(infid) <= 1e-12 and abs(na - 1.0) <= 1e-12 and abs(nb - 1.0) <= 1e-12


@pytest.mark.parametrize("n,prec", [(34, "f64"), (35, "f32")])
def test_tfim_ring_round_trip_largest_register(n, prec):
    """The TFIM ring at the configs' full sizes (256 GiB): 5 Trotter steps
    through the fused form and then their exact inverse return |0...0>
    (every shot 0); the norm holds to 1e-12 (FP64) / 1e-5 (FP32)."""
    say("tfim round trip n=%d %s" % (n, prec))
    fwd = circuits.tfim_step(n) * 5
    with QReg(n, prec=prec) as r:
        r.apply_fused(circuits.build(fwd, op))
        tol = 1e-12 if prec == "f64" else 1e-5
        assert abs(r.get_absolute() - 1.0) < tol
        r.apply_fused(circuits.build(_inverse(fwd), op))
        outs = r.sample_mask((1 << n) - 1, [qvnt_amd.uniform(8, n, k) for k in range(8)])
        assert all(int(x) == 0 for x in outs)
        assert abs(r.get_absolute() - 1.0) < tol


def permutation_circuit(g, n, count):
    """x (some controlled), swap and iswap (some daggered, some controlled):
    every amplitude is moved, or moved and multiplied by +-i, so the plain
    and the fused form must agree bit for bit."""
    out = op.id()
    for _ in range(count):
        kind = int(g.integers(6))
        a = int(g.integers(n))
        b = (a + 1 + int(g.integers(n - 1))) % n
        c = (b + 1 + int(g.integers(n - 2))) % n
        if c == a:
            c = (c + 1) % n
        if kind == 0:
            o = op.x(1 << a)
        elif kind == 1:
            o = op.x(1 << a).c(1 << b)
        elif kind == 2:
            o = op.swap((1 << a) | (1 << b))
        elif kind == 3:
            o = op.i_swap((1 << a) | (1 << b))
            if g.integers(2):
                o = o.dgr()
        elif kind == 4 and c != b:
            o = op.swap((1 << a) | (1 << b)).c(1 << c)
        else:
            o = op.i_swap((1 << a) | (1 << b))
        out *= o
    return out


@pytest.mark.parametrize("n,tq,prec", [(10, 10, "f64"), (16, 10, "f64"), (20, 10, "f64"), (20, 12, "f32"),
                                       (24, 11, "f64")])
def test_relabeled_permutations_bit_exact(n, tq, prec):
    """Uncontrolled swaps and iswaps are relabelings in the fused form and
    uncontrolled flips move to the end of their pass; each pass then
    permutes its tile's qubits in the LDS read-out, and passes without
    gates bring the last qubits home.  A circuit of permutations checks
    all of it exactly, against the plain form (one kernel per SingleOp),
    from a random state."""
    g = np.random.default_rng(7100 + n + tq)
    circ = permutation_circuit(g, n, 300)
    dt = np.complex128 if prec == "f64" else np.complex64
    psi = (g.standard_normal(1 << n) + 1j * g.standard_normal(1 << n)).astype(dt)
    with QReg(n, prec=prec) as r:
        r.set_state(psi)
        r.apply(circ)
        want = r.get_state()[: 1 << n]
        r.set_state(psi)
        passes = r.apply_fused(circ, max_qubits=tq)
        got = r.get_state()[: 1 << n]
    say("permutations n=%d tq=%d %s: %d gates, %d passes" % (n, tq, prec, len(circ), passes))
    assert np.array_equal(got, want)
