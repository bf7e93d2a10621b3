"""Reset, register algebra, readback and the QASM driver loop on the device
against the oracle (quant.rs:156-166, 202-229, 245-431, 513-594;
qasm/sym.rs:40-74)."""
import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import Program, QReg, Sym, op

pytestmark = pytest.mark.gpu


def rand_state(g, n):
    psi = g.normal(size=1 << n) + 1j * g.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


def pair(orc, g, n, prec="f64"):
    psi = rand_state(g, n)
    full = np.zeros(max(1 << n, 8), np.complex128)
    full[: 1 << n] = psi
    r = QReg(n, prec=prec)
    r.set_state(full)
    o = orc.Reg(n)
    o.set_state(full)
    return r, o


@pytest.mark.parametrize("na,nb", [(0, 3), (1, 1), (2, 3), (4, 2), (10, 8)])
def test_tensor(orc, na, nb):
    g = np.random.default_rng(100 + 10 * na + nb)
    (a, oa), (b, ob) = pair(orc, g, na), pair(orc, g, nb)
    t = a * b
    assert t.q_num == na + nb
    assert np.array_equal(t.get_state(), (oa * ob).state())
    # inputs are kept and can be used afterwards
    a.apply(op.x(1) if na else op.id())
    t.close()


def test_tensor_fp32(orc):
    g = np.random.default_rng(3)
    (a, oa), (b, ob) = pair(orc, g, 6, "f32"), pair(orc, g, 5, "f32")
    got = (a * b).get_state()
    assert got.dtype == np.complex64
    assert np.max(np.abs(got - (oa * ob).state())) < 1e-6


@pytest.mark.parametrize("n", [3, 7, 12])
def test_combine(orc, n):
    g = np.random.default_rng(200 + n)
    (a, oa), (b, ob) = pair(orc, g, n), pair(orc, g, n)
    assert np.array_equal(a.combine(b).get_state(), oa.combine(ob).state())
    th = 1.1
    m = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]], np.complex128)
    assert np.array_equal(a.combine_with_unitary(b, m).get_state(), oa.combine_with_unitary(ob, m).state())
    with pytest.raises(qvnt_amd.QvntError):
        a.combine_with_unitary(b, np.array([[1, 1], [0, 1]]))


def test_combine_rejects_what_the_reference_rejects():
    with pytest.raises(qvnt_amd.QvntError):
        QReg(2).combine(QReg(2))
    with pytest.raises(qvnt_amd.QvntError):
        QReg(4).combine(QReg(5))


def test_linear_composition_and_set_num(orc):
    g = np.random.default_rng(9)
    (a, oa), (b, ob) = pair(orc, g, 11), pair(orc, g, 11)
    c0, c1 = 0.6 + 0.1j, -0.3 + 0.7j
    a.linear_composition(b, c0, c1)
    oa.linear_composition(ob, c0, c1)
    assert np.array_equal(a.get_state(), oa.state())
    a.set_num(7)
    oa.set_num(7)
    assert np.array_equal(a.get_state(), oa.state())
    a.set_num(13)
    oa.set_num(13)
    assert np.array_equal(a.get_state(), oa.state()) and a.q_num == 13 and len(a) == 1 << 13


@pytest.mark.parametrize("n,mask", [(5, 0b00101), (12, 0b100000000001), (16, 0xFFFF), (20, 1 << 19)])
def test_reset_by_mask(orc, n, mask):
    g = np.random.default_rng(300 + n)
    r, o = pair(orc, g, n)
    r.reset_by_mask(mask)
    o.reset_by_mask(mask)
    assert np.max(np.abs(r.get_state() - o.state())) <= 1e-15
    r2 = QReg.with_state(4, 0b0001)
    r2.reset_by_mask(0b0001)                    # the only amplitude goes: reset(0)
    assert np.array_equal(r2.get_state(), np.eye(16)[0])


@pytest.mark.parametrize("n,prec", [(4, "f64"), (14, "f64"), (10, "f32")])
def test_get_polar(orc, n, prec):
    g = np.random.default_rng(400 + n)
    r, o = pair(orc, g, n, prec)
    got, want = r.get_polar(), o.get_polar()
    tol = 1e-15 if prec == "f64" else 1e-6
    assert np.max(np.abs(got[:, 0] - want[:, 0])) <= tol
    assert np.max(np.abs(got[:, 1] - want[:, 1])) <= tol * 4


@pytest.mark.parametrize("n,shots", [(8, 2048), (13, 100000)])
def test_sample_all(orc, n, shots):
    g = np.random.default_rng(500 + n)
    r, o = pair(orc, g, n)
    h = r.sample_all(shots, seed=5, stream=n)
    assert len(h) == len(r) and int(h.sum()) == shots
    assert np.array_equal(h, o.sample_all(shots, 5, n))
    q = QReg.with_state(8, 123)                 # quant.rs:713-724
    q.apply(op.h(255))
    for k in range(10):
        hist = q.sample_all(2048, stream=k)
        assert len(hist) == 256 and int(hist.sum()) == 2048


def random_program(g, n, c_num, steps):
    """(device steps, oracle steps, tail) of a random driver program."""
    from test_gpu_parity import build, random_circuit
    dev, ref = [], []
    for k in range(steps):
        specs = random_circuit(g, n, int(g.integers(1, 12)))
        kind = int(g.integers(0, 4))
        a = int(g.integers(1, 1 << n))
        b = int(g.integers(0, 1 << c_num))
        if kind == 2:
            a = int(g.integers(0, 1 << c_num))
            b = int(g.integers(0, 1 << bin(a).count("1")))
        dev.append((build(specs, qvnt_amd.op), kind, a, b))
        ref.append((build(specs, __import__("oracle").op), kind, a, b))
    tail = random_circuit(g, n, 5)
    return dev, ref, tail


@pytest.mark.parametrize("n,seed,xor", [(4, 1, False), (6, 2, True), (9, 3, False), (12, 4, True)])
def test_program_driver(orc, n, seed, xor):
    from test_gpu_parity import build
    g = np.random.default_rng(600 + seed)
    c_num = 5
    dev, ref, tail = random_program(g, n, c_num, 24)
    draws = [qvnt_amd.uniform(seed, 9, k) for k in range(64)]
    r = QReg(n)
    o = orc.Reg(n)
    import ctypes
    from qvnt_amd._lib import Step, check, lib
    arr = (Step * len(dev))()
    for k, (mo, sep, a, b) in enumerate(dev):
        arr[k].op, arr[k].sep, arr[k].a, arr[k].b = mo.handle, sep, a, b
    cval, used = ctypes.c_uint64(0), ctypes.c_size_t(0)
    d = np.array(draws)
    tail_op = build(tail, qvnt_amd.op)          # held: the call borrows its handle
    check(lib.qvnt_reg_run_program(r.handle, arr, len(dev), tail_op.handle, 1 if xor else 0,
                                   ctypes.byref(cval), c_num, d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   len(d), ctypes.byref(used)))
    want_c, want_used = o.run_program(ref, build(tail, orc.op), c_num, measure_xor=xor, draws=draws)
    assert (cval.value, used.value) == (want_c, want_used)
    assert np.max(np.abs(r.get_state() - o.state())) <= 1e-12


def test_sym_teleport_like():
    """Program / Sym: x q0; measure q0 -> c0; if (c == 1) x q1; tail h q2."""
    p = Program(3, 2)
    p.apply(op.x(0b1)).measure(0b1, 0b1).if_eq(0b11, 1, op.x(0b10))
    p.apply(op.h(0b100))
    s = Sym(p)
    s.finish()
    assert s.get_class().get() == 1
    probs = s.get_probabilities()
    assert abs(probs[0b011] - 0.5) < 1e-12 and abs(probs[0b111] - 0.5) < 1e-12
    s.reset()
    assert s.get_class().get() == 0 and s.get_probabilities()[0] == 1.0


def test_reused_blocks_and_streams_start_clean(orc):
    """A freed register hands its state, workspace and stream back to the
    device's cache and the next register of the same sizes takes them: a
    register made on a reused block holds exactly its basis state, a resize
    onto a reused (stale) block holds the kept amplitudes and zeros, and a
    circuit with sampling and a measurement repeats bit for bit on reused
    blocks, also when the sampling workspace grows mid-life."""
    from qvnt_amd import circuits
    (n, specs), = circuits.batched_registers(1, seed=5, n_lo=12, n_hi=12, g_lo=60, g_hi=60)
    ops = circuits.build(specs, op)
    small = np.array([qvnt_amd.uniform(5, 0, j) for j in range(300)])
    big = np.array([qvnt_amd.uniform(5, 0, j) for j in range(5000)])
    first = None
    for it in range(4):
        s = 0x5A5 if it % 2 == 0 else 0x0F3
        with QReg.with_state(n, s) as r:
            psi = r.get_state()
            assert psi[s] == 1 and np.count_nonzero(psi) == 1, "iteration %d" % it
            if it % 2:
                continue
            r.apply(ops)
            r.apply_fused(ops)
            a = r.sample_mask(0xFF0, small)
            b = r.sample_mask(0xFF0, big)
            assert np.array_equal(b[:300], a)
            m = r.measure_mask(0b1010, u=0.25).get()
            got = (r.get_state(), a, b, m)
        if first is None:
            first = got
        else:
            assert all(np.array_equal(x, y) for x, y in zip(got, first)), "iteration %d" % it
    with QReg.with_state(n, 0) as r:
        r.apply(ops)            # leaves a scrambled 12-qubit block in the cache
    o = orc.Reg(10, 3)
    o.set_num(n)
    with QReg.with_state(10, 3) as r:
        r.set_num(n)
        assert np.array_equal(r.get_state(), o.state())
    qvnt_amd.cache_trim()           # everything cached goes back to the device
    with QReg.with_state(n, 0x5A5) as r:
        r.apply(ops)
        r.apply_fused(ops)
        assert np.array_equal(r.sample_mask(0xFF0, big), first[2])
        assert r.measure_mask(0b1010, u=0.25).get() == first[3]
        assert np.array_equal(r.get_state(), first[0])       # (the state after the measurement)
