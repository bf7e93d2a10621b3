"""The batched-register form: every register's result equals the
single-register fused form bit for bit (same tile cap), the plain form and
the oracle to 1e-12 (FP64), and stays within the FP32 bound in FP32."""
import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import QReg, circuits, op
from qvnt_amd.batch import TUNE_CAPS, TUNE_DENSE, Batch

pytestmark = pytest.mark.gpu


def workload(count, seed, n_lo=6, n_hi=14, g_lo=20, g_hi=200):
    regs = circuits.batched_registers(count, seed=seed, n_lo=n_lo, n_hi=n_hi, g_lo=g_lo, g_hi=g_hi)
    states = np.random.default_rng(seed).integers(0, 1 << 62, count, dtype=np.uint64)
    states = [int(s) & ((1 << n) - 1) for s, (n, _) in zip(states, regs)]
    return regs, states, [circuits.build(specs, op) for _, specs in regs]


def infidelity(a, b):
    return 1.0 - abs(np.vdot(a.astype(np.complex128), b.astype(np.complex128))) ** 2


@pytest.mark.parametrize("cap,dense", [(6, 0), (9, 0), (12, 0), (12, 2), (10, 3)])
def test_batch_equals_single_fused(cap, dense):
    regs, states, ops = workload(40, 100 + cap + dense, n_hi=16 if dense else 14)
    with Batch([n for n, _ in regs], states) as b:
        info = b.compile(ops, max_qubits=cap, dense_min=dense)
        b.run()
        assert info["single_ops"] == sum(len(o) for o in ops)
        if dense:
            assert info["dense_phases"] > 0
        for k, (n, _) in enumerate(regs):
            with QReg.with_state(n, states[k]) as r:
                r.set_dense(dense)
                r.apply_fused(ops[k], max_qubits=cap)
                want = r.get_state()
            assert np.array_equal(b.get_state(k), want[: 1 << n]), "register %d (%d qubits)" % (k, n)


@pytest.mark.parametrize("dense", [0, 2])
def test_batch_matches_plain_and_oracle(orc, dense):
    regs, states, ops = workload(24, 7 + dense, n_lo=6, n_hi=12)
    with Batch([n for n, _ in regs], states) as b:
        b.compile(ops, max_qubits=10, dense_min=dense)
        b.run()
        for k, (n, specs) in enumerate(regs):
            with QReg.with_state(n, states[k]) as r:
                r.apply(ops[k])
                plain = r.get_state()[: 1 << n]
            got = b.get_state(k)
            assert np.max(np.abs(got - plain)) <= 1e-12
            if k < 8:
                ref = orc.Reg(n, states[k])
                ref.apply(circuits.build(specs, orc.op))
                assert np.max(np.abs(got - ref.state()[: 1 << n])) <= 1e-12


def test_autotune_and_reset():
    regs, states, ops = workload(48, 3, n_lo=8, n_hi=14)
    with Batch([n for n, _ in regs], states) as b:
        info = b.compile(ops)                 # autotuner: runs the batch, then resets it
        assert len(info["tune_ms"]) == len(TUNE_CAPS) * len(TUNE_DENSE)
        for n, cap in info["tile_cap"].items():
            assert cap in TUNE_CAPS or cap == n   # FP64 register set
            assert info["dense_min"][n] in TUNE_DENSE
        b.run()
        got = [b.get_state(k) for k in range(len(regs))]
    for k, (n, _) in enumerate(regs):
        with QReg.with_state(n, states[k]) as r:
            r.apply(ops[k])
            assert np.max(np.abs(got[k] - r.get_state()[: 1 << n])) <= 1e-12


def test_graph_replays_match_eager():
    regs, states, ops = workload(32, 5, n_lo=6, n_hi=13)
    with Batch([n for n, _ in regs], states) as b:
        b.compile(ops, max_qubits=9)
        b.run(graph=False)
        eager = [b.get_state(k) for k in range(len(regs))]
        for rep in range(3):
            b.reset()
            b.run(graph=True)
            b.sync()
            for k in range(len(regs)):
                assert np.array_equal(b.get_state(k), eager[k]), "replay %d, register %d" % (rep, k)
            # eager work between replays
            b.reset()
            b.run(graph=False)
            assert np.array_equal(b.get_state(0), eager[0])
        # two runs in a row apply the circuits twice
        b.reset()
        b.run(graph=True)
        b.run(graph=True)
        twice = b.get_state(1)
    n = regs[1][0]
    with QReg.with_state(n, states[1]) as r:
        r.apply_fused(ops[1], max_qubits=9)
        r.apply_fused(ops[1], max_qubits=9)
        assert np.array_equal(twice, r.get_state()[: 1 << n])


@pytest.mark.parametrize("prec,cap,dense", [("f64", 12, 0), ("f64", 12, 2), ("f64", 9, 0), ("f32", 13, 3),
                                             ("f32", 10, 0)])
def test_persistent_launch_equals_rounds(prec, cap, dense):
    """The persistent cooperative launch (all rounds in one dispatch, a grid
    barrier between rounds) gives the round launches' amplitudes bit for
    bit.  Registers of 6..16 qubits under one tile cap mix tile sizes in a
    round, so a block runs several tiles of a small register at once; two
    runs in a row, twice after a reset."""
    regs, states, ops = workload(48, 31 + cap + dense, n_lo=6, n_hi=16 if dense else 14)
    with Batch([n for n, _ in regs], states, prec=prec) as b:
        info = b.compile(ops, max_qubits=cap, dense_min=dense)
        assert info["persist_blocks"] > 0
        if dense:
            assert info["dense_phases"] > 0
        b.run()
        b.run()
        want = [b.get_state(k) for k in range(len(regs))]
        assert b.persist_form is None
        for rep in range(2):
            b.reset()
            b.run(graph="persistent")
            b.run(graph="persistent")
            b.sync()
            print("persistent launch form:", b.persist_form)
            # a captured cooperative launch stays cooperative on this runtime
            assert b.persist_form == "graph"
            for k, (n, _) in enumerate(regs):
                assert np.array_equal(b.get_state(k), want[k]), "run %d, register %d (%d qubits)" % (rep, k, n)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_persistent_fixed_and_claimed_units(prec):
    """Rounds with more units than the grid has blocks: each round's units
    split into a fixed part (each XCD's blocks walk their own range) and a
    tail that blocks claim one at a time from the round's counter.  The
    launch equals the round launches bit for bit, three launches in a row
    (the counters must return to zero after each)."""
    regs, states, ops = workload(128, 77, n_lo=15, n_hi=17, g_lo=20, g_hi=80)
    with Batch([n for n, _ in regs], states, prec=prec) as b:
        info = b.compile(ops, max_qubits=10, dense_min=0)
        G = info["persist_blocks"]
        assert G > 0 and G % 8 == 0
        # round 0 holds one unit per tile of every register's first pass
        assert sum(1 << (n - 10) for n, _ in regs) * 0.85 >= 2 * G
        b.run()
        want = [b.get_state(k) for k in range(len(regs))]
        for rep in range(3):
            b.reset()
            b.run(graph="persistent")
            b.sync()
            assert b.persist_form == "graph"
            for k, (n, _) in enumerate(regs):
                assert np.array_equal(b.get_state(k), want[k]), "launch %d, register %d (%d qubits)" % (rep, k, n)


@pytest.mark.parametrize("dense,cap", [(0, 12), (2, 12), (0, 13), (3, 13)])
def test_fp32_batch_error_bound(dense, cap):
    regs, states, ops = workload(24, 9 + dense + cap, n_lo=8, n_hi=16)
    with Batch([n for n, _ in regs], states, prec="f32") as b:
        b.compile(ops, max_qubits=cap, dense_min=dense)
        b.run()
        for k, (n, _) in enumerate(regs):
            with QReg.with_state(n, states[k]) as r:
                r.apply(ops[k])
                ref = r.get_state()[: 1 << n]
            got = b.get_state(k)
            assert got.dtype == np.complex64
            assert np.max(np.abs(got - ref)) <= 1e-5
            assert infidelity(got, ref) <= 1e-5


def test_batch_sampling_matches_register():
    regs, states, ops = workload(6, 11, n_lo=10, n_hi=12)
    draws = np.array([qvnt_amd.uniform(11, 0, j) for j in range(512)])
    with Batch([n for n, _ in regs], states) as b:
        b.compile(ops, max_qubits=8)
        b.run()
        for k, (n, _) in enumerate(regs):
            psi = b.get_state(k)
            outs, total = b.sample_mask(k, 0xF0, draws)
            with QReg(n) as r:
                r.set_state(psi)
                assert np.array_equal(outs, r.sample_mask(0xF0, draws))
                assert total == r.get_absolute()


def test_batch_rejects_bad_input():
    with pytest.raises(qvnt_amd.QvntError):
        Batch([5, 8])
    with Batch([6, 7]) as b:
        with pytest.raises(qvnt_amd.QvntError):
            b.compile([op.x(1), op.x(1 << 7)])     # qubit 7 beyond a 7-qubit register
        with pytest.raises(TypeError):
            b.compile([op.x(1)])


def test_batch_at_config_sizes():
    """512 registers of 10-20 qubits (the batched workload's sizes, 20-120
    gates each), autotuned and run as a hipGraph: every 16th register equals
    the single-register fused form with the settings the autotuner chose,
    bit for bit, and every register keeps its norm."""
    regs, states, ops = workload(512, 4242, n_lo=10, n_hi=20, g_lo=20, g_hi=120)
    with Batch([n for n, _ in regs], states) as b:
        info = b.compile(ops)
        b.run(graph=True)
        b.sync()
        for k in range(0, len(regs), 16):
            n = regs[k][0]
            with QReg.with_state(n, states[k]) as r:
                r.set_dense(info["dense_min"][n])
                r.apply_fused(ops[k], max_qubits=info["tile_cap"][n])
                want = r.get_state()[: 1 << n]
            got = b.get_state(k)
            if n < 20:
                assert np.array_equal(got, want), "register %d (%d qubits)" % (k, n)
            else:   # from 2^20 amplitudes the single-register form runs one-gate passes as plain launches
                assert np.max(np.abs(got - want)) <= 1e-12, "register %d (%d qubits)" % (k, n)
            assert abs(np.vdot(got, got).real - 1.0) < 1e-12
