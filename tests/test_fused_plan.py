"""The fused form's compile-only statistics (qvnt_fused_plan_stats, host
only) and the tile choice from 28 qubits.

The run-time model behind the choice was fitted to measured applies on
MI355X: for the 30-qubit FP64 brickwork (40 layers) one call per layer ran
fastest in tiles of 12 qubits (0.990 s against 1.115 s with 11), the whole
circuit in one call in tiles of 11 (0.626 s against 0.686 s with 12); the
30-qubit TFIM ring likewise (per step 12, 50 steps in one call 11).  FP32
has a model of its own (its rows are longer, see below)."""
import pytest

import qvnt_amd
from qvnt_amd import circuits, fused_plan_stats, op


def merged(ops):
    m = ops[0].clone()
    for o in ops[1:]:
        m *= o
    return m


@pytest.fixture(scope="module")
def brickwork30():
    return [circuits.build(circuits.brickwork_layer(30, l), op) for l in range(40)]


def test_auto_tile_follows_the_measured_optimum_fp64(brickwork30):
    whole = merged(brickwork30)
    assert fused_plan_stats(whole, 30, "f64")["tile"] == 11
    # one call per layer: mostly tiles of 12; a layer that needs as many
    # passes in smaller tiles takes those (cheaper ops), so the layers
    # together still make the 120 passes of tiles of 12
    layers = [fused_plan_stats(o, 30, "f64") for o in brickwork30]
    assert sum(s["tile"] == 12 for s in layers) >= 30
    assert sum(s["passes"] for s in layers) == 120
    for o, s in zip(brickwork30, layers):
        if s["tile"] != 12:
            assert s["passes"] == fused_plan_stats(o, 30, "f64", max_qubits=12)["passes"]
    tfim = circuits.build(circuits.tfim_step(30), op)
    assert fused_plan_stats(tfim, 30, "f64")["tile"] == 12
    assert fused_plan_stats(merged([tfim] * 50), 30, "f64")["tile"] == 11


def test_auto_tile_follows_the_measured_optimum_fp32(brickwork30):
    """FP32 tile rows hold four qubits (128-byte lines), so a pass has one
    free qubit fewer than in FP64; an FP32 tile may hold 13 qubits (64 KiB
    of LDS).  Measured at 30 qubits, tiles of 12 against 13: per layer
    0.514 against 0.483 s; whole 0.303 against 0.304 s; TFIM per step
    0.656 against 0.625 s; TFIM whole 0.409 against 0.428 s (tile 11:
    0.512, 0.328, 0.661 and 0.457 s in an earlier run)."""
    assert fused_plan_stats(merged(brickwork30), 30, "f32")["tile"] in (12, 13)
    assert all(fused_plan_stats(o, 30, "f32")["tile"] == 13 for o in brickwork30)
    tfim = circuits.build(circuits.tfim_step(30), op)
    assert fused_plan_stats(tfim, 30, "f32")["tile"] == 13
    assert fused_plan_stats(merged([tfim] * 50), 30, "f32")["tile"] == 12
    # the rows: every pass of the same circuit holds qubits 0..3 in FP32
    # and needs at least as many passes as in FP64
    for o in brickwork30[:4]:
        assert (fused_plan_stats(o, 30, "f32", max_qubits=12)["passes"]
                >= fused_plan_stats(o, 30, "f64", max_qubits=12)["passes"])


def test_stats_of_explicit_and_small_settings(brickwork30):
    whole = merged(brickwork30)
    fixed = fused_plan_stats(whole, 30, "f64", max_qubits=12)
    assert fixed["tile"] == 12 and fixed["dense_phases"] == 0
    # the same passes the fused benchmark reports for this circuit in tiles of 12
    assert fixed["passes"] == 51
    auto = fused_plan_stats(whole, 30, "f64")
    assert auto["passes"] == 59 and auto["cost"] > 0
    # dense phases or an explicit tile switch the choice off; so does a
    # register below 28 qubits
    assert fused_plan_stats(whole, 30, "f64", dense_min=3)["tile"] == 12
    small = [circuits.build(circuits.brickwork_layer(20, l), op) for l in range(4)]
    assert fused_plan_stats(merged(small), 20, "f64")["tile"] == 12
    assert fused_plan_stats(merged(small), 20, "f32", max_qubits=13)["tile"] == 13
    assert fused_plan_stats(merged(small), 20, "f64", max_qubits=13)["tile"] == 12  # FP64 tiles hold 2^12
    # below 2^6 amplitudes the plain kernels run, one launch per gate
    h = op.h(0b11111)
    st = fused_plan_stats(h, 5, "f64")
    assert st["tile"] == 0 and st["passes"] == len(h.gates())


def test_stats_refuse_bad_arguments():
    with pytest.raises(qvnt_amd.QvntError):
        fused_plan_stats(op.x(1), 0, "f64")
    with pytest.raises(qvnt_amd.QvntError):
        fused_plan_stats(op.x(1 << 12), 10, "f64")  # beyond the register


@pytest.mark.parametrize("prec,rows", [("f64", 3), ("f32", 4)])
def test_every_brickwork_layer_takes_the_fewest_passes(brickwork30, prec, rows):
    """A pass holds the row qubits and tile - rows others, so a layer that
    acts on all 30 qubits needs ceil((30 - rows) / (12 - rows)) passes in
    tiles of 12.  The layer's cx whose control lies above its target puts
    the control's one-qubit gates in a later pass than the target's when
    passes are formed from the queue's start (4 passes in FP64 for layers
    7, 9, 11, 14 and 18); formed from its end they fit in 3."""
    fewest = -(-(30 - rows) // (12 - rows))
    assert all(fused_plan_stats(o, 30, prec, max_qubits=12)["passes"] == fewest for o in brickwork30)
