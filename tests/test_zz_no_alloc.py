"""The allocation-free shots entry point (qvnt_reg_sample_mask_noalloc)
against qvnt_reg_sample_mask: the same outcomes for the same draws, on two
calls in a row (the second on a smaller register and fewer shots, so the
fixed workspace holds the first call's larger contents), in FP64 and FP32,
up to the largest register its workspace is sized for (30 qubits)."""
import ctypes

import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import QReg, circuits, op
from qvnt_amd._lib import E_ARG, lib

pytestmark = pytest.mark.gpu


def prepared(n, prec, seed):
    """A register of n qubits in a seeded non-trivial state."""
    g = np.random.default_rng(seed)
    r = QReg(n, prec=prec)
    r.apply(circuits.build(circuits.random_mixed(n, 60, g), op))
    return r


def draws(seed, shots):
    return np.array([qvnt_amd.uniform(seed, 7, k) for k in range(shots)])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_equals_sample_mask_two_calls(prec):
    for n, shots, mask in ((22, 4096, 0b1011_0110_0000_1111_0101), (14, 300, 0b10_0110_1001)):
        print("noalloc %s: %d qubits, %d shots" % (prec, n, shots), flush=True)
        with prepared(n, prec, n) as r:
            u = draws(n, shots)
            want = r.sample_mask(mask, u)
            got = r.sample_mask_noalloc(mask, u)
            assert np.array_equal(got, want)
            assert np.all((got & ~np.uint64(mask)) == 0)
            # the state is untouched: a second pass gives the same outcomes
            assert np.array_equal(r.sample_mask_noalloc(mask, u), want)


def test_noalloc_largest_register():
    print("noalloc: 30 qubits FP64", flush=True)
    with QReg(30) as r:
        r.apply(op.h((1 << 30) - 1))
        r.apply(op.rz(0.7, 1 << 29))
        r.apply(op.ry(1.1, 1 << 3))
        u = draws(30, 1024)
        mask = (1 << 30) - 1
        want = r.sample_mask(mask, u)
        assert np.array_equal(r.sample_mask_noalloc(mask, u), want)


def test_noalloc_refuses_what_its_workspace_cannot_hold():
    with QReg(20) as r:
        r.apply(op.h(0xFF))
        shots = 1 << 20   # 2 x 8 MiB of draws and outcomes: more than the 16 MiB workspace
        u = np.zeros(shots)
        out = np.empty(shots, np.uint64)
        rc = lib.qvnt_reg_sample_mask_noalloc(r.handle, 0xFF, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), shots,
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert rc == E_ARG
        assert "workspace" in qvnt_amd.last_error()
        # and a call that fits still works afterwards
        u = draws(3, 64)
        assert np.array_equal(r.sample_mask_noalloc(0xFF, u), r.sample_mask(0xFF, u))
