"""QReg / CReg over the device register of libqvnt_hip.so.

Mirrors qvnt 0.4.4's src/register/quant.rs (QReg) and class.rs (CReg, the
measurement result).  The amplitudes stay in HBM; get_state/set_state copy
them out/in.  Measurement draws come from a seeded generator (the
reference's thread_rng is unseeded): a register created with ``seed`` draws
``qvnt_uniform(seed, stream, k)`` for its k-th measurement unless the caller
passes the draw ``u`` itself.
"""
import ctypes

import numpy as np

from ._lib import F32, F64, KCLASSES, MEASURE_RENORM, KStats, check, lib
from .ops import MultiOp

DEFAULT_SEED = 0x51564E54


def uniform(seed, stream, index):
    """The seeded draw in [0, 1) (SplitMix64; include/qvnt.h qvnt_uniform)."""
    return lib.qvnt_uniform(seed, stream, index)


class CReg:
    """Classical result of a measurement (src/register/class.rs:8-12, 66-74)."""

    def __init__(self, q_num, value=0):
        self.q_num = q_num
        self.q_mask = (1 << q_num) - 1 if q_num < 64 else (1 << 64) - 1
        self.value = value

    def get(self):
        return self.value

    def __repr__(self):
        return "(" + "".join("1" if self.value >> k & 1 else "0" for k in reversed(range(self.q_num))) + ")"

    def __eq__(self, other):
        return isinstance(other, CReg) and (self.q_num, self.value) == (other.q_num, other.value)


class QReg:
    """A state vector of max(2^n, 8) amplitudes in HBM (quant.rs:103-108)."""

    def __init__(self, q_num, prec="f64", state=0, seed=DEFAULT_SEED, stream=0):
        self._h = None
        p = {"f64": F64, "f32": F32}[prec]
        h = ctypes.c_void_p()
        check(lib.qvnt_reg_with_state(q_num, state, p, ctypes.byref(h)))
        self._h = h
        self.q_num = q_num
        self.prec = prec
        self.seed = seed
        self.stream = stream
        self._draws = 0

    @classmethod
    def new(cls, q_num, prec="f64", **kw):
        """QReg::new (quant.rs:113-125)."""
        return cls(q_num, prec=prec, **kw)

    @classmethod
    def with_state(cls, q_num, state, prec="f64", **kw):
        """QReg::with_state (quant.rs:129-150)."""
        return cls(q_num, prec=prec, state=state, **kw)

    def close(self):
        if self._h:
            lib.qvnt_reg_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def num(self):
        return self.q_num

    def __len__(self):
        return lib.qvnt_reg_len(self._h)

    def apply(self, op):
        """QReg::apply (quant.rs:376-395): one kernel per SingleOp, queue order."""
        if not isinstance(op, MultiOp):
            raise TypeError("apply expects a MultiOp")
        check(lib.qvnt_reg_apply(self._h, op.handle))
        return self

    def apply_fused(self, op, max_qubits=0):
        """The fused form: returns the number of passes over the state."""
        if not isinstance(op, MultiOp):
            raise TypeError("apply_fused expects a MultiOp")
        passes = ctypes.c_uint32(0)
        check(lib.qvnt_reg_apply_fused(self._h, op.handle, max_qubits, ctypes.byref(passes)))
        return passes.value

    def set_dense(self, min_ops):
        """Fused form: fold phases of >= min_ops primitives on four register
        qubits into dense 16x16 unitaries on the matrix cores (0: never)."""
        check(lib.qvnt_reg_set_dense(self._h, min_ops))

    def apply_gates(self, gates):
        arr = (type(gates[0]) * len(gates))(*gates) if gates else None
        check(lib.qvnt_reg_apply_gates(self._h, arr, len(gates)))
        return self

    def get_absolute(self):
        out = ctypes.c_double()
        check(lib.qvnt_reg_get_absolute(self._h, ctypes.byref(out)))
        return out.value

    def get_probabilities(self):
        n = 1 << self.q_num
        out = np.empty(n, np.float64)
        check(lib.qvnt_reg_get_probabilities(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out

    def next_draw(self):
        u = uniform(self.seed, self.stream, self._draws)
        self._draws += 1
        return u

    def measure_mask(self, mask, u=None, renorm=False):
        """QReg::measure_mask (quant.rs:490-501) with a seeded draw."""
        if u is None and (mask & ((1 << self.q_num) - 1)):
            u = self.next_draw()
        out = ctypes.c_uint64()
        check(lib.qvnt_reg_measure_mask(self._h, mask, 0.0 if u is None else u,
                                        MEASURE_RENORM if renorm else 0, ctypes.byref(out)))
        return CReg(self.q_num, out.value)

    def measure(self, u=None, renorm=False):
        """QReg::measure (quant.rs:505-507)."""
        return self.measure_mask((1 << self.q_num) - 1, u=u, renorm=renorm)

    def sample_mask(self, mask, u):
        """Non-destructive shots: measure_mask on copies of the register."""
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(len(u), np.uint64)
        check(lib.qvnt_reg_sample_mask(self._h, mask, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(u),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return out

    def compare(self, other):
        """(max |a_i - b_i|, infidelity, <a|b>, |a|^2, |b|^2), computed on the device in FP64."""
        out = (ctypes.c_double * 5)()
        check(lib.qvnt_reg_compare(self._h, other.handle, out))
        ov = complex(out[0], out[1])
        infid = 1.0 - abs(ov) ** 2 / (out[2] * out[3]) if out[2] > 0 and out[3] > 0 else float("nan")
        return out[4], infid, ov, out[2], out[3]

    def get_state(self):
        dt = np.complex128 if self.prec == "f64" else np.complex64
        out = np.empty(len(self), dt)
        check(lib.qvnt_reg_get_state(self._h, out.ctypes.data, out.nbytes))
        return out

    def set_state(self, psi):
        dt = np.complex128 if self.prec == "f64" else np.complex64
        psi = np.ascontiguousarray(psi, dt)
        check(lib.qvnt_reg_set_state(self._h, psi.ctypes.data, psi.nbytes))
        return self

    def sync(self):
        check(lib.qvnt_reg_sync(self._h))

    def set_timing(self, enable=True):
        check(lib.qvnt_reg_set_timing(self._h, 1 if enable else 0))

    def kernel_stats(self, reset=False):
        arr = (KStats * len(KCLASSES))()
        check(lib.qvnt_reg_kernel_stats(self._h, arr, 1 if reset else 0))
        return {k: {"launches": s.launches, "ms": s.ms, "bytes": s.bytes} for k, s in zip(KCLASSES, arr)}
