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


def fused_plan_stats(op, q_num, prec="f64", max_qubits=0, dense_min=0):
    """What QReg.apply_fused would compile for a register of q_num qubits
    (host only): passes over the state, the tile size used, the compiler's
    summed op-cost estimate and the number of dense phases."""
    out = (ctypes.c_double * 4)()
    check(lib.qvnt_fused_plan_stats(op.handle, q_num, F64 if prec == "f64" else F32, max_qubits, dense_min, out))
    return {"passes": int(out[0]), "tile": int(out[1]), "cost": out[2], "dense_phases": int(out[3])}


def cache_trim():
    """Give the device blocks and idle streams that freed registers left in
    the library's per-device cache back to the device (qvnt_cache_trim)."""
    check(lib.qvnt_cache_trim())


def uniform(seed, stream, index):
    """The seeded draw in [0, 1) (SplitMix64; include/qvnt.h qvnt_uniform)."""
    return lib.qvnt_uniform(seed, stream, index)


def _bits(mask):
    """The set bits of mask as single-bit masks, ascending (math/bits_iter.rs)."""
    out = []
    while mask:
        low = mask & -mask
        out.append(low)
        mask ^= low
    return out


class CReg:
    """Classical register (src/register/class.rs): a value of q_num bits."""

    def __init__(self, q_num, value=0):
        self.q_num = q_num
        self.q_mask = (1 << q_num) - 1 if q_num < 64 else (1 << 64) - 1
        self.value = value

    @classmethod
    def new(cls, q_num):
        return cls(q_num, 0)

    @classmethod
    def with_state(cls, q_num, state):
        return cls(q_num, state)

    def get(self):
        return self.value

    def num(self):
        return self.q_num

    def set_num(self, q_num):
        """class.rs:75-78: the value is kept as it is."""
        self.q_num = q_num
        self.q_mask = (1 << q_num) - 1 if q_num < 64 else (1 << 64) - 1

    def reset(self, i_state=0):
        self.value = i_state & self.q_mask

    def set(self, bit, mask):
        """class.rs:84-90: the bits of mask := bit."""
        self.value = (self.value | mask) if bit else (self.value & ~mask)

    def xor(self, bit, mask):
        if bit:
            self.value ^= mask

    def get_by_mask(self, mask):
        """class.rs:106-117: the value's bits at the set bits of mask & q_mask, packed from bit 0."""
        return sum(1 << k for k, b in enumerate(_bits(mask & self.q_mask)) if self.value & b)

    def __mul__(self, other):
        """Tensor product (class.rs:92-96): self on the low bits."""
        return CReg(self.q_num + other.q_num, self.value | (other.value << self.q_num))

    def __repr__(self):
        return "(" + "".join("1" if self.value >> k & 1 else "0" for k in reversed(range(self.q_num))) + ")"

    def __eq__(self, other):
        return isinstance(other, CReg) and (self.q_num, self.value) == (other.q_num, other.value)


class VReg:
    """Virtual register (src/register/virtl.rs): a list of single-bit masks,
    indexed by position, by a predicate on positions, by a list of
    positions, or whole ([:]), giving the OR of the selected masks."""

    def __init__(self, num):
        self.masks = _bits((1 << num) - 1)

    @classmethod
    def from_mask(cls, mask):
        v = cls.__new__(cls)
        v.masks = _bits(mask)
        return v

    def __len__(self):
        return len(self.masks)

    def __getitem__(self, key):
        if isinstance(key, int):
            return self.masks[key]
        if isinstance(key, slice):
            if key != slice(None):
                raise TypeError("only the full range [:] is supported")
            sel = range(len(self.masks))
        elif callable(key):
            sel = [i for i in range(len(self.masks)) if key(i)]
        else:
            sel = [i for i in range(len(self.masks)) if i in set(key)]
        out = 0
        for i in sel:
            out |= self.masks[i]
        return out

    def __repr__(self):
        return "[" + ", ".join("%032x" % m for m in self.masks) + "]"


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

    def sample_mask_noalloc(self, mask, u):
        """sample_mask through the allocation-free entry point (a fixed
        device workspace shared by every call on the GPU)."""
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(len(u), np.uint64)
        check(lib.qvnt_reg_sample_mask_noalloc(self._h, mask, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(u),
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

    def debug(self):
        """Debug of QReg (quant.rs:603-623): the first min(2^n, 8) amplitudes."""
        need = ctypes.c_size_t(0)
        check(lib.qvnt_reg_debug(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        check(lib.qvnt_reg_debug(self._h, buf, need.value, ctypes.byref(need)))
        return buf.value.decode()

    def __repr__(self):
        return self.debug() if self._h else "QReg(closed)"

    # -------------------------------------------- reset, algebra, readback
    @classmethod
    def _adopt(cls, h, q_num, prec, seed=DEFAULT_SEED):
        r = cls.__new__(cls)
        r._h, r.q_num, r.prec, r.seed, r.stream, r._draws = h, q_num, prec, seed, 0, 0
        return r

    def reset(self, i_state=0):
        """QReg::reset (quant.rs:202-205)."""
        check(lib.qvnt_reg_reset(self._h, i_state))

    def reset_by_mask(self, mask):
        """QReg::reset_by_mask (quant.rs:207-229): zero idx & mask != 0, then normalize."""
        check(lib.qvnt_reg_reset_by_mask(self._h, mask))

    def __mul__(self, other):
        """Tensor product (quant.rs:330-371): self on the low qubits; both are kept."""
        h = ctypes.c_void_p()
        check(lib.qvnt_reg_tensor(self._h, other.handle, ctypes.byref(h)))
        return QReg._adopt(h, self.q_num + other.q_num, self.prec, self.seed)

    def __imul__(self, other):
        out = self * other
        self.close()
        self._h, self.q_num = out._h, out.q_num
        out._h = None
        return self

    def combine(self, other):
        """QReg::combine (quant.rs:245-268): self, then other, one qubit larger."""
        h = ctypes.c_void_p()
        check(lib.qvnt_reg_combine(self._h, other.handle, ctypes.byref(h)))
        return QReg._adopt(h, self.q_num + 1, self.prec, self.seed)

    def combine_with_unitary(self, other, m):
        """QReg::combine_with_unitary (quant.rs:271-308); m: 2x2 complex."""
        c = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(4)).view(np.float64)
        h = ctypes.c_void_p()
        check(lib.qvnt_reg_combine_with_unitary(self._h, other.handle,
                                                c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h)))
        return QReg._adopt(h, self.q_num + 1, self.prec, self.seed)

    def linear_composition(self, other, c0, c1):
        """QReg::linear_composition (quant.rs:311-328): self = self * c0 + other * c1."""
        c = np.array([complex(c0).real, complex(c0).imag, complex(c1).real, complex(c1).imag], np.float64)
        check(lib.qvnt_reg_linear_composition(self._h, other.handle, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def set_num(self, q_num):
        """QReg::set_num (quant.rs:156-166): zero-extend or truncate to 2^q_num amplitudes."""
        check(lib.qvnt_reg_set_num(self._h, q_num))
        self.q_num = q_num

    def get_polar(self):
        """QReg::get_polar (quant.rs:417-431): rows (|psi_i|, arg psi_i)."""
        out = np.empty(2 << self.q_num, np.float64)
        check(lib.qvnt_reg_get_polar(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(out)))
        return out.reshape(-1, 2)

    def sample_all(self, count, seed=None, stream=None):
        """QReg::sample_all (quant.rs:513-594) as an exact multinomial of
        `count` seeded draws; one count per buffer element."""
        seed = self.seed if seed is None else seed
        stream = self.stream + 1 if stream is None else stream
        out = np.zeros(len(self), np.uint64)
        check(lib.qvnt_reg_sample_all(self._h, count, seed, stream, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                      len(out)))
        return out

    def get_vreg(self):
        """QReg::get_vreg (quant.rs:232-235)."""
        return VReg.from_mask((1 << self.q_num) - 1)

    def get_vreg_by(self, mask):
        """QReg::get_vreg_by (quant.rs:237-243): None for bits beyond the register."""
        if mask & ~((1 << self.q_num) - 1):
            return None
        return VReg.from_mask(mask)

    def set_timing(self, enable=True):
        check(lib.qvnt_reg_set_timing(self._h, 1 if enable else 0))

    def kernel_stats(self, reset=False):
        arr = (KStats * len(KCLASSES))()
        check(lib.qvnt_reg_kernel_stats(self._h, arr, 1 if reset else 0))
        return {k: {"launches": s.launches, "ms": s.ms, "bytes": s.bytes} for k, s in zip(KCLASSES, arr)}
