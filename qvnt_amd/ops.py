"""The op:: gate set and the MultiOp queue, mirroring qvnt 0.4.4's Rust API.

Reference (paths relative to its root): constructors src/operator/mod.rs:113-522,
queue semantics src/operator/multi/mod.rs:63-191, `.c()`/`.dgr()`
src/operator/applicable.rs:5-47.  Constructors take the reference argument
order (phase first, then mask).  Where the reference panics (a rotation or
swap mask with the wrong bit count) a ValueError is raised; where it returns
``None`` (``.c()`` overlapping the operator, checked matrix constructors),
``None`` is returned.
"""
import ctypes

import numpy as np

from ._lib import E_CTRL_OVERLAP, E_INVALID_MASK, Gate, QvntError, check, last_error, lib


def _new(fn, *args):
    h = ctypes.c_void_p()
    rc = fn(*args, ctypes.byref(h))
    if rc == E_INVALID_MASK:
        raise ValueError(last_error())
    check(rc)
    return MultiOp(h)


class MultiOp:
    """A queue of SingleOps applied front to back (multi/mod.rs:63)."""

    __slots__ = ("_h",)

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.qvnt_op_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def clone(self):
        return _new(lib.qvnt_op_clone, self._h)

    def __mul__(self, other):
        """`a * b`: b's gates run after a's (MulAssign appends, multi/mod.rs:166-170)."""
        out = self.clone()
        check(lib.qvnt_op_mul(out._h, other._h))
        return out

    def __imul__(self, other):
        check(lib.qvnt_op_mul(self._h, other._h))
        return self

    def c(self, mask):
        """Controlled version, or None if `mask` overlaps act|ctrl (multi/mod.rs:125-132)."""
        out = self.clone()
        rc = lib.qvnt_op_c(out._h, mask)
        if rc == E_CTRL_OVERLAP:
            return None
        check(rc)
        return out

    def dgr(self):
        """Reversed queue of daggered gates (multi/mod.rs:120-123)."""
        out = self.clone()
        check(lib.qvnt_op_dgr(out._h))
        return out

    def act_on(self):
        return lib.qvnt_op_act_on(self._h)

    def __len__(self):
        return lib.qvnt_op_len(self._h)

    def gates(self):
        n = len(self)
        if n == 0:
            return []
        p = lib.qvnt_op_gates(self._h)
        return [Gate.from_buffer_copy(p[i]) for i in range(n)]

    def __getitem__(self, i):
        gates = self.gates()
        g = gates[i]
        return from_gates([g] if isinstance(g, Gate) else g)

    def name(self):
        need = ctypes.c_size_t(0)
        check(lib.qvnt_op_name(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        check(lib.qvnt_op_name(self._h, buf, need.value, ctypes.byref(need)))
        return buf.value.decode()

    __repr__ = name

    def __eq__(self, other):
        if not isinstance(other, MultiOp):
            return NotImplemented
        return [bytes(g) for g in self.gates()] == [bytes(g) for g in other.gates()]

    def ends_with(self, suffix):
        """multi/mod.rs:65-71."""
        a, b = self.gates(), suffix.gates()
        return all(bytes(x) == bytes(y) for x, y in zip(reversed(a), reversed(b)))


def from_gates(gates):
    arr = (Gate * max(1, len(gates)))(*gates)
    return _new(lib.qvnt_op_from_gates, arr, len(gates))


class _OpNamespace:
    """`op::` (src/operator/mod.rs)."""

    @staticmethod
    def id():
        return _new(lib.qvnt_op_id)

    @staticmethod
    def x(a_mask):
        return _new(lib.qvnt_op_x, a_mask)

    @staticmethod
    def y(a_mask):
        return _new(lib.qvnt_op_y, a_mask)

    @staticmethod
    def z(a_mask):
        return _new(lib.qvnt_op_z, a_mask)

    @staticmethod
    def s(a_mask):
        return _new(lib.qvnt_op_s, a_mask)

    @staticmethod
    def t(a_mask):
        return _new(lib.qvnt_op_t, a_mask)

    @staticmethod
    def rx(phase, a_mask):
        return _new(lib.qvnt_op_rx, phase, a_mask)

    @staticmethod
    def ry(phase, a_mask):
        return _new(lib.qvnt_op_ry, phase, a_mask)

    @staticmethod
    def rz(phase, a_mask):
        return _new(lib.qvnt_op_rz, phase, a_mask)

    @staticmethod
    def rxx(phase, ab_mask):
        return _new(lib.qvnt_op_rxx, phase, ab_mask)

    @staticmethod
    def ryy(phase, ab_mask):
        return _new(lib.qvnt_op_ryy, phase, ab_mask)

    @staticmethod
    def rzz(phase, ab_mask):
        return _new(lib.qvnt_op_rzz, phase, ab_mask)

    @staticmethod
    def phi(phase, a_mask):
        return _new(lib.qvnt_op_phi, phase, a_mask)

    @staticmethod
    def swap(ab_mask):
        return _new(lib.qvnt_op_swap, ab_mask)

    @staticmethod
    def sqrt_swap(ab_mask):
        return _new(lib.qvnt_op_sqrt_swap, ab_mask)

    @staticmethod
    def i_swap(ab_mask):
        return _new(lib.qvnt_op_i_swap, ab_mask)

    @staticmethod
    def sqrt_i_swap(ab_mask):
        return _new(lib.qvnt_op_sqrt_i_swap, ab_mask)

    @staticmethod
    def h(a_mask):
        return _new(lib.qvnt_op_h, a_mask)

    @staticmethod
    def u1(lam, a_mask):
        return _new(lib.qvnt_op_u1, lam, a_mask)

    @staticmethod
    def u2(phi, lam, a_mask):
        return _new(lib.qvnt_op_u2, phi, lam, a_mask)

    @staticmethod
    def u3(the, phi, lam, a_mask):
        return _new(lib.qvnt_op_u3, the, phi, lam, a_mask)

    @staticmethod
    def qft(a_mask):
        return _new(lib.qvnt_op_qft, a_mask)

    @staticmethod
    def qft_swapped(a_mask):
        return _new(lib.qvnt_op_qft_swapped, a_mask)

    @staticmethod
    def matrix1(a_mask, m):
        """Checked 2x2 matrix gate (single/pauli.rs:32-34); None if invalid."""
        arr = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(4)).view(np.float64)
        try:
            return _new(lib.qvnt_op_matrix1, a_mask, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        except ValueError:
            return None

    @staticmethod
    def matrix2(a_mask, b_mask, m):
        """Checked 4x4 matrix gate (single/pauli.rs:37-39); None if invalid."""
        arr = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(16)).view(np.float64)
        try:
            return _new(lib.qvnt_op_matrix2, a_mask, b_mask, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        except ValueError:
            return None


op = _OpNamespace()

__all__ = ["MultiOp", "op", "from_gates", "Gate", "QvntError"]
