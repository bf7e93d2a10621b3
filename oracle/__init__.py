"""ctypes loader for the CPU oracle (oracle/qvnt_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it as the checker.  The product
package ``qvnt_amd`` never imports it.

The oracle restates qvnt 0.4.4's CPU algorithm (src/operator/atomic/*,
src/operator/{single,multi}, src/register/quant.rs); see qvnt_oracle.c for
the per-function citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libqvnt_oracle.so")

# status codes (include/qvnt.h)
E_INVALID_MASK = -1
E_CTRL_OVERLAP = -2
E_ZERO_NORM = -5
MEASURE_RENORM = 1

# gate kinds (include/qvnt.h, enum qvnt_kind)
KINDS = ["ID", "X", "RX", "RXX", "Y", "RY", "RYY", "Z", "S", "T", "RZ", "RZZ",
         "U1", "U2", "H1", "H2", "SWAP", "ISWAP", "SQRT_SWAP", "SQRT_ISWAP", "PHI"]
KIND = {k: i for i, k in enumerate(KINDS)}


class Gate(ctypes.Structure):
    """Mirror of ``qvnt_gate`` (include/qvnt.h)."""
    _fields_ = [("kind", ctypes.c_uint32), ("dagger", ctypes.c_uint32),
                ("act", ctypes.c_uint64), ("a", ctypes.c_uint64),
                ("b", ctypes.c_uint64), ("ctrl", ctypes.c_uint64),
                ("re", ctypes.c_double), ("im", ctypes.c_double),
                ("m", ctypes.c_double * 32)]


assert ctypes.sizeof(Gate) == 312

SEP_NOP, SEP_MEASURE, SEP_IF, SEP_RESET = 0, 1, 2, 3


class Step(ctypes.Structure):
    """Mirror of ``orc_step``: one (op, separator) entry of qasm's ExtOp."""
    _fields_ = [("g", ctypes.POINTER(Gate)), ("n", ctypes.c_size_t), ("sep", ctypes.c_uint32),
                ("pad", ctypes.c_uint32), ("a", ctypes.c_uint64), ("b", ctypes.c_uint64)]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, d, gp = ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(Gate)
        for name in ["x", "y", "z", "s", "t", "h", "qft", "qft_swapped",
                     "swap", "sqrt_swap", "i_swap", "sqrt_i_swap"]:
            f = getattr(L, "orc_op_" + name)
            f.argtypes, f.restype = [u64, gp], ctypes.c_long
        for name in ["rx", "ry", "rz", "rxx", "ryy", "rzz", "phi", "u1"]:
            f = getattr(L, "orc_op_" + name)
            f.argtypes, f.restype = [d, u64, gp], ctypes.c_long
        L.orc_op_id.argtypes, L.orc_op_id.restype = [gp], ctypes.c_long
        L.orc_op_u2.argtypes, L.orc_op_u2.restype = [d, d, u64, gp], ctypes.c_long
        L.orc_op_u3.argtypes, L.orc_op_u3.restype = [d, d, d, u64, gp], ctypes.c_long
        L.orc_op_matrix1.argtypes = [u64, ctypes.POINTER(d), gp]
        L.orc_op_matrix1.restype = ctypes.c_long
        L.orc_op_matrix2.argtypes = [u64, u64, ctypes.POINTER(d), gp]
        L.orc_op_matrix2.restype = ctypes.c_long
        L.orc_op_c.argtypes, L.orc_op_c.restype = [gp, ctypes.c_size_t, u64], ctypes.c_int
        L.orc_op_dgr.argtypes, L.orc_op_dgr.restype = [gp, ctypes.c_size_t], None
        L.orc_op_act_on.argtypes, L.orc_op_act_on.restype = [gp, ctypes.c_size_t], u64
        L.orc_reg_with_state.argtypes, L.orc_reg_with_state.restype = [ctypes.c_uint32, u64], ctypes.c_void_p
        L.orc_reg_free.argtypes, L.orc_reg_free.restype = [ctypes.c_void_p], None
        L.orc_reg_psi.argtypes, L.orc_reg_psi.restype = [ctypes.c_void_p], ctypes.POINTER(d)
        L.orc_reg_len.argtypes, L.orc_reg_len.restype = [ctypes.c_void_p], u64
        L.orc_reg_num_threads.argtypes, L.orc_reg_num_threads.restype = [ctypes.c_void_p, ctypes.c_int], ctypes.c_int
        L.orc_reg_apply.argtypes, L.orc_reg_apply.restype = [ctypes.c_void_p, gp, ctypes.c_size_t], None
        L.orc_reg_get_absolute.argtypes, L.orc_reg_get_absolute.restype = [ctypes.c_void_p], d
        L.orc_reg_get_probabilities.argtypes = [ctypes.c_void_p, ctypes.POINTER(d)]
        L.orc_reg_get_probabilities.restype = None
        L.orc_reg_measure_mask.argtypes = [ctypes.c_void_p, u64, d, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]
        L.orc_reg_measure_mask.restype = u64
        L.orc_tree_total.argtypes, L.orc_tree_total.restype = [ctypes.POINTER(d), u64], d
        L.orc_tree_sample.argtypes = [ctypes.POINTER(d), u64, d, ctypes.POINTER(ctypes.c_int)]
        L.orc_tree_sample.restype = u64
        L.orc_splitmix64.argtypes, L.orc_splitmix64.restype = [ctypes.POINTER(u64)], u64
        L.orc_uniform.argtypes, L.orc_uniform.restype = [u64, u64, u64], d
        vp, pd = ctypes.c_void_p, ctypes.POINTER(d)
        L.orc_reg_reset.argtypes, L.orc_reg_reset.restype = [vp, u64], None
        L.orc_reg_reset_by_mask.argtypes, L.orc_reg_reset_by_mask.restype = [vp, u64], None
        L.orc_reg_tensor.argtypes, L.orc_reg_tensor.restype = [vp, vp], vp
        L.orc_reg_combine.argtypes, L.orc_reg_combine.restype = [vp, vp], vp
        L.orc_reg_combine_with_unitary.argtypes, L.orc_reg_combine_with_unitary.restype = [vp, vp, pd], vp
        L.orc_reg_linear_composition.argtypes, L.orc_reg_linear_composition.restype = [vp, vp, pd], ctypes.c_int
        L.orc_reg_set_num.argtypes, L.orc_reg_set_num.restype = [vp, ctypes.c_uint32], ctypes.c_int
        L.orc_reg_get_polar.argtypes, L.orc_reg_get_polar.restype = [vp, pd], None
        L.orc_reg_sample_all.argtypes = [vp, u64, u64, u64, ctypes.POINTER(u64)]
        L.orc_reg_sample_all.restype = ctypes.c_int
        L.orc_run_program.argtypes = [vp, ctypes.POINTER(Step), ctypes.c_size_t, gp, ctypes.c_size_t, ctypes.c_uint32,
                                      ctypes.POINTER(u64), ctypes.c_uint32, pd, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t)]
        L.orc_run_program.restype = ctypes.c_int
        _lib = L
    return _lib


class Op:
    """A reference MultiOp restated as an array of gates (oracle side)."""

    def __init__(self, gates=()):
        self.gates = list(gates)

    def _arr(self):
        return (Gate * max(1, len(self.gates)))(*self.gates)

    def __mul__(self, other):
        return Op(self.gates + other.gates)

    def c(self, mask):
        arr = self._arr()
        if lib().orc_op_c(arr, len(self.gates), mask) != 0:
            return None
        return Op(list(arr)[:len(self.gates)])

    def dgr(self):
        arr = self._arr()
        lib().orc_op_dgr(arr, len(self.gates))
        return Op(list(arr)[:len(self.gates)])

    def act_on(self):
        return lib().orc_op_act_on(self._arr(), len(self.gates))

    def __len__(self):
        return len(self.gates)


def _make(fn, *args, cap=4096):
    buf = (Gate * cap)()
    n = fn(*args, buf)
    if n < 0:
        raise ValueError("oracle builder rejected mask (reference panics): %d" % n)
    return Op(list(buf)[:n])


class _OpNS:
    """op:: constructors with the reference argument order."""

    def id(self):
        return _make(lib().orc_op_id)

    def __getattr__(self, name):
        f = getattr(lib(), "orc_op_" + name)
        return lambda *args: _make(f, *args)

    def matrix1(self, a, m):
        arr = (ctypes.c_double * 8)(*np.asarray(m, np.complex128).view(np.float64).ravel())
        return _make(lib().orc_op_matrix1, a, arr)

    def matrix2(self, a, b, m):
        arr = (ctypes.c_double * 32)(*np.asarray(m, np.complex128).view(np.float64).ravel())
        return _make(lib().orc_op_matrix2, a, b, arr)


op = _OpNS()


class Reg:
    """QReg restated: out-of-place gather engine on host memory."""

    def __init__(self, q_num, state=0, threads=1):
        self.h = lib().orc_reg_with_state(q_num, state)
        if not self.h:
            raise MemoryError("oracle register allocation failed")
        self.q_num = q_num
        if threads != 1 and lib().orc_reg_num_threads(self.h, threads) != 0:
            raise ValueError("num_threads out of range")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_reg_free(self.h)
            self.h = None

    def __len__(self):
        return lib().orc_reg_len(self.h)

    def apply(self, o):
        lib().orc_reg_apply(self.h, o._arr(), len(o.gates))

    def state(self):
        n = len(self)
        p = lib().orc_reg_psi(self.h)
        return np.ctypeslib.as_array(p, shape=(2 * n,)).view(np.complex128).copy()

    def set_state(self, psi):
        psi = np.ascontiguousarray(psi, np.complex128)
        assert psi.shape == (len(self),)
        p = lib().orc_reg_psi(self.h)
        ctypes.memmove(p, psi.ctypes.data, psi.nbytes)

    def get_absolute(self):
        return lib().orc_reg_get_absolute(self.h)

    def get_probabilities(self):
        out = np.empty(1 << self.q_num, np.float64)
        lib().orc_reg_get_probabilities(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def measure_mask(self, mask, u, renorm=False):
        err = ctypes.c_int(0)
        r = lib().orc_reg_measure_mask(self.h, mask, u, MEASURE_RENORM if renorm else 0, ctypes.byref(err))
        if err.value:
            raise ZeroDivisionError("measure on a zero-norm state")
        return r

    @classmethod
    def _wrap(cls, h, q_num):
        if not h:
            return None
        r = cls.__new__(cls)
        r.h, r.q_num = h, q_num
        return r

    def reset(self, i_state=0):
        lib().orc_reg_reset(self.h, i_state)

    def reset_by_mask(self, mask):
        lib().orc_reg_reset_by_mask(self.h, mask)

    def __mul__(self, other):
        """Tensor product: self on the low qubits (quant.rs:330-371)."""
        return Reg._wrap(lib().orc_reg_tensor(self.h, other.h), self.q_num + other.q_num)

    def combine(self, other):
        return Reg._wrap(lib().orc_reg_combine(self.h, other.h), self.q_num + 1)

    def combine_with_unitary(self, other, m):
        c = np.ascontiguousarray(np.asarray(m, np.complex128).reshape(4).view(np.float64))
        return Reg._wrap(lib().orc_reg_combine_with_unitary(self.h, other.h,
                                                            c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                         self.q_num + 1)

    def linear_composition(self, other, c0, c1):
        c = np.array([c0.real, c0.imag, c1.real, c1.imag], np.float64)
        if lib().orc_reg_linear_composition(self.h, other.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise ValueError("buffers of different length")

    def set_num(self, q_num):
        if lib().orc_reg_set_num(self.h, q_num):
            raise MemoryError("set_num")
        self.q_num = q_num

    def get_polar(self):
        out = np.empty(2 << self.q_num, np.float64)
        lib().orc_reg_get_polar(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out.reshape(-1, 2)

    def sample_all(self, shots, seed, stream):
        out = np.zeros(len(self), np.uint64)
        if lib().orc_reg_sample_all(self.h, shots, seed, stream, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))):
            raise ZeroDivisionError("zero-norm state")
        return out

    def run_program(self, steps, tail, c_num, c_value=0, measure_xor=False, draws=()):
        """Sym::finish (qasm/sym.rs:40-74).  steps: (Op, sep, a, b) tuples;
        returns (classical value, draws used)."""
        keep = [st[0]._arr() for st in steps]
        arr = (Step * max(1, len(steps)))()
        for k, (o, sep, a, b) in enumerate(steps):
            arr[k].g = ctypes.cast(keep[k], ctypes.POINTER(Gate))
            arr[k].n = len(o.gates)
            arr[k].sep, arr[k].a, arr[k].b = sep, a, b
        d = np.ascontiguousarray(draws, np.float64)
        cval = ctypes.c_uint64(c_value)
        used = ctypes.c_size_t(0)
        rc = lib().orc_run_program(self.h, arr, len(steps), tail._arr(), len(tail.gates), 1 if measure_xor else 0,
                                   ctypes.byref(cval), c_num, d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   len(d), ctypes.byref(used))
        if rc:
            raise RuntimeError("program failed with status %d" % rc)
        return cval.value, used.value


def tree_total(psi, count=None):
    psi = np.ascontiguousarray(psi, np.complex128)
    count = len(psi) if count is None else count
    return lib().orc_tree_total(psi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), count)


def tree_sample(psi, u, count=None):
    psi = np.ascontiguousarray(psi, np.complex128)
    count = len(psi) if count is None else count
    err = ctypes.c_int(0)
    r = lib().orc_tree_sample(psi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), count, u, ctypes.byref(err))
    if err.value:
        raise ZeroDivisionError("zero-norm state")
    return r


def uniform(seed, stream, index):
    return lib().orc_uniform(seed, stream, index)


def splitmix64_sequence(seed, count):
    st = ctypes.c_uint64(seed)
    return [lib().orc_splitmix64(ctypes.byref(st)) for _ in range(count)]
