"""ctypes binding of libqvnt_hip.so (the C ABI declared in include/qvnt.h).

The library is built in-tree (``python qvnt_amd/build.py``).  There is no
fallback: if the library is missing or cannot be loaded, importing the
package raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqvnt_hip.so")

OK = 0
E_INVALID_MASK = -1
E_CTRL_OVERLAP = -2
E_OOM = -3
E_HIP = -4
E_ZERO_NORM = -5
E_ARG = -6
E_NO_DEVICE = -7

F64, F32 = 0, 1
MEASURE_RENORM = 1

KINDS = ["ID", "X", "RX", "RXX", "Y", "RY", "RYY", "Z", "S", "T", "RZ", "RZZ",
         "U1", "U2", "H1", "H2", "SWAP", "ISWAP", "SQRT_SWAP", "SQRT_ISWAP", "PHI"]
KIND = {k: i for i, k in enumerate(KINDS)}
KCLASSES = ["diag", "pair", "quad", "fused", "reduce", "collapse", "other"]


class Gate(ctypes.Structure):
    """``qvnt_gate``: one reference SingleOp (src/operator/single/mod.rs:42-47)."""
    _fields_ = [("kind", ctypes.c_uint32), ("dagger", ctypes.c_uint32),
                ("act", ctypes.c_uint64), ("a", ctypes.c_uint64),
                ("b", ctypes.c_uint64), ("ctrl", ctypes.c_uint64),
                ("re", ctypes.c_double), ("im", ctypes.c_double),
                ("m", ctypes.c_double * 32)]


class KStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("ms", ctypes.c_double), ("bytes", ctypes.c_double)]


assert ctypes.sizeof(Gate) == 312


class QvntError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("qvnt error %d: %s" % (code, msg))
        self.code = code


def _declare(L):
    u32, u64, i, d = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    sz, vp, cp = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p
    opp = ctypes.POINTER(ctypes.c_void_p)
    gp = ctypes.POINTER(Gate)
    dp = ctypes.POINTER(d)
    u64p = ctypes.POINTER(u64)
    sig = {
        "qvnt_last_error": ([], cp),
        "qvnt_version": ([], cp),
        "qvnt_uniform": ([u64, u64, u64], d),
        "qvnt_op_id": ([opp], i),
        "qvnt_op_h": ([u64, opp], i),
        "qvnt_op_u1": ([d, u64, opp], i),
        "qvnt_op_u2": ([d, d, u64, opp], i),
        "qvnt_op_u3": ([d, d, d, u64, opp], i),
        "qvnt_op_matrix1": ([u64, dp, opp], i),
        "qvnt_op_matrix2": ([u64, u64, dp, opp], i),
        "qvnt_op_from_gates": ([gp, sz, opp], i),
        "qvnt_op_clone": ([vp, opp], i),
        "qvnt_op_free": ([vp], None),
        "qvnt_op_mul": ([vp, vp], i),
        "qvnt_op_c": ([vp, u64], i),
        "qvnt_op_dgr": ([vp], i),
        "qvnt_op_act_on": ([vp], u64),
        "qvnt_op_len": ([vp], sz),
        "qvnt_op_gates": ([vp], gp),
        "qvnt_op_name": ([vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)], i),
        "qvnt_reg_new": ([u32, i, opp], i),
        "qvnt_reg_with_state": ([u32, u64, i, opp], i),
        "qvnt_reg_free": ([vp], i),
        "qvnt_cache_trim": ([], i),
        "qvnt_reg_num": ([vp], u32),
        "qvnt_reg_len": ([vp], u64),
        "qvnt_reg_prec": ([vp], i),
        "qvnt_reg_apply": ([vp, vp], i),
        "qvnt_reg_apply_gates": ([vp, gp, sz], i),
        "qvnt_reg_apply_fused": ([vp, vp, u32, ctypes.POINTER(u32)], i),
        "qvnt_fused_plan_stats": ([vp, u32, i, u32, u32, dp], i),
        "qvnt_reg_get_absolute": ([vp, dp], i),
        "qvnt_reg_get_probabilities": ([vp, dp, u64], i),
        "qvnt_reg_measure_mask": ([vp, u64, d, u32, u64p], i),
        "qvnt_reg_measure": ([vp, d, u32, u64p], i),
        "qvnt_reg_sample_mask": ([vp, u64, dp, sz, u64p], i),
        "qvnt_reg_sample_mask_noalloc": ([vp, u64, dp, sz, u64p], i),
        "qvnt_reg_compare": ([vp, vp, dp], i),
        "qvnt_reg_get_state": ([vp, vp, sz], i),
        "qvnt_reg_set_state": ([vp, vp, sz], i),
        "qvnt_reg_sync": ([vp], i),
        "qvnt_reg_debug": ([vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)], i),
        "qvnt_reg_set_timing": ([vp, i], i),
        "qvnt_reg_set_dense": ([vp, u32], i),
        "qvnt_reg_kernel_stats": ([vp, ctypes.POINTER(KStats), i], i),
        "qvnt_reg_reset": ([vp, u64], i),
        "qvnt_reg_reset_by_mask": ([vp, u64], i),
        "qvnt_reg_tensor": ([vp, vp, opp], i),
        "qvnt_reg_combine": ([vp, vp, opp], i),
        "qvnt_reg_combine_with_unitary": ([vp, vp, dp, opp], i),
        "qvnt_reg_linear_composition": ([vp, vp, dp], i),
        "qvnt_reg_set_num": ([vp, u32], i),
        "qvnt_reg_get_polar": ([vp, dp, u64], i),
        "qvnt_reg_sample_all": ([vp, u64, u64, u64, u64p, u64], i),
        "qvnt_reg_run_program": ([vp, ctypes.POINTER(Step), sz, vp, u32, u64p, u32, dp, sz, ctypes.POINTER(sz)], i),
    }
    for name in ["x", "y", "z", "s", "t", "swap", "sqrt_swap", "i_swap", "sqrt_i_swap", "qft", "qft_swapped"]:
        sig["qvnt_op_" + name] = ([u64, opp], i)
    for name in ["rx", "ry", "rz", "rxx", "ryy", "rzz", "phi"]:
        sig["qvnt_op_" + name] = ([d, u64, opp], i)
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    return L


SEP_NOP, SEP_MEASURE, SEP_IF, SEP_RESET = 0, 1, 2, 3


class Step(ctypes.Structure):
    """Mirror of ``qvnt_step`` (one (op, separator) entry of the QASM driver)."""
    _fields_ = [("op", ctypes.c_void_p), ("sep", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("a", ctypes.c_uint64), ("b", ctypes.c_uint64)]


def load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("qvnt_amd: %s is not built (run `python qvnt_amd/build.py`); "
                          "there is no CPU fallback" % LIB_PATH)
    return _declare(ctypes.CDLL(LIB_PATH))


lib = load()


def last_error():
    return lib.qvnt_last_error().decode()


def check(rc):
    if rc != OK:
        raise QvntError(rc, last_error())
    return rc
