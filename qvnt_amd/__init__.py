"""qvnt_amd — an MI355X state-vector engine with qvnt 0.4.4's register and
operator API (op::, MultiOp, QReg, CReg), backed by hand-written HIP kernels
for gfx950 behind the C ABI in include/qvnt.h.

The package loads ``qvnt_amd/lib/libqvnt_hip.so``; there is no CPU path.
"""
from ._lib import QvntError, Gate, KIND, KINDS, lib, last_error
from .ops import MultiOp, from_gates, op
from .register import CReg, QReg, VReg, cache_trim, fused_plan_stats, uniform
from .program import Program, Sym
from . import qasm

__all__ = ["QvntError", "Gate", "KIND", "KINDS", "MultiOp", "from_gates", "op",
           "CReg", "QReg", "VReg", "Program", "Sym", "qasm", "fused_plan_stats", "cache_trim", "uniform", "lib",
           "last_error"]
