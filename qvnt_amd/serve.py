"""Python side of the continuous-batching request server (include/qvnt.h,
qvnt_server_*): submit circuits with their shots, poll finished results."""
import ctypes

import numpy as np

from ._lib import F32, F64, check, lib
from .ops import MultiOp


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("q_num", ctypes.c_uint32), ("prec", ctypes.c_uint32),
                ("init_state", ctypes.c_uint64), ("op", ctypes.c_void_p), ("fused", ctypes.c_uint32),
                ("max_qubits", ctypes.c_uint32), ("dense_min", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("mask", ctypes.c_uint64), ("shots", ctypes.c_uint64),
                ("draws", ctypes.POINTER(ctypes.c_double)), ("outcomes", ctypes.POINTER(ctypes.c_uint64))]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("passes", ctypes.c_uint32),
                ("total", ctypes.c_double), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


def _declare():
    vp, u32, u64, i, d = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    lib.qvnt_server_new.argtypes, lib.qvnt_server_new.restype = [u64, u32, ctypes.POINTER(vp)], i
    lib.qvnt_server_submit.argtypes, lib.qvnt_server_submit.restype = [vp, ctypes.POINTER(Request)], i
    lib.qvnt_server_poll.argtypes = [vp, ctypes.POINTER(Result), u32, ctypes.POINTER(u32), i]
    lib.qvnt_server_poll.restype = i
    lib.qvnt_server_pending.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u64)]
    lib.qvnt_server_pending.restype = i
    lib.qvnt_server_now.argtypes, lib.qvnt_server_now.restype = [vp], d
    lib.qvnt_server_free.argtypes, lib.qvnt_server_free.restype = [vp], i


_declare()


class Server:
    """Requests are admitted as slots and arena space free up; outcomes
    equal those of the same request run alone."""

    def __init__(self, hbm_bytes, slots=4):
        h = ctypes.c_void_p()
        check(lib.qvnt_server_new(hbm_bytes, slots, ctypes.byref(h)))
        self._h = h
        self._live = {}  # id -> (draws, outcomes, op): kept alive until polled

    def submit(self, rid, q_num, op, mask, draws, fused=True, prec="f64", init_state=0, max_qubits=0, dense_min=0):
        if not isinstance(op, MultiOp):
            raise TypeError("submit expects a MultiOp")
        if rid in self._live:
            # the first request's outcome buffer must stay alive until it is polled
            raise ValueError("request id %d is already in flight" % rid)
        draws = np.ascontiguousarray(draws, np.float64)
        outs = np.zeros(len(draws), np.uint64)
        req = Request(rid, q_num, {"f64": F64, "f32": F32}[prec], init_state, op.handle, 1 if fused else 0,
                      max_qubits, dense_min, 0, mask, len(draws), draws.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      outs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        self._live[rid] = (draws, outs, op)
        try:
            check(lib.qvnt_server_submit(self._h, ctypes.byref(req)))
        except Exception:
            del self._live[rid]
            raise

    def poll(self, wait_ms=0, cap=512):
        arr = (Result * cap)()
        n = ctypes.c_uint32(0)
        check(lib.qvnt_server_poll(self._h, arr, cap, ctypes.byref(n), wait_ms))
        out = []
        for r in arr[:n.value]:
            draws, outs, _ = self._live.pop(r.id)
            out.append({"id": r.id, "status": r.status, "passes": r.passes, "total": r.total,
                        "t_submit": r.t_submit, "t_admit": r.t_admit, "t_done": r.t_done, "outcomes": outs})
        return out

    def drain(self, count, timeout_ms=600000):
        """Polls until `count` results have arrived."""
        res = []
        waited = 0
        while len(res) < count and waited < timeout_ms:
            got = self.poll(wait_ms=50)
            res += got
            waited = 0 if got else waited + 50
        return res

    def now(self):
        return lib.qvnt_server_now(self._h)

    def pending(self):
        q, r, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib.qvnt_server_pending(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(p)))
        return q.value, r.value, p.value

    def close(self):
        if getattr(self, "_h", None):
            lib.qvnt_server_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
