"""Python side of the batched-register form (include/qvnt.h, qvnt_batch_*):
many independent registers of 6..30 qubits, each with its own circuit,
compiled into fused passes and run round by round on one stream."""
import ctypes

import numpy as np

from ._lib import F32, F64, check, lib
from .ops import MultiOp


class BatchInfo(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint32), ("launches", ctypes.c_uint32), ("passes", ctypes.c_uint64),
                ("tiles", ctypes.c_uint64), ("single_ops", ctypes.c_uint64), ("bytes", ctypes.c_double),
                ("phases", ctypes.c_uint64), ("dense_phases", ctypes.c_uint64), ("dense_flops", ctypes.c_double),
                ("tile_cap", ctypes.c_uint32 * 32), ("dense_min", ctypes.c_uint32 * 32),
                ("tune_caps", ctypes.c_uint32 * 3), ("persist_blocks", ctypes.c_uint32),
                ("tune_ms", (ctypes.c_double * 3) * 3), ("single_ms", ctypes.c_double * 32)]


TUNE_CAPS = (10, 11, 12)          # FP64; FP32 tries 11, 12, 13
TUNE_DENSE = (0, 3, 4)


def _declare():
    vp, u32, u64, i, d = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "qvnt_batch_new": ([u32, P(u32), P(u64), u32, P(vp)], i),
        "qvnt_batch_count": ([vp], u32),
        "qvnt_batch_reset": ([vp], i),
        "qvnt_batch_compile": ([vp, P(vp), u32, u32, P(BatchInfo)], i),
        "qvnt_batch_run": ([vp, i], i),
        "qvnt_batch_persist_form": ([vp, P(i)], i),
        "qvnt_batch_time": ([vp, u32, i, P(d)], i),
        "qvnt_batch_sync": ([vp], i),
        "qvnt_batch_get_state": ([vp, u32, vp, ctypes.c_size_t], i),
        "qvnt_batch_sample_mask": ([vp, u32, u64, P(d), ctypes.c_size_t, P(u64), P(d)], i),
        "qvnt_batch_free": ([vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = args, res


_declare()


class Batch:
    """Registers k = 0..len(q_nums)-1 with q_nums[k] qubits in basis state
    states[k] (default 0)."""

    def __init__(self, q_nums, states=None, prec="f64"):
        q = np.ascontiguousarray(q_nums, np.uint32)
        st = None if states is None else np.ascontiguousarray(states, np.uint64)
        self.q_nums = q.tolist()
        self.prec = prec
        h = ctypes.c_void_p()
        check(lib.qvnt_batch_new(len(q), q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                 None if st is None else st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                 {"f64": F64, "f32": F32}[prec], ctypes.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.q_nums)

    def compile(self, ops, max_qubits=0, dense_min=0):
        """Compiles ops[k] for register k with tiles of at most max_qubits
        qubits and dense phases of >= dense_min ops (0: none); max_qubits = 0
        runs the autotuner instead (the registers are reset afterwards).
        Returns the schedule summary."""
        if len(ops) != len(self.q_nums) or not all(isinstance(o, MultiOp) for o in ops):
            raise TypeError("compile expects one MultiOp per register")
        arr = (ctypes.c_void_p * len(ops))(*[o.handle for o in ops])
        info = BatchInfo()
        check(lib.qvnt_batch_compile(self._h, arr, max_qubits, dense_min, ctypes.byref(info)))
        sizes = sorted(set(self.q_nums))
        tune = {"cap%d_dense%d" % (info.tune_caps[i], d): round(info.tune_ms[i][j], 3)
                for i in range(3) for j, d in enumerate(TUNE_DENSE) if info.tune_ms[i][j] > 0}
        single = {n: round(info.single_ms[n], 3) for n in sizes if info.single_ms[n] > 0}
        return {"rounds": info.rounds, "launches": info.launches, "passes": info.passes, "tiles": info.tiles,
                "single_ops": info.single_ops, "bytes": info.bytes, "phases": info.phases,
                "dense_phases": info.dense_phases, "dense_flops": info.dense_flops,
                "persist_blocks": info.persist_blocks,
                "tile_cap": {n: info.tile_cap[n] for n in sizes}, "dense_min": {n: info.dense_min[n] for n in sizes},
                "tune_ms": tune, "single_ms": single}

    def reset(self):
        check(lib.qvnt_batch_reset(self._h))

    @staticmethod
    def _mode(graph):
        return 2 if graph == "persistent" else (1 if graph else 0)

    def run(self, graph=False):
        """One application of every compiled circuit: graph False -- the
        rounds as launches; True -- the same as one replayed hipGraph;
        "persistent" -- one persistent cooperative launch over all rounds,
        captured into a hipGraph and replayed (see persist_form)."""
        check(lib.qvnt_batch_run(self._h, self._mode(graph)))

    @property
    def persist_form(self):
        """How run(graph="persistent") runs: None (not yet since the last
        compile), "graph" (the cooperative launch captured into a hipGraph)
        or "direct" (hipLaunchCooperativeKernel)."""
        f = ctypes.c_int()
        check(lib.qvnt_batch_persist_form(self._h, ctypes.byref(f)))
        return {0: None, 1: "graph", 2: "direct"}[f.value]

    def time(self, runs, graph=False):
        ms = ctypes.c_double()
        check(lib.qvnt_batch_time(self._h, runs, self._mode(graph), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(lib.qvnt_batch_sync(self._h))

    def get_state(self, k):
        dt = np.complex128 if self.prec == "f64" else np.complex64
        out = np.empty(1 << self.q_nums[k], dt)
        check(lib.qvnt_batch_get_state(self._h, k, out.ctypes.data, out.nbytes))
        return out

    def sample_mask(self, k, mask, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(len(u), np.uint64)
        total = ctypes.c_double()
        check(lib.qvnt_batch_sample_mask(self._h, k, mask, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(u),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(total)))
        return out, total.value

    def close(self):
        if getattr(self, "_h", None):
            lib.qvnt_batch_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
