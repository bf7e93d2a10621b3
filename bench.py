#!/usr/bin/env python3
"""Benchmark: gates/s and % of HBM roofline of the plain FP64 apply on the
30-qubit brickwork random circuit (16 GiB state), one MI355X, with the
fused form on the same layers beside it (key "fused"), and the rest of the
metric in the same JSON line.

A step is one brickwork layer: u3 on every qubit (3 SingleOps each), rzz on
alternating nearest-neighbour pairs, one controlled x -- 105 or 106
reference SingleOps.  --steps K layers of the depth-40 circuit are timed
after --warmup W layers.  `value` counts reference SingleOps (u3 = 3), the
unit the plain form launches one kernel for.

The roofline object reports the kernel class that dominates device time,
from HIP events recorded on the register's stream around every launch in
the timed region: achieved = algorithmic bytes 2 * 16 B * 2^(n - #controls)
per launch / average launch time.  The CPU baseline times the oracle (the C
restatement of qvnt's out-of-place gather with its OpenMP twin of the Rayon
loop) on the host on a bounded sample of the same workload, plus the
20-qubit case and the reference's own benchmark workload.

Further keys of the default run (skipped with --no-extra), each leg freeing
its device memory before the next:
  tfim       TFIM ring, 34 qubits FP64 and 35 qubits FP32 (256 GiB each):
             one plain Trotter step with its dominant kernel's roofline,
             4 steps through the fused form one call per step, and 50 steps
             as one call;
  tfim_pair  28 qubits, FP32 against FP64 after every one of 50 steps
             (fused and plain): max |dpsi| and infidelity;
  batch      4096 registers of 10-20 qubits, FP64 and FP32, autotuned and
             run as one hipGraph: gates/s, ms per run, HBM and matrix-core
             rates (the plain register-by-register loop beside FP64);
  served     256 serving requests (12-30 qubits, 100-2000 gates, 1024
             shots): one-at-a-time fused loop (and plain loop on the first
             128), the continuous-batching server at saturation, and
             open-loop Poisson arrivals at 50/80/100 % of the serial fused
             capacity into a one-slot and a four-slot server (p50/p99
             latency).
"""
import argparse
import json
import os
import sys
import time

# The CPU baseline's OpenMP threads, one per core; set before the oracle's
# OpenMP runtime starts.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}   # MI355X dense matrix-core peaks
METRIC = "gates/s + % HBM roofline at 30q FP64; served circuits/s at p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baseline runs on: lscpu model, sockets, physical
    cores, NUMA nodes; the CPUs this process may use; OpenMP pinning."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        sockets = int(kv.get("Socket(s)", "1") or 1)
        cores = int(kv.get("Core(s) per socket", "0") or 0)
        info = {"model": kv.get("Model name"), "sockets": sockets, "cores_per_socket": cores,
                "physical_cores": sockets * cores, "threads_per_core": int(kv.get("Thread(s) per core", "1") or 1),
                "numa_nodes": int(kv.get("NUMA node(s)", "1") or 1)}
    except (OSError, ValueError, subprocess.SubprocessError) as e:
        info = {"lscpu_error": str(e)}
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    info["omp_num_threads_env"] = os.environ.get("OMP_NUM_THREADS")
    info["pinning"] = "OMP_PROC_BIND=%s OMP_PLACES=%s" % (os.environ.get("OMP_PROC_BIND"), os.environ.get("OMP_PLACES"))
    return info


def cpu_threads(host):
    """Threads for the CPU baseline: one per physical core this process may
    use, capped by OMP_NUM_THREADS when the environment sets it (a GPU lease
    grants a share of the host's cores: 16 per GPU on the MI355X boxes)."""
    n = min(host.get("affinity_cpus") or 1, host.get("physical_cores") or host.get("affinity_cpus") or 1)
    env = host.get("omp_num_threads_env")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(n, threads, ks=(1, 4)):
    """The oracle (qvnt's algorithm in C, OpenMP twin of the Rayon loop) on
    the brickwork at full size: the first k SingleOps of layer 0 as one
    MultiOp::apply, for each k in ks.  Each apply pays the reference's
    per-apply costs (a fresh output buffer and a copy of the input,
    multi/mod.rs:98, quant.rs:382-383) plus one out-of-place gather per
    SingleOp, so time = a + b k is fitted to separate the two; the layer rate
    is 106 / (a + 106 b), the reference applying a whole layer as one
    MultiOp."""
    import oracle
    from qvnt_amd import circuits
    specs = circuits.brickwork_layer(n, 0)
    full = circuits.build(specs, oracle.op)
    reg = oracle.Reg(n, 0, threads=threads)
    # untimed: the first apply also pays the first touch of the register's
    # own buffer, which the reference's later applies do not
    reg.apply(oracle.Op(full.gates[:1]))
    runs = []
    for k in ks:
        part = oracle.Op(full.gates[:k])
        t0 = time.perf_counter()
        reg.apply(part)
        runs.append((k, time.perf_counter() - t0))
    del reg
    done = sum(k for k, _ in runs)
    elapsed = sum(t for _, t in runs)
    (k0, t0_), (k1, t1_) = runs[0], runs[-1]
    b = (t1_ - t0_) / (k1 - k0) if k1 != k0 else t1_ / k1
    a = max(0.0, t0_ - b * k0)
    layer = len(full.gates)
    return {"value": done / elapsed, "unit": "gates/s", "cores": threads, "kind": "port",
            "sample": "first %s SingleOps of brickwork layer 0 at %d qubits, one MultiOp::apply each after one "
                      "untimed apply (FP64, out-of-place gather + per-apply copy), %.1f s"
                      % ("/".join(str(k) for k, _ in runs), n, elapsed),
            "apply_seconds": {str(k): round(t, 3) for k, t in runs},
            "per_singleop_s": round(b, 4), "per_apply_overhead_s": round(a, 3),
            "layer_as_one_apply_gates_per_s": round(layer / (a + layer * b), 4) if a + layer * b > 0 else None}


def reference_bench_leg(threads, cpu=True):
    """The reference's own benchmark (benches/performance.rs:4-20): per
    iteration QReg::with_state(q, 0), apply(qft(0b0111) * qft(0b1110)),
    measure_mask(0b100), for q = 18, 19, 20 -- on the GPU (plain form,
    register created and freed every iteration, as the reference allocates
    its Vec) and, with `cpu`, on the oracle on 1 and `threads` host threads.
    Iterations per second; each side runs for about half a second, the GPU
    side after 20 untimed iterations (code objects loaded, a stream and the
    register's blocks cached)."""
    from qvnt_amd import QReg, op, uniform
    out = {}
    mask = 0b100
    for q in (18, 19, 20):
        gops = op.qft(0b0111) * op.qft(0b1110)
        row = {}
        for it in range(20):
            with QReg.with_state(q, 0) as r:
                r.apply(gops)
                r.measure_mask(mask, u=uniform(1, q, it)).get()
        it, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 0.5 or it < 3:
            with QReg.with_state(q, 0) as r:
                r.apply(gops)
                got = r.measure_mask(mask, u=uniform(0x51564E54, q, it)).get()
            assert got & ~mask == 0
            it += 1
        row["gpu_iter_per_s"] = round(it / (time.perf_counter() - t0), 1)
        if cpu:
            import oracle
            cops = oracle.op.qft(0b0111) * oracle.op.qft(0b1110)
        for th in (sorted({threads, 1}) if cpu else []):
            it, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 0.5 or it < 3:
                ref = oracle.Reg(q, 0, threads=th)
                ref.apply(cops)
                got = ref.measure_mask(mask, uniform(0x51564E54, q, it))
                assert got & ~mask == 0
                del ref
                it += 1
            row["cpu_port_iter_per_s_%dthreads" % th] = round(it / (time.perf_counter() - t0), 1)
        out["q%d" % q] = row
        log("reference bench q=%d: %s" % (q, row))
    return {"workload": "benches/performance.rs: with_state(q,0); apply(qft(0b0111)*qft(0b1110)); "
                        "measure_mask(0b100); iterations/s", "single_ops": 12, "by_qubits": out}


def config1_leg(threads):
    """The CPU-runnable case (20 qubits: qft on all qubits, h on all, 200
    random gates; the state fits the CPU caches, so the oracle is not
    page-fault bound here): the plain GPU form and the fused form (compiled
    again on every call) beside the oracle on all the host threads used for
    cpu_baseline and on one thread."""
    import oracle
    from qvnt_amd import QReg, circuits, op
    n = 20
    state, specs, _ = circuits.plain_20q()
    gops = circuits.build(specs, op)
    count = len(gops)
    reg = QReg.with_state(n, state)
    reg.apply(gops)
    reg.sync()
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 1.0:
        reg.apply(gops)
        reg.sync()
        reps += 1
    gpu = reps * count / (time.perf_counter() - t0)
    reg.close()
    # the fused form of the same queue (each call compiles it again), and
    # its agreement with the plain form from the same start state
    import numpy as np
    with QReg.with_state(n, state) as a, QReg.with_state(n, state) as b:
        a.apply(gops)
        b.apply_fused(gops)
        fused_diff = float(np.max(np.abs(a.get_state() - b.get_state())))
    if not fused_diff <= 1e-12:
        raise RuntimeError("20-qubit case: fused and plain forms differ by %.3g" % fused_diff)
    with QReg.with_state(n, state) as reg:
        reg.apply_fused(gops)
        reg.sync()
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 1.0:
            reg.apply_fused(gops)
            reg.sync()
            reps += 1
        fused = reps * count / (time.perf_counter() - t0)
    cpu = {}
    cops = circuits.build(specs, oracle.op)
    for th in sorted({threads, 1}, reverse=True):
        ref = oracle.Reg(n, state, threads=th)
        t0 = time.perf_counter()
        ref.apply(cops)
        cpu[str(th)] = round(count / (time.perf_counter() - t0), 1)
        del ref
    return {"workload": "plain-20q-fp64: qft(2^20-1), h on all qubits, 200 random gates (%d SingleOps)" % count,
            "gpu_plain_gates_per_s": round(gpu, 1), "gpu_fused_gates_per_s": round(fused, 1),
            "fused_vs_plain_max_abs_diff": fused_diff, "cpu_port_gates_per_s_by_threads": cpu}


def fused_leg(QReg, n, ops, warmup, timed_gates, plain_value, dense=0):
    """The same layers through qvnt_reg_apply_fused, one call per layer,
    then the whole timed circuit as one call (dense: the dense-phase
    threshold, 0 for none)."""
    out = {}
    for label, groups in (("per_layer", [[o] for o in ops[warmup:]]), ("whole_circuit", [ops[warmup:]])):
        reg = QReg(n)
        reg.set_dense(dense)
        for o in ops[:warmup]:
            reg.apply_fused(o)
        merged = []
        for grp in groups:
            m = grp[0].clone()
            for o in grp[1:]:
                m *= o
            merged.append(m)
        reg.sync()
        reg.set_timing(True)
        reg.kernel_stats(reset=True)
        passes = 0
        t0 = time.perf_counter()
        for m in merged:
            passes += reg.apply_fused(m)
        reg.sync()
        wall = time.perf_counter() - t0
        st = reg.kernel_stats(reset=True)["fused"]
        norm = reg.get_absolute()
        reg.close()
        v = timed_gates / wall
        ach = st["bytes"] / (st["ms"] * 1e-3) / 1e9 if st["ms"] else 0.0
        speedup = v / plain_value if plain_value else None
        log("fused (%s): %d passes for %d SingleOps in %.3f s -> %.1f gates/s (x%s plain); k_fused %.1f GB/s; |psi|^2-1 = %.1e"
            % (label, passes, timed_gates, wall, v, "%.1f" % speedup if speedup else "?", ach, norm - 1.0))
        out[label] = {"value": round(v, 3), "passes": passes, "seconds": round(wall, 4),
                      "speedup_vs_plain": round(speedup, 2) if speedup else None,
                      "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": "k_fused",
                                   "launches": st["launches"],
                                   "avg_ms": round(st["ms"] / st["launches"], 4) if st["launches"] else None,
                                   "bytes_per_launch": st["bytes"] / st["launches"] if st["launches"] else None}}
    return out


def tfim_leg(n, prec, plain_steps, fused_steps, per_step=False, dense=0, per_step_steps=0):
    """Transverse-field Ising Trotter steps on a ring (rzz on every edge, then
    rx on every qubit; J = 1, dt = 0.05, h_i ~ U[0.5, 1.5]); a step is 2n
    SingleOps.  The plain form runs plain_steps steps; the fused form runs
    fused_steps steps as one call (and one call per step: fused_steps steps
    with per_step, else per_step_steps steps).  One register serves every
    form (reset to |0> in between)."""
    from qvnt_amd import QReg, circuits, op
    step = circuits.build(circuits.tfim_step(n), op)
    amp = 16 if prec == "f64" else 8
    out = {"workload": "tfim-ring-%dq-%s" % (n, prec), "qubits": n, "state_gib": (amp << n) / 2**30,
           "single_ops_per_step": len(step)}
    log("tfim: %d qubits %s (%.1f GiB)" % (n, prec, (amp << n) / 2**30))
    reg = QReg(n, prec=prec)
    try:
        if plain_steps > 0:
            reg.set_timing(True)
            reg.sync()
            t0 = time.perf_counter()
            for _ in range(plain_steps):
                reg.apply(step)
            reg.sync()
            wall = time.perf_counter() - t0
            st = reg.kernel_stats(reset=True)
            reg.set_timing(False)
            norm = reg.get_absolute()
            gates = plain_steps * len(step)
            dom = max((k for k in st if st[k]["launches"]), key=lambda k: st[k]["ms"])
            d = st[dom]
            ach = d["bytes"] / (d["ms"] * 1e-3) / 1e9
            out["plain"] = {"value": round(gates / wall, 3), "unit": "gates/s", "steps": plain_steps,
                            "seconds": round(wall, 3), "norm_minus_1": norm - 1.0,
                            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                         "kernel": "k_" + dom, "launches": d["launches"],
                                         "avg_ms": round(d["ms"] / d["launches"], 3)}}
            log("tfim plain: %d SingleOps in %.2f s -> %.2f gates/s, k_%s %.0f GB/s; |psi|^2-1 = %.1e"
                % (gates, wall, gates / wall, dom, ach, norm - 1.0))
        out["fused"] = {}
        forms = [("whole", fused_steps, fused_steps)]
        if per_step or per_step_steps:
            forms.insert(0, ("per_step", 1, fused_steps if per_step else per_step_steps))
        for label, per_call, steps in forms if fused_steps > 0 else ():
            full = step.clone()
            for _ in range(per_call - 1):
                full *= step
            reg.reset(0)
            reg.set_dense(dense)
            reg.set_timing(True)
            reg.sync()
            reg.kernel_stats(reset=True)
            passes = 0
            t0 = time.perf_counter()
            for _ in range(steps // per_call):
                passes += reg.apply_fused(full)
            reg.sync()
            wall = time.perf_counter() - t0
            st = reg.kernel_stats(reset=True)
            reg.set_timing(False)
            norm = reg.get_absolute()
            gates = steps * len(step)
            f = st["fused"]
            ach = f["bytes"] / (f["ms"] * 1e-3) / 1e9 if f["ms"] else 0.0
            plain_v = out.get("plain", {}).get("value")
            out["fused"][label] = {"value": round(gates / wall, 3), "unit": "gates/s", "steps": steps,
                                   "passes": passes, "seconds": round(wall, 3),
                                   "speedup_vs_plain": round(gates / wall / plain_v, 2) if plain_v else None,
                                   "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                                                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                                "kernel": "k_fused", "launches": f["launches"],
                                                "avg_ms": round(f["ms"] / f["launches"], 3) if f["launches"] else None},
                                   "norm_minus_1": norm - 1.0}
            log("tfim fused (%s): %d steps, %d passes, %.2f s -> %.1f gates/s, k_fused %.0f GB/s; |psi|^2-1 = %.1e"
                % (label, steps, passes, wall, gates / wall, ach, norm - 1.0))
    finally:
        reg.close()
    return out


def tfim_pair_leg(n, steps, forms=("fused", "plain")):
    """The same TFIM evolution in an FP64 and an FP32 register side by side;
    after every step the device compares them in FP64 (max |dpsi|,
    infidelity 1 - |<a|b>|^2 / (|a|^2 |b|^2))."""
    from qvnt_amd import QReg, circuits, op
    step = circuits.build(circuits.tfim_step(n), op)
    out = {"workload": "tfim-ring-%dq fp32 vs fp64, %d steps" % (n, steps)}
    for form in forms:
        r64, r32 = QReg(n, prec="f64"), QReg(n, prec="f32")
        rows = []
        try:
            for k in range(steps):
                if form == "plain":
                    r64.apply(step)
                    r32.apply(step)
                else:
                    r64.apply_fused(step)
                    r32.apply_fused(step)
                err, infid, _, n64, n32 = r64.compare(r32)
                rows.append([k + 1, err, infid])
        finally:
            r64.close()
            r32.close()
        out[form] = {"max_abs_err_per_step": [r[1] for r in rows], "infidelity_per_step": [r[2] for r in rows],
                     "max_abs_err": max(r[1] for r in rows), "max_infidelity": max(r[2] for r in rows),
                     "final_max_abs_err": rows[-1][1], "final_infidelity": rows[-1][2]}
        log("tfim pair %d q (%s): max |dpsi| %.3e, max infidelity %.3e over %d steps"
            % (n, form, out[form]["max_abs_err"], out[form]["max_infidelity"], steps))
    return out


def serve_leg(requests, n_hi, slots, arena_gib, dense, plain=True, loads=(0.5, 0.8, 1.0), plain_requests=128):
    """Serving workload (n log-uniform in [12, n_hi], 100-2000 mixed gates,
    1024 seeded shots of measure_mask on 8 random qubits per request):
      serial_plain   closed loop, one request at a time, plain form, on the
                     first plain_requests requests (the plain form takes
                     ~0.3 s a request);
      serial_fused   closed loop, one request at a time, fused form (its
                     circuits/s is the capacity c of serial service);
      saturation     every request submitted at once into the
                     continuous-batching server (`slots` streams);
      open_loop      the same Poisson arrival times at load x c for each
                     load, into a one-slot server (serial fused service,
                     first come first served) and into the `slots` server.
    Latency = completion - submission (submission includes validation and
    the fused-program compile); circuits/s = requests / (last completion -
    first submission).  Every fused run must give the same outcomes."""
    import numpy as np
    from qvnt_amd import circuits, op
    from qvnt_amd.serve import Server
    reqs = circuits.serving_requests(requests, n_lo=12, n_hi=n_hi)
    t0 = time.perf_counter()
    ops = [circuits.build(specs, op) for _, specs, _, _ in reqs]
    qs = [r[0] for r in reqs]
    log("serve: %d requests built in %.1f s; qubits %d..%d (%d of >= 28); gates %d..%d" % (
        len(reqs), time.perf_counter() - t0, min(qs), max(qs), sum(q >= 28 for q in qs),
        min(len(o) for o in ops), max(len(o) for o in ops)))
    arena = int(arena_gib * 2**30)

    def stats(res, label, count=None):
        bad = [r["id"] for r in res if r["status"] != 0]
        if bad or len(res) != (count or len(reqs)):
            raise RuntimeError("%s: %d results, failed requests %s" % (label, len(res), bad[:8]))
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        o = {"circuits_per_s": round(len(res) / span, 3), "seconds": round(span, 3),
             "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
             "p90_ms": round(1e3 * float(np.percentile(lat, 90)), 2),
             "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
             "mean_ms": round(1e3 * float(lat.mean()), 2), "max_ms": round(1e3 * float(lat.max()), 2)}
        log("%s: %.2f circuits/s, p50 %.1f ms, p99 %.1f ms (%.1f s)" % (label, o["circuits_per_s"], o["p50_ms"],
                                                                        o["p99_ms"], span))
        return o

    def closed_loop(fused, count=None):
        res = []
        with Server(arena, slots=1) as srv:
            for k, (n, _, mask, draws) in enumerate(reqs[:count]):
                srv.submit(k, n, ops[k], mask, draws, fused=fused, dense_min=dense if fused else 0)
                res += srv.drain(1)
        return res

    def open_loop(n_slots, gaps):
        with Server(arena, slots=n_slots) as srv:
            start = srv.now()
            due = start + np.cumsum(gaps) - gaps[0]
            for k, (n, _, mask, draws) in enumerate(reqs):
                wait = due[k] - srv.now()
                if wait > 0:
                    time.sleep(wait)
                srv.submit(k, n, ops[k], mask, draws, fused=True, dense_min=dense)
            res = srv.drain(len(reqs))
            peak = srv.pending()[2]
        return res, peak

    out = {"workload": "serving-%d-requests-12..%dq-100..2000-gates-1024-shots" % (len(reqs), n_hi),
           "requests": len(reqs), "slots": slots, "arena_gib": arena_gib, "dense_min": dense,
           "note": "p99 of %d latencies" % len(reqs)}
    outcomes = {}

    def same(res):
        return all(np.array_equal(r["outcomes"], outcomes[r["id"]][0]) and r["total"] == outcomes[r["id"]][1]
                   for r in res)

    totals_plain = None
    if plain:
        n_plain = min(plain_requests, len(reqs))
        res = closed_loop(False, n_plain)
        out["serial_plain"] = stats(res, "serial plain (first %d requests)" % n_plain, n_plain)
        out["serial_plain"]["requests"] = n_plain
        totals_plain = {r["id"]: r["total"] for r in res}
    res = closed_loop(True)
    out["serial_fused"] = stats(res, "serial fused")
    # the overlap bound: requests of 2^27+ amplitudes run one at a time in
    # the server, so its makespan is at least their serial service time
    service = {r["id"]: r["t_done"] - r["t_submit"] for r in res}
    big = sum(t for k, t in service.items() if qs[k] >= 27)
    out["serial_fused"]["service_s_27q_plus"] = round(big, 3)
    out["serial_fused"]["service_s_all"] = round(sum(service.values()), 3)
    out["overlap_bound"] = round(sum(service.values()) / big, 3) if big else None
    log("serial fused: requests of 27+ qubits take %.3f of %.3f s of service; overlap can raise the rate at most %.2fx" % (
        big, sum(service.values()), out["overlap_bound"] or 0.0))
    outcomes = {r["id"]: (r["outcomes"], r["total"]) for r in res}
    if totals_plain:
        out["max_total_diff_fused_vs_plain"] = max(abs(outcomes[k][1] - totals_plain[k]) for k in totals_plain)
    equal = True
    res, peak = open_loop(slots, np.zeros(len(reqs)))
    out["saturation"] = stats(res, "served, all at once (%d slots)" % slots)
    out["saturation"]["arena_peak_gib"] = round(peak / 2**30, 2)
    equal = equal and same(res)
    cap = out["serial_fused"]["circuits_per_s"]
    unit_gaps = np.random.default_rng(circuits.SEED).exponential(1.0, len(reqs))
    out["open_loop"] = []
    for load in loads:
        rate = load * cap
        row = {"load": load, "offered_rate": round(rate, 3)}
        for label, n_slots in (("serial_fused", 1), ("served", slots)):
            res, _ = open_loop(n_slots, unit_gaps / rate)
            row[label] = stats(res, "%s at %.0f%% of serial capacity (%.1f/s)" % (label, 100 * load, rate))
            equal = equal and same(res)
        out["open_loop"].append(row)
    out["outcomes_equal_across_runs"] = bool(equal)
    if not equal:
        raise RuntimeError("served outcomes differ between runs")
    base = out.get("serial_plain", {}).get("circuits_per_s")
    out["saturation_vs_serial_plain"] = round(out["saturation"]["circuits_per_s"] / base, 2) if base else None
    out["saturation_vs_serial_fused"] = round(out["saturation"]["circuits_per_s"] / cap, 3)
    return out


def batch_leg(registers, n_hi, prec, tile_cap, dense, steps, warmup, check, plain=True):
    """Batched uneven registers (n uniform in [10, n_hi], 20-400 mixed gates,
    seeded), each from |0>.  The plain form applies every register's circuit
    one kernel per SingleOp, register after register; the batched form
    compiles all circuits (tile_cap 0: autotuned tile caps and dense-phase
    thresholds) and runs them round by round as one hipGraph, `steps` runs
    timed.  value = SingleOps/s of the batched form over all registers."""
    import numpy as np
    from qvnt_amd import QReg, circuits, op
    from qvnt_amd.batch import Batch
    t0 = time.perf_counter()
    regs = circuits.batched_registers(registers, n_lo=10, n_hi=n_hi)
    ops = [circuits.build(specs, op) for _, specs in regs]
    single = sum(len(o) for o in ops)
    amp = 16 if prec == "f64" else 8
    state_gib = sum(amp << n for n, _ in regs) / 2**30
    log("batch: %d registers, %d SingleOps, %.2f GiB of %s state (built in %.1f s)" % (
        len(regs), single, state_gib, prec, time.perf_counter() - t0))
    out = {"workload": "batched-%d-registers-10..%dq-20..400-gates-%s" % (len(regs), n_hi, prec),
           "registers": len(regs), "single_ops": single, "state_gib": round(state_gib, 2)}
    sample = min(check, len(regs))
    plain_states = []
    if plain:
        # every register, in chunks of 512 live registers (allocation outside the timed spans)
        wall = 0.0
        for c0 in range(0, len(regs), 512):
            live = [QReg(n, prec=prec) for n, _ in regs[c0:c0 + 512]]
            for r in live:
                r.sync()
            t0 = time.perf_counter()
            for r, o in zip(live, ops[c0:c0 + 512]):
                r.apply(o)
            for r in live:
                r.sync()
            wall += time.perf_counter() - t0
            plain_states += [live[k - c0].get_state()[: 1 << regs[k][0]] for k in range(c0, min(sample, c0 + 512))]
            for r in live:
                r.close()
        out["plain"] = {"value": round(single / wall, 1), "unit": "gates/s", "seconds": round(wall, 3)}
        log("plain: %.2f s -> %.0f SingleOps/s" % (wall, single / wall))
    with Batch([n for n, _ in regs], prec=prec) as b:
        t0 = time.perf_counter()
        info = b.compile(ops, max_qubits=tile_cap, dense_min=dense)
        comp = time.perf_counter() - t0
        log("compile%s: %.2f s, %d rounds, %d launches, %d passes, %d of %d phases dense; caps %s; dense %s" % (
            " + autotune" if not tile_cap else "", comp, info["rounds"], info["launches"], info["passes"],
            info["dense_phases"], info["phases"], info["tile_cap"], info["dense_min"]))
        b.run(graph=True)
        b.sync()
        errs, first = [], []
        for k in range(sample):
            got = b.get_state(k)
            first.append(got)
            if plain_states:
                errs.append(float(np.max(np.abs(got.astype(np.complex128) - plain_states[k]))))
        for _ in range(warmup):
            b.run(graph=True)
        ms_graph = b.time(steps, graph=True)
        ms_eager = b.time(max(1, steps // 2), graph=False)
        persist = None
        if info["persist_blocks"]:
            # the persistent form: one cooperative launch over all rounds
            b.reset()
            b.run(graph="persistent")
            b.sync()
            same = all(np.array_equal(b.get_state(k), first[k]) for k in range(sample))
            ms_persist = b.time(steps, graph="persistent")
            persist = {"ms_per_run": round(ms_persist, 3), "blocks": info["persist_blocks"],
                       "value": round(single / (ms_persist * 1e-3), 1), "equal_to_rounds": same,
                       "launch": b.persist_form}
            log("persistent launch (%d blocks, %s): %.2f ms/run; equal to the rounds on %d registers: %s" % (
                info["persist_blocks"], b.persist_form, ms_persist, sample, same))
    gbs = info["bytes"] / (ms_graph * 1e-3) / 1e9
    value = single / (ms_graph * 1e-3)
    tflops = info["dense_flops"] / (ms_graph * 1e-3) / 1e12
    log("batched: %.2f ms/run (graph), %.2f ms/run (eager) -> %.0f SingleOps/s, %.0f GB/s algorithmic, "
        "%.1f TFLOP/s on the matrix cores" % (ms_graph, ms_eager, value, gbs, tflops))
    out["batched"] = {"value": round(value, 1), "unit": "gates/s", "ms_per_run_graph": round(ms_graph, 3),
                      "ms_per_run_eager": round(ms_eager, 3), "runs_timed": steps, "compile_s": round(comp, 2),
                      "rounds": info["rounds"], "launches": info["launches"], "passes": info["passes"],
                      "tiles": info["tiles"], "phases": info["phases"], "dense_phases": info["dense_phases"],
                      "tile_cap": info["tile_cap"], "dense_min": info["dense_min"], "tune_ms": info["tune_ms"],
                      "single_ms": info["single_ms"],
                      "max_abs_err_vs_plain": max(errs) if errs else None, "checked_registers": len(errs),
                      "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel": "k_batch",
                                   "bytes_per_run": info["bytes"]},
                      "mfma": {"achieved": round(tflops, 3), "peak": MFMA_PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                               "frac": round(tflops / MFMA_PEAK_TFLOPS[prec], 4), "flops_per_run": info["dense_flops"],
                               "note": "dense 16x16 phases (real 32x32 GEMMs) only"}}
    out["batched"]["persistent"] = persist
    if "plain" in out:
        out["batched"]["speedup_vs_plain"] = round(value / out["plain"]["value"], 2)
    return out


def _line(workload_line, **extra):
    line = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None}
    line.update(workload_line)
    line.update(extra)
    return line


def tfim(args):
    res = tfim_leg(args.qubits, args.prec, args.steps, args.fused_steps, per_step=True, dense=args.dense)
    v = res.get("plain", {}).get("value")
    print(json.dumps(_line({"value": v, "unit": "gates/s", "steps": args.steps, "warmup": 0, "dtype": args.prec,
                            "data": "synthetic (seeded TFIM ring)",
                            "config": {"workload": res["workload"], "qubits": args.qubits,
                                       "fused_steps": args.fused_steps, "dense_min": args.dense}}, tfim=res)),
          flush=True)


def tfim_pair(args):
    forms = ("plain",) if args.no_fused else ("fused", "plain")
    res = tfim_pair_leg(args.qubits, args.fused_steps, forms)
    print(json.dumps({"config": {"workload": res["workload"]}, "tfim_pair": res}), flush=True)


def serve(args):
    res = serve_leg(args.requests, args.qubits, args.slots, args.arena_gib, args.dense, plain=not args.no_plain)
    print(json.dumps(_line({"value": res["saturation"]["circuits_per_s"], "unit": "circuits/s",
                            "steps": res["requests"], "warmup": 0, "dtype": "f64",
                            "data": "synthetic (seeded serving requests)",
                            "config": {"workload": res["workload"]}}, served=res)), flush=True)


def batch(args):
    res = batch_leg(args.registers, args.qubits, args.prec, args.tile_cap, args.dense, args.steps, args.warmup,
                    args.check_registers, plain=not args.no_plain)
    b = res["batched"]
    print(json.dumps(_line({"value": b["value"], "unit": "gates/s", "steps": args.steps, "warmup": args.warmup,
                            "ms_per_step": b["ms_per_run_graph"], "dtype": args.prec,
                            "data": "synthetic (seeded batched registers)", "config": {"workload": res["workload"]},
                            "roofline": b["roofline"]}, batch=res)), flush=True)


PLAIN_CLASSES = [
    # (label, constructor, args, control): every plain kernel class at the
    # register's size, with controls and targets inside and outside one
    # 128-byte line and the lane range (qubits 0-5)
    ("rz q0", "rz", (0.3, 1 << 0), 0), ("rz q12", "rz", (0.3, 1 << 12), 0), ("rz top", "rz", (0.3, "top"), 0),
    ("t q0", "t", (1 << 0,), 0), ("rzz q0q1", "rzz", (0.3, 0b11), 0),
    ("crz t1 c0", "rz", (0.3, 1 << 1), 1 << 0), ("crz t20 c12", "rz", (0.3, 1 << 20), 1 << 12),
    ("crz top c28", "rz", (0.3, "top"), 1 << 28), ("crz t5 top", "rz", (0.3, 1 << 5), "top"),
    ("ry q3", "ry", (0.3, 1 << 3), 0), ("ry q12", "ry", (0.3, 1 << 12), 0), ("ry top", "ry", (0.3, "top"), 0),
    ("x t12 c0", "x", (1 << 12,), 1 << 0), ("x top c1", "x", ("top",), 1 << 1), ("x t3 c2", "x", (1 << 3,), 1 << 2),
    ("x top c12", "x", ("top",), 1 << 12), ("x t0 top", "x", (1 << 0,), "top"), ("x t3 c7", "x", (1 << 3,), 1 << 7),
    ("x t12 c20", "x", (1 << 12,), 1 << 20), ("swap q0q9", "swap", ((1 << 0) | (1 << 9),), 0),
    ("h2 q0q1", "h", (0b11,), 0), ("h2 q12q13", "h", (3 << 12,), 0), ("h2 top pair", "h", ("top2",), 0),
    ("h2 q2q20", "h", ((1 << 2) | (1 << 20),), 0), ("h2 q5q6", "h", (3 << 5,), 0), ("ch2 q0q1 c2", "h", (0b11,), 1 << 2),
]


def plain_classes(args):
    """Every plain kernel class on one register (default 30 qubits FP64):
    --steps launches of each case, timed with HIP events on the register's
    stream; algorithmic bytes 2 * s * 2^(n - #controls) per launch.  Ends
    with get_absolute, whose k_leaf reads the state once (the FETCH_SIZE
    calibration of the PMC runs; the register's k_init writes it once)."""
    from qvnt_amd import QReg, circuits, op
    n, prec = args.qubits, args.prec
    top, amp = 1 << (n - 1), (16 if prec == "f64" else 8)

    def fix(v):
        return top if v == "top" else (3 << (n - 2)) if v == "top2" else v

    out = {}
    with QReg(n, prec=prec) as r:
        r.apply(op.h(0b111))
        for label, ctor, a, ctrl in PLAIN_CLASSES:
            o = circuits.build([(ctor, tuple(fix(x) for x in a), fix(ctrl))], op)
            r.apply(o)
            r.sync()
            r.set_timing(True)
            r.kernel_stats(reset=True)
            for _ in range(args.steps):
                r.apply(o)
            r.sync()
            st = r.kernel_stats(reset=True)
            r.set_timing(False)
            k = max((c for c in st if st[c]["launches"]), key=lambda c: st[c]["ms"])
            d = st[k]
            gbs = d["bytes"] / d["ms"] / 1e6
            out[label] = {"class": k, "avg_ms": round(d["ms"] / d["launches"], 4), "gbs": round(gbs, 1),
                          "frac": round(gbs / HBM_PEAK_GBS, 4)}
            log("%-12s %-5s %8.4f ms %8.1f GB/s %.3f" % (label, k, d["ms"] / d["launches"], gbs, gbs / HBM_PEAK_GBS))
        norm = r.get_absolute()
    print(json.dumps({"workload": "plain-classes-%dq-%s" % (n, prec), "launches_per_case": args.steps,
                      "norm": norm, "state_bytes": amp << n, "cases": out}), flush=True)


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get(kernel, {}).get("bytes_per_launch")


def tiles(args):
    """Tile sizes of the fused form at --qubits (default 30): four workloads
    (the brickwork, 40 layers, one call per layer and one call in all; the
    TFIM ring, 50 steps, one call per step and one call in all) run in tiles
    of 10, 11 and 12 qubits and with the automatic choice, each timed
    (wall clock around the calls, compile included).  Per tile size, a least
    squares fit of time = passes x a + cost x b over the four workloads
    (passes and the compiler's summed op cost from qvnt_fused_plan_stats):
    the model behind the automatic choice."""
    import numpy as np
    from qvnt_amd import QReg, circuits, fused_plan_stats, op
    n, prec = args.qubits, args.prec

    def merged(lst):
        m = lst[0].clone()
        for o in lst[1:]:
            m *= o
        return m

    bw = [circuits.build(circuits.brickwork_layer(n, l), op) for l in range(40)]
    tf = circuits.build(circuits.tfim_step(n), op)
    work = {"brickwork_per_layer": bw, "brickwork_whole": [merged(bw)], "tfim_per_step": [tf] * 50,
            "tfim_whole": [merged([tf] * 50)]}
    times, stats = {}, {}
    for name, calls in work.items():
        for cap in (10, 11, 12, 0):
            reg = QReg(n, prec=prec)
            reg.apply_fused(calls[0], max_qubits=cap)  # first-call allocations
            reg.sync()
            t0 = time.perf_counter()
            passes = 0
            for c in calls:
                passes += reg.apply_fused(c, max_qubits=cap)
            reg.sync()
            dt = time.perf_counter() - t0
            reg.close()
            st = [fused_plan_stats(c, n, prec, max_qubits=cap) for c in calls]
            stats[(name, cap)] = (sum(x["passes"] for x in st), sum(x["cost"] for x in st))
            assert stats[(name, cap)][0] == passes, (name, cap)
            times[(name, cap)] = dt
            log("%s %s tile %s: %d passes, %.4f s" % (prec, name, cap or "auto", passes, dt))
    fit = {}
    for cap in (10, 11, 12):
        A = np.array([stats[(w, cap)] for w in work], dtype=float)
        y = np.array([times[(w, cap)] for w in work])
        coef = np.linalg.lstsq(A, y, rcond=None)[0]
        fit[cap] = {"pass_s": float(coef[0]), "cost_s": float(coef[1]),
                    "max_rel_err": float(np.max(np.abs(A @ coef - y) / y))}
        log("fit tile %d: %.4e s/pass + %.4e s/cost, max rel. error %.3f" % (cap, coef[0], coef[1],
                                                                             fit[cap]["max_rel_err"]))
    best = {w: min((10, 11, 12), key=lambda c: times[(w, c)]) for w in work}
    print(json.dumps({"workload": "tiles", "qubits": n, "dtype": prec,
                      "seconds": {"%s/%s" % (w, c or "auto"): round(t, 4) for (w, c), t in times.items()},
                      "fastest_fixed_tile": best, "fit": fit}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--qubits", type=int, default=None,
                    help="register size (largest size for serve/batch); default per workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain form (profiling the fused kernel alone)")
    ap.add_argument("--no-extra", action="store_true",
                    help="brickwork: skip the served / TFIM / FP32-vs-FP64 / batched legs")
    ap.add_argument("--workload", default="brickwork",
                    choices=["brickwork", "tfim", "tfim-pair", "serve", "batch", "tiles", "plain-classes",
                             "reference"])
    ap.add_argument("--registers", type=int, default=4096, help="batch: number of registers")
    ap.add_argument("--tile-cap", type=int, default=0, help="batch: tile cap (0: autotune)")
    ap.add_argument("--dense", type=int, default=None,
                    help="fused forms: fold phases of >= this many ops into dense 16x16 unitaries (0: none; "
                         "default 0, serve 3; batch: used with --tile-cap, else autotuned)")
    ap.add_argument("--check-registers", type=int, default=64, help="batch: registers compared with the plain form")
    ap.add_argument("--requests", type=int, default=256, help="serve: number of requests")
    ap.add_argument("--slots", type=int, default=4,
                    help="serve: concurrent requests on the device (one stream each; the runtime has 4 hardware queues)")
    ap.add_argument("--arena-gib", type=float, default=160.0, help="serve: HBM admission budget")
    ap.add_argument("--prec", default="f64", choices=["f64", "f32"])
    ap.add_argument("--fused-steps", type=int, default=50, help="tfim: Trotter steps through the fused form")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.dense is None:
        args.dense = 3 if args.workload == "serve" else 0
    if args.qubits is None:
        args.qubits = {"brickwork": 30, "tfim": 34 if args.prec == "f64" else 35, "tfim-pair": 28, "serve": 30,
                       "batch": 20, "tiles": 30, "plain-classes": 30, "reference": 20}[args.workload]

    import qvnt_amd  # noqa: F401  (fails loudly without the HIP library)
    from qvnt_amd import QReg, circuits, op

    if args.workload == "tfim":
        return tfim(args)
    if args.workload == "tfim-pair":
        return tfim_pair(args)
    if args.workload == "serve":
        return serve(args)
    if args.workload == "batch":
        return batch(args)
    if args.workload == "tiles":
        return tiles(args)
    if args.workload == "plain-classes":
        return plain_classes(args)
    if args.workload == "reference":
        # the reference's Criterion loop alone (its GPU side without --no-cpu
        # adds the oracle's), e.g. for a HIP API trace of the per-register calls
        host = host_info()
        refb = reference_bench_leg(cpu_threads(host), cpu=not args.no_cpu)
        print(json.dumps({"config": {"workload": refb["workload"]}, "reference_bench": refb}), flush=True)
        return

    n = args.qubits
    depth = max(40, args.warmup + args.steps)   # the depth-40 brickwork circuit
    layers = [circuits.brickwork_layer(n, l) for l in range(args.warmup + args.steps)]
    ops = [circuits.build(s, op) for s in layers]
    timed_gates = sum(len(o) for o in ops[args.warmup:])
    label = "brickwork-%dq-fp64-depth%d, layers %d-%d timed" % (n, depth, args.warmup, args.warmup + args.steps - 1)

    if args.no_plain:
        fused = fused_leg(QReg, n, ops, args.warmup, timed_gates, None, args.dense)
        print(json.dumps({"config": {"workload": label}, "fused": fused}), flush=True)
        return

    reg = QReg(n)
    log("register: %d qubits FP64, %.1f GiB" % (n, (16 << n) / 2**30))
    for o in ops[:args.warmup]:
        reg.apply(o)
    reg.sync()
    reg.set_timing(True)
    reg.kernel_stats(reset=True)
    t0 = time.perf_counter()
    for o in ops[args.warmup:]:
        reg.apply(o)
    reg.sync()
    t1 = time.perf_counter()
    stats = reg.kernel_stats(reset=True)
    reg.set_timing(False)
    norm = reg.get_absolute()
    reg.close()
    wall = t1 - t0
    value = timed_gates / wall
    log("plain: %d SingleOps in %.3f s -> %.1f gates/s; |psi|^2 = %.15f" % (timed_gates, wall, value, norm))
    for k, s in stats.items():
        if s["launches"]:
            log("  %-8s %6d launches %10.2f ms  %8.1f GB/s" % (k, s["launches"], s["ms"], s["bytes"] / s["ms"] / 1e6))

    dom = max((k for k in stats if stats[k]["launches"]), key=lambda k: stats[k]["ms"])
    d = stats[dom]
    per_launch_bytes = d["bytes"] / d["launches"]
    avg_ms = d["ms"] / d["launches"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    kernel_name = {"diag": "k_diag", "pair": "k_pair", "quad": "k_quad"}.get(dom, dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(kernel_name),
                "traffic_source": "profiles/pmc_traffic.json (not measured in this run): rocprofv3 --pmc FETCH_SIZE "
                                  "x 2.000 (calibrated on k_leaf, which reads the 16 GiB state once and reports 8 GiB) "
                                  "+ WRITE_SIZE x 1.000 (k_init), rz on qubit 12 of a 30-qubit FP64 register "
                                  "(bench.py --workload plain-classes)",
                "kernel": kernel_name, "launches": d["launches"], "avg_ms": round(avg_ms, 4),
                "bytes_per_launch": per_launch_bytes,
                "timing": "HIP events on the register's stream around every launch of the timed layers"}
    all_ms = sum(s["ms"] for s in stats.values())
    all_bytes = sum(s["bytes"] for s in stats.values())

    fused = None
    if not args.no_fused:
        fused = fused_leg(QReg, n, ops, args.warmup, timed_gates, value, args.dense)

    host = host_info()
    threads = cpu_threads(host)
    cpu = c1 = refb = None
    if not args.no_cpu:
        log("cpu baseline (oracle, %d host threads; %s) ..." % (threads, host))
        cpu = cpu_baseline(n, threads)
        cpu["host"] = host
        log("cpu: %.3f gates/s on %d threads (%s)" % (cpu["value"], cpu["cores"], cpu["apply_seconds"]))
        c1 = config1_leg(threads)
        log("20-qubit case: GPU plain %.0f gates/s, fused %.0f gates/s (max |d| vs plain %.1e); oracle %s gates/s "
            "by threads" % (c1["gpu_plain_gates_per_s"], c1["gpu_fused_gates_per_s"], c1["fused_vs_plain_max_abs_diff"],
                            c1["cpu_port_gates_per_s_by_threads"]))
        refb = reference_bench_leg(threads)

    extra = {}
    if not args.no_extra:
        # each leg frees its device memory before the next allocates
        extra["tfim"] = {"f64": tfim_leg(34, "f64", 1, 50, per_step_steps=4),
                         "f32": tfim_leg(35, "f32", 1, 50, per_step_steps=4)}
        extra["tfim_pair"] = tfim_pair_leg(28, 50)
        extra["batch"] = {"f64": batch_leg(4096, 20, "f64", 0, 0, 20, 2, 64),
                          "f32": batch_leg(4096, 20, "f32", 0, 0, 20, 2, 64, plain=False)}
        extra["served"] = serve_leg(256, 30, 4, 160.0, 3)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded brickwork circuit, angles U[0,2pi))",
        "config": {"workload": label, "qubits": n, "state_gib": (16 << n) / 2**30, "singleops_timed": timed_gates,
                   "form": "plain (one in-place kernel per SingleOp)", "step": "one brickwork layer (105-106 SingleOps)"},
        "roofline": roofline,
        "all_kernels_gbs": round(all_bytes / all_ms / 1e6, 1) if all_ms else None,
        "fused": fused,
        "cpu_baseline": cpu,
        "config1": c1,
        "reference_bench": refb,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
