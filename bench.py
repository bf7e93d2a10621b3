#!/usr/bin/env python3
"""Benchmark: gates/s and % of HBM roofline of the plain FP64 apply on the
30-qubit brickwork random circuit (16 GiB state), one MI355X, with the
fused form on the same layers beside it (key "fused").

A step is one brickwork layer: u3 on every qubit (3 SingleOps each), rzz on
alternating nearest-neighbour pairs, one controlled x -- 105 or 106
reference SingleOps.  --steps K layers are timed after --warmup W layers
(default 40: the depth-40 circuit).  `value` counts reference SingleOps
(u3 = 3), the unit the plain form launches one kernel for.

The roofline object reports the kernel class that dominates device time,
from HIP events recorded on the register's stream around every launch in
the timed region: achieved = algorithmic bytes 2 * 16 B * 2^(n - #controls)
per launch / average launch time.  The CPU baseline times the oracle (the C
restatement of qvnt's out-of-place gather with its OpenMP twin of the Rayon
loop) on the host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}   # MI355X dense matrix-core peaks
METRIC = "gates/s + % HBM roofline at 30q FP64; served circuits/s at p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, budget_s):
    """Oracle (qvnt's algorithm in C, OpenMP) on the first SingleOps of
    brickwork layer 0 at full size, grown until ~budget_s of work."""
    import oracle
    from qvnt_amd import circuits
    threads = min(os.cpu_count() or 1, 16)
    specs = circuits.brickwork_layer(n, 0)
    full = circuits.build(specs, oracle.op)
    reg = oracle.Reg(n, 0, threads=threads if threads > 1 else 1)
    done, elapsed, count = 0, 0.0, 1
    while elapsed < budget_s and done < len(full.gates):
        count = min(count, len(full.gates) - done)
        part = oracle.Op(full.gates[done:done + count])
        t0 = time.perf_counter()
        reg.apply(part)
        dt = time.perf_counter() - t0
        elapsed += dt
        done += count
        per = dt / count
        count = max(1, int((budget_s - elapsed) / per / 2)) if per > 0 else 1
    del reg
    return {"value": done / elapsed, "unit": "gates/s", "cores": threads, "kind": "port",
            "sample": "first %d SingleOps of brickwork layer 0 at %d qubits (FP64, one MultiOp::apply per batch, "
                      "out-of-place gather + per-apply copy), %.1f s" % (done, n, elapsed)}


def config1_leg():
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
    for threads in sorted({min(os.cpu_count() or 1, 16), 1}, reverse=True):
        ref = oracle.Reg(n, state, threads=threads)
        t0 = time.perf_counter()
        ref.apply(cops)
        cpu[str(threads)] = round(count / (time.perf_counter() - t0), 1)
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


def tfim(args):
    """Transverse-field Ising Trotter steps on a ring (rzz on every edge, then
    rx on every qubit; J = 1, dt = 0.05, h_i ~ U[0.5, 1.5]); a step is one
    Trotter step of 2n SingleOps.  The plain form runs --steps steps; the
    fused form runs all --fused-steps steps, one call per step and then the
    whole evolution as one call."""
    from qvnt_amd import QReg, circuits, op
    n, prec = args.qubits, args.prec
    step = circuits.build(circuits.tfim_step(n), op)
    amp = 16 if prec == "f64" else 8
    out = {"plain": None, "fused": {}}
    log("tfim: %d qubits %s (%.1f GiB)" % (n, prec, (amp << n) / 2**30))
    if args.steps > 0 and not args.no_plain:
        reg = QReg(n, prec=prec)
        reg.set_timing(True)
        reg.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            reg.apply(step)
        reg.sync()
        wall = time.perf_counter() - t0
        st = reg.kernel_stats(reset=True)
        norm = reg.get_absolute()
        reg.close()
        gates = args.steps * len(step)
        dom = max((k for k in st if st[k]["launches"]), key=lambda k: st[k]["ms"])
        d = st[dom]
        ach = d["bytes"] / (d["ms"] * 1e-3) / 1e9
        out["plain"] = {"value": round(gates / wall, 3), "steps": args.steps, "seconds": round(wall, 3),
                        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": dom,
                                     "avg_ms": round(d["ms"] / d["launches"], 3)}}
        log("tfim plain: %d SingleOps in %.2f s -> %.2f gates/s, %s %.0f GB/s; |psi|^2-1 = %.1e"
            % (gates, wall, gates / wall, dom, ach, norm - 1.0))
    for label, per_call in (("per_step", 1), ("whole", args.fused_steps)) if args.fused_steps > 0 else ():
        full = step.clone()
        for _ in range(per_call - 1):
            full *= step
        reg = QReg(n, prec=prec)
        reg.set_dense(args.dense)
        reg.set_timing(True)
        reg.sync()
        passes = 0
        t0 = time.perf_counter()
        for _ in range(args.fused_steps // per_call):
            passes += reg.apply_fused(full)
        reg.sync()
        wall = time.perf_counter() - t0
        st = reg.kernel_stats(reset=True)["fused"]
        norm = reg.get_absolute()
        reg.close()
        gates = args.fused_steps * len(step)
        ach = st["bytes"] / (st["ms"] * 1e-3) / 1e9
        out["fused"][label] = {"value": round(gates / wall, 3), "passes": passes, "seconds": round(wall, 3),
                               "k_fused_gbs": round(ach, 1), "norm_minus_1": norm - 1.0}
        log("tfim fused (%s): %d steps, %d passes, %.2f s -> %.1f gates/s, k_fused %.0f GB/s; |psi|^2-1 = %.1e"
            % (label, args.fused_steps, passes, wall, gates / wall, ach, norm - 1.0))
    line = {"metric": METRIC, "value": out["plain"]["value"] if out["plain"] else None, "unit": "gates/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 0, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": prec, "data": "synthetic (seeded TFIM ring)",
            "config": {"workload": "tfim-ring-%dq-%s" % (n, prec), "qubits": n, "state_gib": (amp << n) / 2**30,
                       "fused_steps": args.fused_steps, "dense_min": args.dense},
            "plain": out["plain"], "fused": out["fused"]}
    print(json.dumps(line), flush=True)


def tfim_pair(args):
    """The same TFIM evolution in an FP64 and an FP32 register side by side;
    after every step the device compares them (max |dpsi|, infidelity)."""
    from qvnt_amd import QReg, circuits, op
    n = args.qubits
    step = circuits.build(circuits.tfim_step(n), op)
    r64, r32 = QReg(n, prec="f64"), QReg(n, prec="f32")
    rows = []
    for k in range(args.fused_steps):
        if args.no_fused:
            r64.apply(step)
            r32.apply(step)
        else:
            r64.apply_fused(step)
            r32.apply_fused(step)
        err, infid, _, n64, n32 = r64.compare(r32)
        rows.append({"step": k + 1, "max_abs_err": err, "infidelity": infid, "norm64_minus_1": n64 - 1.0,
                     "norm32_minus_1": n32 - 1.0})
        log("step %2d: max|dpsi| %.3e  infidelity %.3e" % (k + 1, err, infid))
    print(json.dumps({"config": {"workload": "tfim-ring-%dq fp32 vs fp64" % n, "form": "plain" if args.no_fused else "fused"},
                      "steps": rows}), flush=True)


def serve(args):
    """Serving workload: --requests circuits (n log-uniform in [12, 30],
    100-2000 mixed gates, 1024 seeded shots of measure_mask on 8 random
    qubits), three ways:
      serial_plain  one request at a time, plain form (a closed loop: the next
                    request is submitted when the last has finished);
      serial_fused  one request at a time, fused form;
      served        Poisson arrivals at --rate requests/s into the
                    continuous-batching server (--slots streams, an HBM arena
                    of --arena-gib), fused form.
    Latency = completion - submission (submission includes validation and
    the fused-program compile); circuits/s = requests / (last completion -
    first submission)."""
    import numpy as np
    from qvnt_amd import circuits, op
    from qvnt_amd.serve import Server
    reqs = circuits.serving_requests(args.requests, n_lo=12, n_hi=args.qubits)
    t0 = time.perf_counter()
    ops = [circuits.build(specs, op) for _, specs, _, _ in reqs]
    log("serve: %d requests built in %.1f s; qubits %s; gates %d..%d" % (
        len(reqs), time.perf_counter() - t0, np.bincount([r[0] for r in reqs]).nonzero()[0].tolist(),
        min(len(o) for o in ops), max(len(o) for o in ops)))
    arena = int(args.arena_gib * 2**30)

    def stats(res, label):
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        bad = [r["id"] for r in res if r["status"] != 0]
        if bad:
            raise RuntimeError("%s: requests %s failed" % (label, bad[:8]))
        out = {"circuits_per_s": round(len(res) / span, 3), "seconds": round(span, 3),
               "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
               "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
               "mean_ms": round(1e3 * float(lat.mean()), 2)}
        log("%s: %.2f circuits/s, p50 %.1f ms, p99 %.1f ms (%.1f s)" % (label, out["circuits_per_s"], out["p50_ms"],
                                                                        out["p99_ms"], span))
        return out

    def closed_loop(fused, slots=1):
        res = []
        with Server(arena, slots=slots) as srv:
            for k, (n, _, mask, draws) in enumerate(reqs):
                srv.submit(k, n, ops[k], mask, draws, fused=fused, dense_min=args.dense if fused else 0)
                res += srv.drain(1)
        return res

    out = {}
    totals = {}
    if not args.no_plain:
        res = closed_loop(False)
        out["serial_plain"] = stats(res, "serial plain")
        totals["plain"] = {r["id"]: r["total"] for r in res}
    res = closed_loop(True)
    out["serial_fused"] = stats(res, "serial fused")
    totals["fused"] = {r["id"]: r["total"] for r in res}
    serial_outcomes = {r["id"]: r["outcomes"] for r in res}
    # open loop: Poisson arrivals at the offered rate
    gaps = np.random.default_rng(circuits.SEED).exponential(1.0 / args.rate, len(reqs))
    with Server(arena, slots=args.slots) as srv:
        start = srv.now()
        due = start + np.cumsum(gaps) - gaps[0]
        for k, (n, _, mask, draws) in enumerate(reqs):
            wait = due[k] - srv.now()
            if wait > 0:
                time.sleep(wait)
            srv.submit(k, n, ops[k], mask, draws, fused=True, dense_min=args.dense)
        res = srv.drain(len(reqs))
        _, _, peak = srv.pending()
    out["served"] = stats(res, "served (rate %.1f/s, %d slots)" % (args.rate, args.slots))
    out["served"]["offered_rate"] = args.rate
    out["served"]["slots"] = args.slots
    out["served"]["arena_peak_gib"] = round(peak / 2**30, 2)
    same = all(np.array_equal(r["outcomes"], serial_outcomes[r["id"]]) and r["total"] == totals["fused"][r["id"]]
               for r in res)
    out["served"]["outcomes_equal_serial_fused"] = bool(same)
    if "plain" in totals:
        out["max_total_diff_fused_vs_plain"] = max(abs(totals["fused"][k] - totals["plain"][k]) for k in totals["plain"])
    base = out.get("serial_plain", out["serial_fused"])["circuits_per_s"]
    line = {"metric": METRIC, "value": out["served"]["circuits_per_s"], "unit": "circuits/s", "n_gpus": 1,
            "steps": len(reqs), "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (seeded serving requests)",
            "config": {"workload": "serving-%d-requests-12..%dq-100..2000-gates-1024-shots" % (len(reqs), args.qubits),
                       "arena_gib": args.arena_gib, "dense_min": args.dense},
            "speedup_vs_serial_plain": round(out["served"]["circuits_per_s"] / base, 2) if base else None}
    line.update(out)
    print(json.dumps(line), flush=True)


def batch(args):
    """Batched uneven registers: --registers registers (n uniform in [10, 20],
    20-400 mixed gates, seeded), each from |0>.  The plain form applies every
    register's circuit one kernel per SingleOp, register after register; the
    batched form compiles all circuits (autotuned tile caps) and runs them
    round by round, as one hipGraph, --steps times.  value = SingleOps/s of
    the batched form over all registers."""
    import numpy as np
    from qvnt_amd import QReg, circuits, op
    from qvnt_amd.batch import Batch
    prec = args.prec
    t0 = time.perf_counter()
    regs = circuits.batched_registers(args.registers, n_lo=10, n_hi=args.qubits)
    ops = [circuits.build(specs, op) for _, specs in regs]
    single = sum(len(o) for o in ops)
    amp = 16 if prec == "f64" else 8
    state_gib = sum(amp << n for n, _ in regs) / 2**30
    log("batch: %d registers, %d SingleOps, %.2f GiB of %s state (built in %.1f s)" % (
        len(regs), single, state_gib, prec, time.perf_counter() - t0))
    out = {}
    sample = min(args.check_registers, len(regs))
    plain_states = []
    if not args.no_plain:
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
        out["plain"] = {"value": round(single / wall, 1), "seconds": round(wall, 3)}
        log("plain: %.2f s -> %.0f SingleOps/s" % (wall, single / wall))
    with Batch([n for n, _ in regs], prec=prec) as b:
        t0 = time.perf_counter()
        info = b.compile(ops, max_qubits=args.tile_cap, dense_min=args.dense)
        comp = time.perf_counter() - t0
        log("compile%s: %.2f s, %d rounds, %d launches, %d passes, %d of %d phases dense; caps %s; dense %s" % (
            " + autotune" if not args.tile_cap else "", comp, info["rounds"], info["launches"], info["passes"],
            info["dense_phases"], info["phases"], info["tile_cap"], info["dense_min"]))
        b.run(graph=True)
        b.sync()
        errs = []
        for k in range(sample):
            got = b.get_state(k)
            if plain_states:
                errs.append(float(np.max(np.abs(got.astype(np.complex128) - plain_states[k]))))
        for _ in range(args.warmup):
            b.run(graph=True)
        ms_graph = b.time(args.steps, graph=True)
        ms_eager = b.time(max(1, args.steps // 2), graph=False)
    gbs = info["bytes"] / (ms_graph * 1e-3) / 1e9
    value = single / (ms_graph * 1e-3)
    log("batched: %.2f ms/run (graph), %.2f ms/run (eager) -> %.0f SingleOps/s, %.0f GB/s algorithmic" % (
        ms_graph, ms_eager, value, gbs))
    out["batched"] = {"value": round(value, 1), "ms_per_run_graph": round(ms_graph, 3),
                      "ms_per_run_eager": round(ms_eager, 3), "compile_s": round(comp, 2),
                      "rounds": info["rounds"], "launches": info["launches"], "passes": info["passes"],
                      "tiles": info["tiles"], "phases": info["phases"], "dense_phases": info["dense_phases"],
                      "tile_cap": info["tile_cap"], "dense_min": info["dense_min"], "tune_ms": info["tune_ms"],
                      "single_ms": info["single_ms"],
                      "max_abs_err_vs_plain": max(errs) if errs else None, "checked_registers": len(errs)}
    line = {"metric": METRIC, "value": round(value, 1), "unit": "gates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_graph, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": prec, "data": "synthetic (seeded batched registers)",
            "config": {"workload": "batched-%d-registers-10..%dq-20..400-gates-%s" % (len(regs), args.qubits, prec),
                       "registers": len(regs), "single_ops": single, "state_gib": round(state_gib, 2)},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_batch"},
            "mfma": {"achieved": round(info["dense_flops"] / (ms_graph * 1e-3) / 1e12, 3),
                     "peak": MFMA_PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                     "frac": round(info["dense_flops"] / (ms_graph * 1e-3) / 1e12 / MFMA_PEAK_TFLOPS[prec], 4),
                     "flops_per_run": info["dense_flops"], "note": "dense 16x16 phases (real 32x32 GEMMs) only"},
            "speedup_vs_plain": round(value / out["plain"]["value"], 2) if "plain" in out else None}
    line.update(out)
    print(json.dumps(line), flush=True)


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
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=None,
                    help="register size (largest size for serve/batch); default per workload")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain form (profiling the fused kernel alone)")
    ap.add_argument("--workload", default="brickwork", choices=["brickwork", "tfim", "tfim-pair", "serve", "batch", "tiles"])
    ap.add_argument("--registers", type=int, default=4096, help="batch: number of registers")
    ap.add_argument("--tile-cap", type=int, default=0, help="batch: tile cap (0: autotune)")
    ap.add_argument("--dense", type=int, default=None,
                    help="fused forms: fold phases of >= this many ops into dense 16x16 unitaries (0: none; "
                         "default 0, serve 3; batch: used with --tile-cap, else autotuned)")
    ap.add_argument("--check-registers", type=int, default=64, help="batch: registers compared with the plain form")
    ap.add_argument("--requests", type=int, default=64, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: offered Poisson rate, requests/s")
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
                       "batch": 20, "tiles": 30}[args.workload]

    import qvnt_amd
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

    n = args.qubits
    layers = [circuits.brickwork_layer(n, l) for l in range(args.warmup + args.steps)]
    ops = [circuits.build(s, op) for s in layers]
    timed_gates = sum(len(o) for o in ops[args.warmup:])

    if args.no_plain:
        fused = fused_leg(QReg, n, ops, args.warmup, timed_gates, None, args.dense)
        print(json.dumps({"fused": fused}), flush=True)
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
                "kernel": kernel_name, "launches": d["launches"], "avg_ms": round(avg_ms, 4),
                "bytes_per_launch": per_launch_bytes}
    all_ms = sum(s["ms"] for s in stats.values())
    all_bytes = sum(s["bytes"] for s in stats.values())

    fused = None
    if not args.no_fused:
        fused = fused_leg(QReg, n, ops, args.warmup, timed_gates, value, args.dense)

    cpu = None
    c1 = None
    if not args.no_cpu:
        log("cpu baseline (oracle, host cores) ...")
        cpu = cpu_baseline(n, args.cpu_seconds)
        log("cpu: %.3f gates/s on %d threads" % (cpu["value"], cpu["cores"]))
        c1 = config1_leg()
        log("20-qubit case: GPU plain %.0f gates/s, fused %.0f gates/s (max |d| vs plain %.1e); oracle %s gates/s "
            "by threads" % (c1["gpu_plain_gates_per_s"], c1["gpu_fused_gates_per_s"], c1["fused_vs_plain_max_abs_diff"],
                            c1["cpu_port_gates_per_s_by_threads"]))

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded brickwork circuit, angles U[0,2pi))",
        "config": {"workload": "brickwork-%dq-fp64-depth%d-plain" % (n, args.steps), "qubits": n,
                   "state_gib": (16 << n) / 2**30, "singleops_timed": timed_gates,
                   "form": "plain (one in-place kernel per SingleOp)"},
        "roofline": roofline,
        "all_kernels_gbs": round(all_bytes / all_ms / 1e6, 1) if all_ms else None,
        "fused": fused,
        "cpu_baseline": cpu,
        "config1": c1,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
