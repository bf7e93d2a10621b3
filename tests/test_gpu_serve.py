"""The continuous-batching server: a request's outcomes are the same alone
and in the mix, and equal the oracle's for plain-form requests."""
import numpy as np
import pytest

import qvnt_amd
from qvnt_amd import QReg, circuits, op
from qvnt_amd.serve import Server

pytestmark = pytest.mark.gpu


def make_requests(count, seed, n_lo=6, n_hi=16, g_lo=20, g_hi=200, shots=256):
    g = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        n = int(g.integers(n_lo, n_hi + 1))
        specs = circuits.random_mixed(n, int(g.integers(g_lo, g_hi + 1)), g)
        mask = int(sum(1 << int(q) for q in g.choice(n, min(8, n), replace=False)))
        draws = np.array([qvnt_amd.uniform(seed, k, j) for j in range(shots)])
        reqs.append((n, specs, mask, draws, bool(k % 3)))
    return reqs


def run(server, reqs, one_at_a_time):
    results = {}
    for k, (n, specs, mask, draws, fused) in enumerate(reqs):
        # every other fused request folds phases into dense 16x16 blocks
        server.submit(k, n, circuits.build(specs, op), mask, draws, fused=fused, dense_min=3 if k % 2 else 0)
        if one_at_a_time:
            for r in server.drain(1):
                results[r["id"]] = r
    if not one_at_a_time:
        for r in server.drain(len(reqs)):
            results[r["id"]] = r
    assert len(results) == len(reqs)
    return results


def test_alone_equals_mix(orc):
    reqs = make_requests(48, 11)
    with Server(4 << 30, slots=1) as s1:
        alone = run(s1, reqs, True)
    with Server(4 << 30, slots=8) as s8:
        mix = run(s8, reqs, False)
    for k, (n, specs, mask, draws, fused) in enumerate(reqs):
        a, m = alone[k], mix[k]
        assert a["status"] == 0 and m["status"] == 0
        assert np.array_equal(a["outcomes"], m["outcomes"])
        assert a["total"] == m["total"]
        assert np.all((a["outcomes"] & ~np.uint64(mask)) == 0)
        if not fused and n <= 14:
            # plain requests are bit-exact with the oracle: same outcomes
            ref = orc.Reg(n)
            ref.apply(circuits.build(specs, orc.op))
            psi = ref.state()
            want = [orc.tree_sample(psi[: 1 << n], u) & mask for u in draws]
            assert np.array_equal(a["outcomes"], np.array(want, np.uint64))


def test_server_matches_register_path():
    reqs = make_requests(12, 21, n_lo=10, n_hi=14)
    with Server(1 << 30, slots=4) as s:
        res = run(s, reqs, False)
    for k, (n, specs, mask, draws, fused) in enumerate(reqs):
        with QReg(n) as r:
            if fused:
                r.set_dense(3 if k % 2 else 0)
                r.apply_fused(circuits.build(specs, op))
            else:
                r.apply(circuits.build(specs, op))
            want = r.sample_mask(mask, draws)
            assert r.get_absolute() == res[k]["total"]
        assert np.array_equal(res[k]["outcomes"], want)


def test_admission_by_budget():
    """A request larger than the arena is refused; the rest are admitted as
    space frees up (some wait) and the arena's peak stays within its size."""
    budget = 40 << 20
    g = np.random.default_rng(3)
    small = circuits.build(circuits.random_mixed(18, 100, g), op)
    body = small.clone()
    for _ in range(3):
        body *= body                             # 800 plain gates: each 20-qubit request runs for milliseconds
    sizes = [20, 20, 19, 22, 18, 20, 19]         # 22 qubits (64 MiB) exceeds the arena; two 20-qubit states fill it
    with Server(budget, slots=8) as s:
        for k, n in enumerate(sizes):
            s.submit(k, n, body, 0xF, np.full(16, 0.5), fused=False)
        res = {r["id"]: r for r in s.drain(len(sizes))}
        _, _, peak = s.pending()
    assert res[3]["status"] == -3
    assert all(res[k]["status"] == 0 for k in range(len(sizes)) if k != 3)
    assert 0 < peak <= budget
    waited = [k for k in res if k != 3 and res[k]["t_admit"] > res[k]["t_submit"] + 1e-3]
    assert waited, "every request was admitted at once although the arena holds two 20-qubit states"
    for k in res:
        assert res[k]["t_submit"] <= res[k]["t_admit"] <= res[k]["t_done"]


def test_duplicate_live_id_is_refused():
    import ctypes
    from qvnt_amd._lib import E_ARG, lib
    from qvnt_amd.serve import Request
    o = op.h(0b111)
    draws = np.array([0.25, 0.75])
    with Server(1 << 28, slots=2) as s:
        s.submit(5, 3, o, 0b111, draws)
        # the Python side refuses it ...
        with pytest.raises(ValueError):
            s.submit(5, 3, o, 0b111, draws)
        # ... and so does the C side, for a caller that binds it directly
        outs = np.zeros(2, np.uint64)
        req = Request(5, 3, 0, 0, o.handle, 1, 0, 0, 0, 0b111, 2, draws.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      outs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        assert lib.qvnt_server_submit(s._h, ctypes.byref(req)) == E_ARG
        res = s.drain(1)
        assert len(res) == 1 and res[0]["id"] == 5 and res[0]["status"] == 0
        # once polled, the id may be used again
        s.submit(5, 3, o, 0b111, draws)
        res = s.drain(1)
        assert res[0]["status"] == 0


def test_large_requests_run_one_at_a_time():
    """Requests of 2^27 amplitudes or more are admitted one at a time (their
    [admit, done] intervals on the server clock do not overlap), while small
    requests run beside them; outcomes equal those served alone."""
    g = np.random.default_rng(77)
    reqs = []
    for k, n in enumerate([27, 10, 27, 12, 11, 27, 9]):
        specs = circuits.random_mixed(n, 30, g)
        mask = int(sum(1 << int(q) for q in g.choice(n, 4, replace=False)))
        draws = np.array([qvnt_amd.uniform(90, k, j) for j in range(64)])
        reqs.append((n, specs, mask, draws))
    ops = [circuits.build(specs, op) for _, specs, _, _ in reqs]
    with Server(16 << 30, slots=4) as s:
        for k, (n, _, mask, draws) in enumerate(reqs):
            s.submit(k, n, ops[k], mask, draws)
        mix = {r["id"]: r for r in s.drain(len(reqs))}
    assert len(mix) == len(reqs) and all(r["status"] == 0 for r in mix.values())
    big = sorted((mix[k]["t_admit"], mix[k]["t_done"]) for k, (n, _, _, _) in enumerate(reqs) if n >= 27)
    for (a0, d0), (a1, d1) in zip(big, big[1:]):
        assert a1 >= d0, big
    with Server(16 << 30, slots=1) as s:
        for k, (n, _, mask, draws) in enumerate(reqs):
            s.submit(k, n, ops[k], mask, draws)
            r = s.drain(1)[0]
            assert np.array_equal(r["outcomes"], mix[k]["outcomes"]) and r["total"] == mix[k]["total"]


def test_serving_sizes_alone_equals_mix():
    """Requests at the serving workload's sizes (log-uniform 12-30 qubits,
    two of them 28 and 30 qubits) with 100-300 gates of the mixed set and
    1024 shots: outcomes and totals bit for bit the same served one at a
    time and all at once."""
    reqs = circuits.serving_requests(22, seed=2024, n_lo=12, n_hi=26, g_lo=100, g_hi=300)
    g = np.random.default_rng(2025)
    for n in (28, 30):
        specs = circuits.random_mixed(n, 150, g)
        mask = int(sum(1 << int(q) for q in g.choice(n, 8, replace=False)))
        reqs.append((n, specs, mask, np.array([qvnt_amd.uniform(2026, n, k) for k in range(1024)])))
    ops = [circuits.build(specs, op) for _, specs, _, _ in reqs]
    print("serving sizes:", sorted(n for n, _, _, _ in reqs), flush=True)
    with Server(48 << 30, slots=1) as s1:
        alone = {}
        for k, (n, _, mask, draws) in enumerate(reqs):
            s1.submit(k, n, ops[k], mask, draws, dense_min=3)
            r = s1.drain(1)[0]
            alone[r["id"]] = r
    with Server(48 << 30, slots=4) as s4:
        for k, (n, _, mask, draws) in enumerate(reqs):
            s4.submit(k, n, ops[k], mask, draws, dense_min=3)
        mix = {r["id"]: r for r in s4.drain(len(reqs))}
    for k, (n, _, mask, _) in enumerate(reqs):
        a, m = alone[k], mix[k]
        assert a["status"] == 0 and m["status"] == 0
        assert np.array_equal(a["outcomes"], m["outcomes"]) and a["total"] == m["total"]
        assert abs(a["total"] - 1.0) < 1e-12
        assert np.all((a["outcomes"] & ~np.uint64(mask)) == 0)
