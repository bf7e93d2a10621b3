"""Seeded synthetic circuits for the benchmark and parity workloads.

A circuit is a list of *specs*: ``(ctor, args, ctrl)`` with ``ctor`` an op::
constructor name, ``args`` its arguments in the reference order and ``ctrl``
a control mask (0 for none).  ``build(spec, ns)`` turns specs into one
MultiOp with any namespace exposing the op:: constructors, so the same
circuit can be built for the device engine and for a checker.

Workloads:
  plain_20q     20 qubits: qft(all), h(all), 200 random gates from
                {rx, ry, rz, phi, controlled x}, then measure_mask(0xFF).
  brickwork     n qubits, per layer: u3 on every qubit, rzz on alternating
                even/odd nearest-neighbour pairs, one controlled x.
  tfim          transverse-field Ising Trotter steps on a ring: rzz(2 J dt)
                on every ring edge, then rx(2 h_i dt) on every qubit.
  random_mixed  gates drawn uniformly from {h, x, y, z, s, t, rx, ry, rz, u3,
                rxx, ryy, rzz, swap, iswap, controlled x}, angles U[0, 2pi):
                the batched-register and serving workloads.
  serving       requests with n log-uniform in [12, 30], 100-2000
                random_mixed gates, shots on 8 random qubits.
"""
import math

import numpy as np

TWO_PI = 2.0 * math.pi
SEED = 0x51564E54


def rng(seed, tag):
    return np.random.Generator(np.random.PCG64([seed, tag]))


def plain_20q(seed=SEED, n=20, random_gates=200):
    """(initial basis state, specs, measure mask) of the plain 20-qubit workload."""
    g = rng(seed, 1)
    state = int(g.integers(0, 1 << n))
    full = (1 << n) - 1
    specs = [("qft", (full,), 0), ("h", (full,), 0)]
    for _ in range(random_gates):
        kind = int(g.integers(0, 5))
        q = int(g.integers(0, n))
        if kind < 4:
            angle = float(g.uniform(0.0, TWO_PI))
            specs.append((("rx", "ry", "rz", "phi")[kind], (angle, 1 << q), 0))
        else:
            c = int(g.integers(0, n - 1))
            c = c + 1 if c >= q else c
            specs.append(("x", (1 << q,), 1 << c))
    return state, specs, 0xFF


def brickwork_layer(n, layer, seed=SEED):
    """Specs of brickwork layer `layer` on n qubits (open chain)."""
    g = rng(seed, 1000 + layer)
    specs = []
    for q in range(n):
        the, phi, lam = (float(v) for v in g.uniform(0.0, TWO_PI, 3))
        specs.append(("u3", (the, phi, lam, 1 << q), 0))
    for q in range(layer & 1, n - 1, 2):
        specs.append(("rzz", (float(g.uniform(0.0, TWO_PI)), (1 << q) | (1 << (q + 1))), 0))
    c, t = (int(v) for v in g.choice(n, 2, replace=False))
    specs.append(("x", (1 << t,), 1 << c))
    return specs


def tfim_step(n, dt=0.05, J=1.0, seed=SEED):
    """Specs of one Trotter step on a ring of n qubits (h_i ~ U[0.5, 1.5], seeded once)."""
    h = rng(seed, 3).uniform(0.5, 1.5, n)
    specs = [("rzz", (2.0 * J * dt, (1 << q) | (1 << ((q + 1) % n))), 0) for q in range(n)]
    specs += [("rx", (2.0 * float(h[q]) * dt, 1 << q), 0) for q in range(n)]
    return specs


MIXED = ("h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "u3", "rxx", "ryy", "rzz", "swap", "i_swap", "cx")


def random_mixed(n, count, g):
    """`count` gates on n >= 2 qubits from the mixed set, drawn from generator g."""
    ctors = g.integers(len(MIXED), size=count)
    qa = g.integers(n, size=count)
    qb = (qa + 1 + g.integers(n - 1, size=count)) % n      # a second, distinct qubit
    ang = g.uniform(0.0, TWO_PI, size=(count, 3))
    specs = []
    for ctor, a, b, (t0, t1, t2) in zip(ctors.tolist(), qa.tolist(), qb.tolist(), ang.tolist()):
        ctor = MIXED[ctor]
        if ctor in ("h", "x", "y", "z", "s", "t"):
            specs.append((ctor, (1 << a,), 0))
        elif ctor in ("rx", "ry", "rz"):
            specs.append((ctor, (t0, 1 << a), 0))
        elif ctor == "u3":
            specs.append(("u3", (t0, t1, t2, 1 << a), 0))
        elif ctor in ("rxx", "ryy", "rzz"):
            specs.append((ctor, (t0, (1 << a) | (1 << b)), 0))
        elif ctor in ("swap", "i_swap"):
            specs.append((ctor, ((1 << a) | (1 << b),), 0))
        else:
            specs.append(("x", (1 << a,), 1 << b))
    return specs


def serving_requests(count, seed=SEED, n_lo=12, n_hi=30, g_lo=100, g_hi=2000, shots=1024):
    """Requests of the serving workload: (n, specs, mask, draws)."""
    g = rng(seed, 5)
    out = []
    for k in range(count):
        n = int(round(math.exp(g.uniform(math.log(n_lo), math.log(n_hi)))))
        specs = random_mixed(n, int(g.integers(g_lo, g_hi + 1)), g)
        qs = g.choice(n, min(8, n), replace=False)
        mask = sum(1 << int(q) for q in qs)
        draws = rng(seed, 100000 + k).random(shots)
        out.append((n, specs, mask, draws))
    return out


def build(specs, ns):
    """One MultiOp (queue order) from specs, with namespace `ns` (op::)."""
    out = ns.id()
    for ctor, args, ctrl in specs:
        o = getattr(ns, ctor)(*args)
        if ctrl:
            o = o.c(ctrl)
            if o is None:
                raise ValueError("control overlaps the gate in spec %r" % ((ctor, args, ctrl),))
        out *= o
    return out


def single_ops(specs):
    """Reference SingleOp count of the specs (u3/u2 = 3, h(k) = ceil(k/2), qft(k) = k(k+1)/2)."""
    total = 0
    for ctor, args, _ in specs:
        if ctor in ("u3", "u2"):
            total += 3
        elif ctor == "h":
            k = bin(args[0]).count("1")
            total += (k + 1) // 2
        elif ctor == "qft":
            k = bin(args[0]).count("1")
            total += k * (k + 1) // 2 if k > 1 else k
        elif ctor == "id":
            total += 0
        else:
            total += 1
    return total


def batched_registers(count, seed=SEED, n_lo=10, n_hi=20, g_lo=20, g_hi=400):
    """The batched-register workload: (n, specs) per register, n uniform in
    [n_lo, n_hi], g_lo..g_hi random_mixed gates."""
    g = rng(seed, 4)
    out = []
    for _ in range(count):
        n = int(g.integers(n_lo, n_hi + 1))
        out.append((n, random_mixed(n, int(g.integers(g_lo, g_hi + 1)), g)))
    return out
