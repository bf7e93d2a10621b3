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


def build(specs, ns):
    """One MultiOp (queue order) from specs, with namespace `ns` (op::)."""
    out = ns.id()
    for ctor, args, ctrl in specs:
        o = getattr(ns, ctor)(*args)
        if ctrl:
            o = o.c(ctrl)
            if o is None:
                raise ValueError("control overlaps the gate in spec %r" % ((ctor, args, ctrl),))
        out = out * o
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
