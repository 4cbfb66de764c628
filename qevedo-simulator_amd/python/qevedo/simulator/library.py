"""Workload generators for the state-vector engine.

A circuit here is ``Circuit`` (see ``qevedo.simulator``) or, for the
lower-level helpers, a plain list of ``(name, qubits, params)`` gate tuples
with OpenQASM 2 ``qelib1`` names. Generators mirror ``src/library.rs`` of
the reference where it defines the circuit (``ghz`` :9-16, ``qft`` :19-29,
``random_gate_circuit`` :103-141, ``eagle_heavy_hex`` :392-425,
``kicked_ising`` :430-446). Randomness comes from NumPy's PCG64 generator
seeded with the given seed; the reference's own ``rand_pcg`` streams are not
reproduced, so a seed names a circuit of this module, not of the reference.
"""
from __future__ import annotations

import math

import numpy as np

HALF_PI = math.pi / 2
QUARTER_PI = math.pi / 4

# The three single-qubit gates of the random benchmark circuit: sqrt(X) as
# RX(pi/2), sqrt(Y) as RY(pi/2) and sqrt(W) as U(pi/2, -pi/4, pi/4)
# (library.rs:41-50).
SQRT_GATES = (
    ("rx", (HALF_PI,)),
    ("ry", (HALF_PI,)),
    ("u3", (HALF_PI, -QUARTER_PI, QUARTER_PI)),
)


def ghz(n: int) -> list:
    """H on qubit 0, then a CX chain (library.rs:9-16)."""
    gates = [("h", (0,), ())]
    gates += [("cx", (q - 1, q), ()) for q in range(1, n)]
    return gates


def qft(n: int) -> list:
    """Quantum Fourier transform without the final reversal (library.rs:19-29)."""
    gates = []
    for target in reversed(range(n)):
        gates.append(("h", (target,), ()))
        for control in reversed(range(target)):
            gates.append(("cp", (control, target), (math.pi / float(1 << (target - control)),)))
    return gates


def random_circuit(n: int, depth: int, seed: int) -> list:
    """The random benchmark circuit with the no-repeat rule.

    Each layer applies one of sqrt(X), sqrt(Y), sqrt(W) to every qubit, then
    CZ on the brick pattern (q, q+1), q = layer % 2, layer % 2 + 2, ...
    (library.rs:53-57). Layer 0 draws each gate uniformly from the three;
    every later layer draws uniformly from the two gates that differ from the
    qubit's previous one (listed in SQRT_GATES order).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    gates = []
    previous = [None] * n
    for layer in range(depth):
        for q in range(n):
            if previous[q] is None:
                g = int(rng.integers(0, 3))
            else:
                others = [i for i in range(3) if i != previous[q]]
                g = others[int(rng.integers(0, 2))]
            previous[q] = g
            name, params = SQRT_GATES[g]
            gates.append((name, (q,), params))
        q = layer % 2
        while q + 1 < n:
            gates.append(("cz", (q, q + 1), ()))
            q += 2
    return gates


def random_gate_circuit(n: int, count: int, seed: int) -> list:
    """Random gates over the standard set (library.rs:103-141, own PRNG)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gates = []
    for _ in range(count):
        qubits = list(range(n))
        for i in range(min(n, 3)):
            j = int(rng.integers(i, n))
            qubits[i], qubits[j] = qubits[j], qubits[i]
        angle = float(rng.random()) * 2 * math.pi
        choice = int(rng.integers(0, 14 if n >= 3 else 12 if n == 2 else 8))
        table = [
            ("h", 1, ()), ("x", 1, ()), ("rx", 1, (angle,)), ("ry", 1, (angle,)),
            ("rz", 1, (angle,)), ("u3", 1, (angle, angle * 0.5, angle * 0.25)), ("t", 1, ()),
            ("sx", 1, ()), ("cx", 2, ()), ("cz", 2, ()), ("cp", 2, (angle,)), ("swap", 2, ()),
            ("ccx", 3, ()), ("cswap", 3, ()),
        ]
        name, arity, params = table[choice]
        gates.append((name, tuple(qubits[:arity]), params))
    return gates


def eagle_heavy_hex() -> list:
    """Couplers of IBM's 127-qubit Eagle heavy-hex lattice (library.rs:392-425)."""
    rows = [(0, 14), (0, 15), (0, 15), (0, 15), (0, 15), (0, 15), (1, 14)]
    edges, starts, nxt = [], [], 0
    for r, (first, length) in enumerate(rows):
        starts.append(nxt)
        edges += [(q, q + 1) for q in range(nxt, nxt + length - 1)]
        nxt += length
        if r + 1 < len(rows):
            below_start = nxt + 4
            below_first = rows[r + 1][0]
            for b in range(4):
                col = 4 * b + 2 * (r % 2)
                bridge = nxt + b
                edges.append((starts[r] + col - first, bridge))
                edges.append((bridge, below_start + col - below_first))
            nxt += 4
    return edges


def heavy_hex_edges(n: int) -> list:
    """Eagle couplers restricted to qubits 0..n-1 (33 qubits: 35 edges)."""
    return [(a, b) for a, b in eagle_heavy_hex() if a < n and b < n]


def kicked_ising(n: int, edges, steps: int, theta: float) -> list:
    """RX(theta) on every qubit then RZZ(-pi/2) on every edge, per step (library.rs:430-446)."""
    gates = []
    for _ in range(steps):
        gates += [("rx", (q,), (theta,)) for q in range(n)]
        gates += [("rzz", (a, b), (-HALF_PI,)) for a, b in edges]
    return gates


def hardware_efficient_ansatz(n: int, layers: int, seed: int) -> list:
    """Layers of RY(theta), RZ(phi) on every qubit then the CZ ladder (0,1), (1,2), ...

    Angles are uniform in [0, 2 pi), drawn per gate in program order.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    gates = []
    for _ in range(layers):
        for q in range(n):
            gates.append(("ry", (q,), (float(rng.random()) * 2 * math.pi,)))
            gates.append(("rz", (q,), (float(rng.random()) * 2 * math.pi,)))
        gates += [("cz", (q, q + 1), ()) for q in range(n - 1)]
    return gates


def to_qasm(num_qubits: int, gates: list) -> str:
    """OpenQASM 2 text with qelib1 names and 17-digit angles (circuit.rs:678-726)."""
    names = {"p": "u1", "u": "u3", "cp": "cu1"}
    lines = ['OPENQASM 2.0;', 'include "qelib1.inc";', f"qreg q[{num_qubits}];"]
    for name, qubits, params in gates:
        name = names.get(name, name)
        p = "(" + ",".join(f"{x:.17f}" for x in params) + ")" if params else ""
        lines.append(f"{name}{p} " + ",".join(f"q[{q}]" for q in qubits) + ";")
    return "\n".join(lines) + "\n"
